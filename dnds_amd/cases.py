"""Benchmark and test cases: meshes, settings and initial conditions.

* `isentropic_vortex` -- the reference's vortex initializer
  (specialBuiltinInitializer 2, DNDS_FV_EulerSolver.hpp:536-582): cell means
  of the exact vortex (strength 5, centre (5, 5), free stream (1, 1, 1, 1))
  from the cell quadrature (TRI_4 / QUAD_9), on the periodic [0,10]^2.
* `riemann_2d` -- the four-quadrant Riemann problem of the reference's
  2D Riemann configuration (data/euler_config_2DRiemann.json: far-field
  state plus three box initializers, applied in order at cell barycentres;
  Special_2DRiemannFar boundaries, split at x = y = 0.8).
* `bump_channel`, `flat_plate` -- steady cases for the implicit solver
  (Euler over a bump; laminar Navier-Stokes over a flat plate).
* `mach3_step` -- the Mach-3 step on a mixed tri/quad mesh.
"""
from __future__ import annotations

import numpy as np

from . import _capi as A
from .host import RIEMANN_VR_SETTING, CompactFacedMeshSerialRW, VRFiniteVolume2D, vr_setting

GAMMA = 1.4

TRI4 = np.array([[1. / 3., 1. / 3., -27. / 96.], [0.6, 0.2, 25. / 96.], [0.2, 0.6, 25. / 96.],
                 [0.2, 0.2, 25. / 96.]])
_g = 0.7745966692414833
_gw = np.array([5. / 9., 8. / 9., 5. / 9.])
_gx = np.array([-_g, 0.0, _g])
QUAD9 = np.array([[_gx[i], _gx[j], _gw[i] * _gw[j]] for j in range(3) for i in range(3)])


def cell_means(mesh: CompactFacedMeshSerialRW, volume: np.ndarray, fstate) -> np.ndarray:
    """(1/V) * sum_g fstate(x_g) detJ_g w_g with TRI_4 / QUAD_9 (as the reference's initializer)."""
    m = mesh.arrays()
    cp, c2n, xy = m["cell_ptr"], m["cell2node"], m["node_xy"]
    nv = np.diff(cp)
    out = np.zeros((m["n_cell"], 4))
    for n, rule in ((3, TRI4), (4, QUAD9)):
        idx = np.nonzero(nv == n)[0]
        if idx.size == 0:
            continue
        nodes = c2n[cp[idx][:, None] + np.arange(n)[None, :]]
        X, Y = xy[nodes, 0], xy[nodes, 1]
        acc = np.zeros((idx.size, 4))
        for xi, eta, w in rule:
            if n == 3:
                N = np.array([1 - xi - eta, xi, eta])
                Nx = np.array([-1.0, 1.0, 0.0])
                Ny = np.array([-1.0, 0.0, 1.0])
            else:
                N = np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta),
                              (1 - xi) * (1 + eta)]) * 0.25
                Nx = np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)]) * 0.25
                Ny = np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)]) * 0.25
            px, py = X @ N, Y @ N
            det = (X @ Nx) * (Y @ Ny) - (X @ Ny) * (Y @ Nx)
            acc += fstate(px, py) * (det * w)[:, None]
        out[idx] = acc / volume[idx][:, None]
    return out


def vortex_state(px, py, gamma=GAMMA, chi=5.0, rm=8.0):
    """Exact isentropic vortex at t = 0 (DNDS_FV_EulerSolver.hpp:556-573)."""
    r = np.sqrt((px - 5) ** 2 + (py - 5) ** 2)
    with np.errstate(over="ignore"):
        damp = 1 - 1.0 / np.exp(np.maximum(rm ** 2 - r ** 2, 0.0))
    dT = -(gamma - 1) / (8 * gamma * np.pi ** 2) * chi ** 2 * np.exp(1 - r ** 2) * damp
    dux = chi / 2 / np.pi * np.exp((1 - r ** 2) / 2) * -(py - 5) * damp
    duy = chi / 2 / np.pi * np.exp((1 - r ** 2) / 2) * +(px - 5) * damp
    T = dT + 1
    ux, uy = dux + 1, duy + 1
    rho = T ** (1 / (gamma - 1))
    p = T * rho
    E = 0.5 * (ux ** 2 + uy ** 2) * rho + p / (gamma - 1)
    return np.stack([rho, rho * ux, rho * uy, E], axis=1)


def vortex_exact(px, py, t, gamma=GAMMA, chi=5.0):
    """Vortex convected by the free stream (1, 1) on the periodic [0,10]^2
    (nearest periodic image; the far field is undisturbed to < 1e-10)."""
    cx = np.mod(5 + t, 10.0)
    cy = np.mod(5 + t, 10.0)
    dx = np.mod(px - cx + 5, 10.0) - 5
    dy = np.mod(py - cy + 5, 10.0) - 5
    return vortex_state(dx + 5, dy + 5, gamma, chi, rm=1e6)


# conservative states (rho, rho u, rho v, E) of the 2D Riemann boxes
RIEMANN_FAR = (1.5, 0.0, 0.0, 3.75)
RIEMANN_BOXES = [  # (x0, x1, y0, y1, state)
    (-1e100, 0.8, -1e100, 1e100, (0.532258064516129, 0.641927378778465, 0.0, 1.137096774193549)),
    (-1e100, 1e100, -1e100, 0.8, (0.532258064516129, 0.0, 0.641927378778465, 1.137096774193549)),
    (-1e100, 0.8, -1e100, 0.8, (0.137992831541219, 0.166425616720343, 0.166425616720343, 0.273297491039427)),
]


def riemann_state(bary: np.ndarray) -> np.ndarray:
    """Box initializers in order (DNDS_FV_EulerSolver.hpp:484-497)."""
    u = np.tile(np.array(RIEMANN_FAR), (bary.shape[0], 1))
    for x0, x1, y0, y1, st in RIEMANN_BOXES:
        sel = (bary[:, 0] > x0) & (bary[:, 0] < x1) & (bary[:, 1] > y0) & (bary[:, 1] < y1)
        u[sel] = np.array(st)
    return u


class Case:
    """A mesh, its VR setup, the Euler settings and the initial state."""

    def __init__(self, name, mesh, vfv, riemann, far, u0, steps, sweeps, cfl, limiter, gas=None):
        self.name = name
        self.gas = dict(gas or {})  # EulerEvaluator keyword arguments of the viscous model
        self.mesh = mesh
        self.vfv = vfv
        self.riemann = riemann
        self.far = far
        self.u0 = u0
        self.steps = steps
        self.sweeps = sweeps
        self.cfl = cfl
        self.limiter = limiter

    @property
    def n_cell(self):
        return self.mesh.n_cell


def isentropic_vortex(n=40, seed=1, nthreads=0, build_limiter_data=False) -> Case:
    """Periodic [0,10]^2, n x n quads cut along seeded diagonals (2 n^2 cells);
    Roe, no limiter, CFL 0.5, 5 sweeps per stage, 100 steps."""
    mesh = CompactFacedMeshSerialRW.generate(A.MESH_TRI_SPLIT, n, n, 0.0, 10.0, 0.0, 10.0, periodic=(True, True),
                                             jitter=0.0, seed=seed)
    vfv = VRFiniteVolume2D(mesh, vr_setting(**RIEMANN_VR_SETTING, build_limiter_data=int(build_limiter_data)))
    vfv.Initialization(nthreads)
    vol = vfv.arrays(["volume"])["volume"]
    u0 = cell_means(mesh, vol, vortex_state)
    return Case("isentropic_vortex", mesh, vfv, "Roe", (1.0, 1.0, 1.0, 2.5 + 1.0), u0, 100, 5, 0.5, False)


def riemann_2d(n=1024, seed=2, jitter=0.2, nthreads=0, build_limiter_data=True, wbap_scale=None) -> Case:
    """[0,1]^2, n x n quads cut along seeded diagonals (2 n^2 cells), interior
    nodes jittered by U(-0.2h, 0.2h); HLLC, SSP-RK3, 200 steps.  wbap_scale:
    WBAP_SmoothIndicatorScale (default that of the reference's 2D Riemann
    configuration, 0.1)."""
    bc = A.BC_SPECIAL_2DRIEMANNFAR
    mesh = CompactFacedMeshSerialRW.generate(A.MESH_TRI_SPLIT, n, n, 0.0, 1.0, 0.0, 1.0, jitter=jitter, seed=seed,
                                             bc=(bc, bc, bc, bc))
    st = dict(RIEMANN_VR_SETTING)
    if wbap_scale is not None:
        st["wbap_smooth_indicator_scale"] = wbap_scale
    vfv = VRFiniteVolume2D(mesh, vr_setting(**st, build_limiter_data=int(build_limiter_data)))
    vfv.Initialization(nthreads)
    bary = vfv.arrays(["cell_bary"])["cell_bary"].reshape(-1, 2)
    u0 = riemann_state(bary)
    return Case("riemann_2d", mesh, vfv, "HLLC", RIEMANN_FAR, u0, 200, 1, 0.5, True)


def bump_far(mach=0.5, gamma=GAMMA):
    """Free stream rho = 1, p = 1/gamma (sound speed 1), u = mach, v = 0."""
    p = 1.0 / gamma
    return (1.0, mach, 0.0, p / (gamma - 1) + 0.5 * mach * mach)


def bump_channel(nx=48, ny=16, kind=A.MESH_TRI_SPLIT, height=0.1, mach=0.5, seed=4, nthreads=0) -> Case:
    """Steady subsonic channel flow over a sine bump: [0,3] x [0,1] with the
    lower wall raised by height * sin^2(pi (x - 1)) on 1 <= x <= 2 (nodes
    moved as y' = b(x) + y (1 - b(x))), Wall_Euler top and bottom, Farfield
    inlet and outlet at Mach `mach`.  Roe flux, no limiter; the initial state
    is the free stream.  VR weights as the reference's implicit 2D Euler
    configuration (data/euler2D_config.json: farWeight 1, wallWeight 0.5),
    1 Jacobi sweep per iteration, CFL 10; the implicit LU-SGS solver drives
    it to steady state."""
    bcs = (A.BC_FARFIELD, A.BC_FARFIELD, A.BC_WALL_EULER, A.BC_WALL_EULER)
    base = CompactFacedMeshSerialRW.generate(kind, nx, ny, 0.0, 3.0, 0.0, 1.0, jitter=0.0, seed=seed, bc=bcs)
    m = base.arrays()
    xy = m["node_xy"].copy()
    x = xy[:, 0]
    b = np.where((x >= 1.0) & (x <= 2.0), height * np.sin(np.pi * (x - 1.0)) ** 2, 0.0)
    xy[:, 1] = b + xy[:, 1] * (1.0 - b)
    bnd = m["face2cell"][:, 1] == A.NO_CELL
    mesh = CompactFacedMeshSerialRW.from_cells(xy, m["cell_ptr"], m["cell2node"], m["face2node"][bnd],
                                               m["face_bc"][bnd])
    vfv = VRFiniteVolume2D(mesh, vr_setting(**dict(RIEMANN_VR_SETTING, far_weight=1.0, wall_weight=0.5),
                                            build_limiter_data=0))
    vfv.Initialization(nthreads)
    far = bump_far(mach)
    u0 = np.tile(np.array(far), (mesh.n_cell, 1))
    return Case("bump_channel", mesh, vfv, "Roe", far, u0, 0, 1, 10.0, False)


# the reference's boundary-layer vfvSetting (data/eulerSA_config_BL.json:61-77):
# isotropic lengths, no tangential weight, no geometric weighting, far-field
# weight 1, wall weight 0.5
BL_VR_SETTING = dict(anisotropic_lengths=0, tang_weight=0.0, tang_weight_direction=0, wall_weight=0.5,
                     far_weight=1.0, weight_len_scale=1.0, weight_scheme_geom=0, weight_scheme_dir=0,
                     jacobi_relax=1.0)


def flat_plate(nx=1536, ny=320, mach=0.2, reynolds=1e5, y_plus=1.0, height=1.0, upstream=0.0, rec_order=3,
               cfl=20.0, nthreads=0) -> Case:
    """Laminar flat plate (the NS_2D model: the 2D solver has no turbulence
    model): [0,1] x [0,height], nx x ny quads, the plate (Wall_NoSlip) along
    y = 0 from the leading edge at x = 0, Farfield on the other three sides.
    Free stream rho = 1, sound speed 1, u = mach; Re = rho u L / mu per unit
    length, mu from Sutherland's law with TRef the free-stream temperature and
    CSutherland 110.4 K / 273.15 K of it (air), Prandtl 0.72, Rgas 1.  The y
    lines are spaced geometrically (one ratio for the whole height) so that
    the first cell's height puts the wall at y+ = y_plus for the turbulent
    skin friction cf = 0.0576 Re^-0.2 at x = 1.  Roe flux, no limiter, the
    reference's boundary-layer VR settings (BL_VR_SETTING), P_ORDER
    rec_order (3 or 1); initial state the free stream.  1 Jacobi sweep per
    iteration; the implicit LU-SGS solver with local time steps at CFL `cfl`
    drives it to steady state.  At Re 5e6 the laminar case diverges (the
    wall cells are dominated by the inviscid terms); Re 1e5 converges."""
    far = bump_far(mach)
    mu = mach / reynolds
    u_tau = mach * np.sqrt(0.5 * 0.0576 * reynolds ** -0.2)
    h0 = y_plus * mu / u_tau
    # ratio r with h0 (r^ny - 1) / (r - 1) = height, by bisection
    lo, hi = 1.0 + 1e-12, 3.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if h0 * (r ** ny - 1.0) / (r - 1.0) > height:
            hi = r
        else:
            lo = r
    bcs = (A.BC_FARFIELD, A.BC_FARFIELD, A.BC_WALL_NOSLIP, A.BC_FARFIELD)
    mesh = CompactFacedMeshSerialRW.generate(A.MESH_QUAD, nx, ny, -upstream, 1.0, 0.0, height, y_ratio=r, bc=bcs)
    if upstream > 0:
        # symmetry (Wall_Euler) ahead of the leading edge
        m = mesh.arrays()
        bnd = m["face2cell"][:, 1] == A.NO_CELL
        f2n, bc = m["face2node"][bnd], m["face_bc"][bnd].copy()
        xm = 0.5 * (m["node_xy"][f2n[:, 0], 0] + m["node_xy"][f2n[:, 1], 0])
        bc[(bc == A.BC_WALL_NOSLIP) & (xm < 0.0)] = A.BC_WALL_EULER
        mesh = CompactFacedMeshSerialRW.from_cells(m["node_xy"], m["cell_ptr"], m["cell2node"], f2n, bc)
    vfv = VRFiniteVolume2D(mesh, vr_setting(**BL_VR_SETTING, rec_order=rec_order, build_limiter_data=0))
    vfv.Initialization(nthreads)
    t_inf = (far[3] - 0.5 * mach * mach) * (GAMMA - 1.0)  # p / (Rgas rho) with rho = Rgas = 1
    gas = dict(muGas=mu, prGas=0.72, Rgas=1.0, TRef=t_inf, CSutherland=110.4 / 273.15 * t_inf)
    u0 = np.tile(np.array(far), (mesh.n_cell, 1))
    return Case("flat_plate", mesh, vfv, "Roe", far, u0, 0, 1, cfl, False, gas)


MACH3_LEFT = (3.857, 2.629, 0.0, 10.33)   # (rho, u, v, p) behind the Mach-3 shock
MACH3_RIGHT = (1.0, 0.0, 0.0, 1.0)


def _cons(rho, u, v, p, gamma=GAMMA):
    return np.stack([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)], axis=-1)


def mach3_step(n=2048, seed=5, jitter=0.15, x_step=0.5, noise=1e-3, nthreads=0) -> Case:
    """The combined large case: [0,4] x [0,1], n x n quads, each cut into two
    triangles or kept (seeded, half each: a mixed mesh of about 1.5 n^2
    cells), interior nodes jittered by U(-0.15h, 0.15h).  A Mach-3 shock
    stands at x = 0.5: left (rho, u, v, p) = (3.857, 2.629, 0, 10.33), right
    (1, 0, 0, 1), with seeded density noise of relative amplitude 1e-3 (u, p
    kept).  Far-field inflow on the left (the left state, supersonic), walls
    (Wall_Euler) elsewhere.  HLLC, WBAP, SSP-RK3, 100 steps."""
    bcs = (A.BC_FARFIELD, A.BC_WALL_EULER, A.BC_WALL_EULER, A.BC_WALL_EULER)
    mesh = CompactFacedMeshSerialRW.generate(A.MESH_MIXED, n, n, 0.0, 4.0, 0.0, 1.0, jitter=jitter, seed=seed, bc=bcs)
    vfv = VRFiniteVolume2D(mesh, vr_setting(**dict(RIEMANN_VR_SETTING, far_weight=1.0), build_limiter_data=1))
    vfv.Initialization(nthreads)
    bary = vfv.arrays(["cell_bary"])["cell_bary"].reshape(-1, 2)
    left = bary[:, 0] < x_step
    prim = np.where(left[:, None], np.array(MACH3_LEFT), np.array(MACH3_RIGHT))
    rng = np.random.default_rng(seed)
    rho = prim[:, 0] * (1 + noise * rng.uniform(-1, 1, prim.shape[0]))
    u0 = _cons(rho, prim[:, 1], prim[:, 2], prim[:, 3])
    far = tuple(_cons(*(np.array(v) for v in MACH3_LEFT)).tolist())
    return Case("mach3_step", mesh, vfv, "HLLC", far, u0, 100, 5, 0.5, True)
