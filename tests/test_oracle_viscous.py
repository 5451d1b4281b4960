"""The oracle's laminar viscous terms (NS_2D with muGas > 0) against what the
mathematics fixes.  CPU only.

When every conservative variable is a global cubic and the cell means and
reconstruction coefficients are its exact ones, both sides of every interior
face reconstruct the same values and gradients, the jump term vanishes, and
the viscous part of EvaluateRHS is the three-point Gauss sum of the
Navier-Stokes viscous flux of that exact state over each cell's faces:

  d rhs_i = (1 / V_i) sum_f sum_g w_g |n_g| F_v(x_g) . n_out

with F_v written here in primitive variables (velocity gradient by the
quotient rule, T = p / (rho R), Sutherland's law, k = Cp mu / Pr), not in
the reference's conservative-gradient form (Gas::ViscousFlux_IdealGas,
DNDS_Gas.hpp:795-826).  At no-slip walls the ghost state (rho, -m, E) makes
the face state's velocity zero and the adiabatic condition removes the
normal conduction (DNDS_Gas.hpp:814-815), so the wall flux is the stress of
the exact momentum gradient divided by rho, with no energy flux; at slip
walls the ghost is the inner state with its normal momentum removed
(DNDS_FV_EulerEvaluator.hpp:1247-1251), so the face state keeps half of it,
and the full conduction.  At both, the jump term is
divided by 3e200 and vanishes.  The
inviscid parts are identical with and without viscosity (HLLC does not use
the Harten-Yee lambda), so d rhs = rhs(mu) - rhs(0).
"""
from math import factorial

import numpy as np
import pytest

import oracle
from conftest import mesh_and_setup, rel_err

MONO = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
# [var][monomial] coefficients of (x - 0.5)^p (y - 0.5)^q
P = np.array([[1.0, 0.10, -0.05, 0.08, 0.03, -0.06, 0.05, -0.02, 0.04, 0.03],
              [0.3, 0.20, 0.15, -0.10, 0.12, 0.05, 0.07, -0.04, 0.02, 0.06],
              [-0.2, 0.10, 0.25, 0.06, -0.08, 0.10, -0.03, 0.05, 0.04, -0.05],
              [2.5, 0.30, -0.20, 0.15, 0.10, -0.12, 0.05, 0.08, -0.06, 0.04]])
# a Mach-2 state flowing in through x = 0 (u near 2, a near 1)
P_SUP = np.array([[1.0, 0.05, -0.03, 0.04, 0.02, -0.03, 0.02, -0.01, 0.02, 0.01],
                  [2.0, 0.10, 0.08, -0.05, 0.06, 0.03, 0.04, -0.02, 0.01, 0.03],
                  [0.1, 0.05, 0.12, 0.03, -0.04, 0.05, -0.02, 0.03, 0.02, -0.03],
                  [3.8, 0.15, -0.10, 0.08, 0.05, -0.06, 0.03, 0.04, -0.03, 0.02]])
FAR_SUP = (1.1, 1.1 * 2.2, 1.1 * 0.05, 0.75 / 0.4 + 0.5 * 1.1 * (2.2 ** 2 + 0.05 ** 2))
GAMMA = 1.4
COEF = {"sub": P, "sup": P_SUP}


def poly(x, y, a=0, b=0, coef=None):
    """d^(a+b) P / dx^a dy^b at (x, y): [..., 4]."""
    coef = P if coef is None else coef
    X, Y = np.asarray(x) - 0.5, np.asarray(y) - 0.5
    out = np.zeros(np.shape(X) + (4,))
    for k, (p, q) in enumerate(MONO):
        if p >= a and q >= b:
            t = (factorial(p) / factorial(p - a) * factorial(q) / factorial(q - b) *
                 X ** (p - a) * Y ** (q - b))
            out += t[..., None] * coef[:, k]
    return out


def exact_state(m, s, coef=None):
    """Exact cell means (degree-3 triangle rule over a fan) and coefficients."""
    xy = m["node_xy"].reshape(-1, 2)
    cp, c2n = m["cell_ptr"], m["cell2node"]
    n = len(cp) - 1
    u = np.zeros((n, 4))
    for i in range(n):
        pts = xy[c2n[cp[i]:cp[i + 1]]]
        tot, area = np.zeros(4), 0.0
        for k in range(1, len(pts) - 1):
            A, B, C = pts[0], pts[k], pts[k + 1]
            ar = 0.5 * ((B[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (B[1] - A[1]))
            q = (np.array([A, B, C]), np.array([(A + B) / 2, (B + C) / 2, (C + A) / 2]), (A + B + C) / 3)
            val = (poly(q[0][:, 0], q[0][:, 1], coef=coef).sum(0) / 20 +
                   poly(q[1][:, 0], q[1][:, 1], coef=coef).sum(0) * 2 / 15 + poly(q[2][0], q[2][1], coef=coef) * 9 / 20)
            tot += val * ar
            area += ar
        u[i] = tot / area
    sc = s["cell_scale"].reshape(-1, 2)
    c = s["cell_bary"].reshape(-1, 2)
    urec = np.zeros((n, 36))
    for r, (a, b) in enumerate(MONO[1:]):
        d = poly(c[:, 0], c[:, 1], a, b, coef=coef)
        for v in range(4):
            urec[:, r + 9 * v] = d[:, v] * sc[:, 0] ** a * sc[:, 1] ** b / (factorial(a) * factorial(b))
    return u, urec


def viscous_flux(U, G, mu_gas, pr, R, tref, cs, wall=None, n=None):
    """[..., d, v] flux of the laminar Navier-Stokes terms in direction d;
    wall 4 (no-slip): the face state has no momentum and no conduction
    through the wall; wall 3 (slip): the face state, the mean of the inner
    state and its projection, keeps half the normal momentum."""
    rho, mom, E = U[..., 0], U[..., 1:3], U[..., 3]
    if wall == 4:
        mom = np.zeros_like(mom)
    elif wall == 3:
        mom = mom - 0.5 * np.sum(mom * n, -1)[..., None] * n
    vel = mom / rho[..., None]
    gv = (G[..., :, 1:3] - vel[..., None, :] * G[..., :, 0:1]) / rho[..., None, None]  # [.., i, j] du_j/dx_i
    p = (GAMMA - 1) * (E - 0.5 * rho * np.sum(vel * vel, -1))
    gq = (2 * np.einsum("...j,...ij->...i", mom, G[..., :, 1:3]) / rho[..., None] -
          np.sum(mom * mom, -1)[..., None] * G[..., :, 0] / rho[..., None] ** 2)
    gp = (GAMMA - 1) * (G[..., :, 3] - 0.5 * gq)
    T = p / (rho * R)
    gT = (rho[..., None] * gp - p[..., None] * G[..., :, 0]) / (rho[..., None] ** 2 * R)
    mu = mu_gas * (T / tref) ** 1.5 * (tref + cs) / (T + cs)
    k = R * GAMMA / (GAMMA - 1) * mu / pr
    div = gv[..., 0, 0] + gv[..., 1, 1]
    tau = mu[..., None, None] * (gv + np.swapaxes(gv, -1, -2)) - (2 / 3) * (mu * div)[..., None, None] * np.eye(2)
    F = np.zeros(U.shape[:-1] + (2, 4))
    F[..., 1:3] = tau
    F[..., 3] = np.einsum("...ij,...j->...i", tau, vel) + (0 if wall == 4 else 1) * k[..., None] * gT
    return F


def expected_viscous_rhs(m, s, consts, coef=None, far=None):
    """Inner and wall faces as the module docstring says; a farfield face
    where the flow enters supersonically has the far state as its ghost, so
    its face state is the mean of the inner value and the far state and its
    gradient the inner one plus n (U_far - U) / distGRP, distGRP = 2 V / A
    (DNDS_FV_EulerEvaluator.cpp:356-395)."""
    xy = m["node_xy"].reshape(-1, 2)
    f2n, f2c, bc = m["face2node"].reshape(-1, 2), m["face2cell"].reshape(-1, 2), m["face_bc"]
    vol = s["volume"]
    n = len(vol)
    out = np.zeros((n, 4))
    xi, w = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)]), np.array([5 / 9, 8 / 9, 5 / 9])
    for f in range(len(bc)):
        a, b = xy[f2n[f, 0]], xy[f2n[f, 1]]
        d = 0.5 * (b - a)
        nrm = np.array([d[1], -d[0]])
        nl = np.linalg.norm(nrm)
        pts = 0.5 * (a + b)[None, :] + xi[:, None] * d[None, :]
        U = poly(pts[:, 0], pts[:, 1], coef=coef)
        G = np.stack([poly(pts[:, 0], pts[:, 1], 1, 0, coef=coef), poly(pts[:, 0], pts[:, 1], 0, 1, coef=coef)], 1)
        if bc[f] == 2:
            dist = vol[f2c[f, 0]] / (2 * nl) * 2  # minVol / faceArea * 2, faceArea = 2 |n|
            jump = np.asarray(far)[None, :] - U
            G = G + (nrm / nl)[None, :, None] * jump[:, None, :] / dist
            U = 0.5 * (U + np.asarray(far)[None, :])
        F = viscous_flux(U, G, *consts, wall=int(bc[f]) if bc[f] in (3, 4) else None, n=nrm / nl)
        flux = np.einsum("g,gdv,d->v", w * nl, F, nrm / nl)
        L, R = f2c[f]
        out[L] += flux / vol[L]
        if R >= 0:
            out[R] -= flux / vol[R]
    return out


CONSTS = {"unit": (0.2, 0.72, 1.0, 1.0, 0.5),            # mu, Pr, R, TRef, CSutherland
          "air": (0.2, 0.72, 1.0 / 300.0, 273.15, 110.4)}   # T near 300 with the reference's TRef, CSutherland


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("consts", ["unit", "air"])
@pytest.mark.parametrize("walls,state", [((4, 4, 4, 4), "sub"), ((3, 4, 3, 3), "sub"), ((2, 4, 3, 4), "sup")])
def test_viscous_rhs_of_an_exact_cubic_state(kind, consts, walls, state):
    mesh, vfv = mesh_and_setup(kind=kind, nx=7, ny=6, periodic=False, jitter=0.2, seed=4, bc=walls)
    m, s = mesh.arrays(), vfv.arrays()
    st = vfv.setting
    mu, pr, R, tref, cs = CONSTS[consts]
    far = FAR_SUP if state == "sup" else (1.0, 0.0, 0.0, 2.5)
    kw = dict(riemann_solver=2, tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
              pr_gas=pr, r_gas=R, t_ref=tref, c_sutherland=cs, far_field=far)
    inv, vis = oracle.Oracle(m, s, mu_gas=0.0, **kw), oracle.Oracle(m, s, mu_gas=mu, **kw)
    u, urec = exact_state(m, s, COEF[state])
    _, _, _, dlf = inv.evaluate_dt(u, 0.5)
    d = vis.evaluate_rhs(u, urec, dlf) - inv.evaluate_rhs(u, urec, dlf)
    ref = expected_viscous_rhs(m, s, CONSTS[consts], COEF[state], far)
    assert np.abs(ref).max() > 1e-2
    assert rel_err(d, ref) < 1e-10


def test_viscous_time_step_on_a_uniform_mesh_at_rest():
    """Quads of side h, a fluid at rest: lambdaFace = a + lamVis A (2 / V)
    with lamVis = mu(T) / rho max(4/3, gamma / Pr) (DNDS_FV_EulerEvaluator.cpp:
    79-119), so dt = CFL h / (4 a + 8 lamVis / h) in every cell."""
    nx = 8
    mesh, vfv = mesh_and_setup(kind=1, nx=nx, ny=nx, periodic=True, jitter=0.0, seed=1)
    m, s = mesh.arrays(), vfv.arrays()
    mu, pr, R, tref, cs = 0.3, 0.7, 2.0, 1.5, 0.4
    orc = oracle.Oracle(m, s, mu_gas=mu, pr_gas=pr, r_gas=R, t_ref=tref, c_sutherland=cs)
    rho, p = 1.3, 0.9
    u = np.tile([rho, 0.0, 0.0, p / (GAMMA - 1)], (mesh.n_cell, 1))
    dt, dtmin, lf, _ = orc.evaluate_dt(u, 0.5, use_local_dt=True)
    h = 1.0 / nx
    T = p / (rho * R)
    muf = mu * (T / tref) ** 1.5 * (tref + cs) / (T + cs)
    lam_vis = muf / rho * max(4 / 3, GAMMA / pr)
    a = np.sqrt(GAMMA * p / rho)
    assert rel_err(dt, np.full(mesh.n_cell, 0.5 * h / (4 * a + 8 * lam_vis / h))) < 1e-13
    assert rel_err(lf, np.full(mesh.n_face, a + lam_vis * 2 / h)) < 1e-13


@pytest.mark.parametrize("bad", [dict(muGas=-1.0), dict(muGas=0.1, prGas=0.0), dict(muGas=0.1, Rgas=0.0),
                                 dict(muGas=0.1, TRef=-1.0), dict(muGas=0.1, CSutherland=-0.1)])
def test_viscous_settings_are_checked_before_the_device_is_touched(bad):
    from dnds_amd import _capi as A
    from dnds_amd.host import DNDSError, EulerEvaluator
    mesh, vfv = mesh_and_setup(kind=0, nx=4, ny=4)
    with pytest.raises(DNDSError) as e:
        EulerEvaluator(mesh, vfv, **bad)
    assert e.value.code == A.ERR_INVALID_ARGUMENT
