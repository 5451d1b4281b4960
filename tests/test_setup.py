"""Host VR setup: against an independent numpy restatement, and against
what the mathematics fixes (polynomial exactness of the reconstruction,
symmetry of A_ii, moments, convergence order).  CPU only."""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err

SETTINGS = {
    # the reference's 2D Riemann vfvSetting (global coordinates)
    "riemann": dict(anisotropic_lengths=1, tang_weight=1.0, tang_weight_direction=0, wall_weight=0.5,
                    far_weight=0.0, weight_len_scale=0.5, weight_scheme_geom=2, weight_scheme_dir=0),
    # defaults of VRFiniteVolume2D::Setting with isotropic lengths
    "default": dict(anisotropic_lengths=0, tang_weight=1.0, tang_weight_direction=1, wall_weight=1.0,
                    far_weight=0.0, weight_len_scale=1.0, weight_scheme_geom=0, weight_scheme_dir=0),
    # SDHQM geometric and optHQM directional weights, far weight 1
    "hqm": dict(anisotropic_lengths=1, tang_weight=0.5, tang_weight_direction=0, wall_weight=0.5,
                far_weight=1.0, weight_len_scale=1.0, weight_scheme_geom=3, weight_scheme_dir=1),
}


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("sname", list(SETTINGS))
@pytest.mark.parametrize("bcset", ["periodic", "mixed"])
def test_setup_matches_numpy_restatement(kind, sname, bcset):
    from oracle import setup_np
    s = SETTINGS[sname]
    periodic = bcset == "periodic"
    bc = None if periodic else (2, 3, 4, 1)
    mesh, vfv = mesh_and_setup(kind=kind, nx=6, ny=5, periodic=periodic, jitter=0.0 if periodic else 0.2, seed=7,
                               bc=bc, **s)
    got = vfv.arrays()
    ref = setup_np.build(mesh.arrays(), s)
    for k in ("volume", "face_area", "cell_bary", "base_moment", "face_center", "face_norm_center", "face_weight",
              "face_d0", "face_dibj_g0", "face_gauss_xy", "cell_scale"):
        assert rel_err(got[k], ref[k]) <= 1e-13, k
    # A_ii^-1 and its products carry the conditioning of A_ii
    for k in ("ainv", "ainv_b", "ainv_bvec", "matrix_secondary"):
        assert rel_err(got[k], ref[k]) <= 1e-12, (k, rel_err(got[k], ref[k]))


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("sname", list(SETTINGS))
def test_exact_cubic_is_a_fixed_point_of_the_sweep(kind, sname):
    """Cell means of a global cubic and its exact coefficients: every face
    jump vanishes, so one Jacobi sweep returns the coefficients."""
    import oracle
    s = SETTINGS[sname]
    mesh, vfv = mesh_and_setup(kind=kind, nx=7, ny=6, jitter=0.2, seed=3, bc=(2, 2, 2, 2), **dict(s, far_weight=0.0))
    m, st = mesh.arrays(), vfv.arrays()
    coef = {(0, 0): 1.0, (1, 0): 2.0, (0, 1): -1.0, (2, 0): 0.5, (1, 1): 0.3, (0, 2): -0.7, (3, 0): 0.2,
            (2, 1): -0.1, (1, 2): 0.4, (0, 3): 0.3}
    from math import comb
    from dnds_amd import cases

    def P(x, y):
        return sum(c * x ** a * y ** b for (a, b), c in coef.items())

    vol = st["volume"]
    means = cases.cell_means(mesh, vol, lambda x, y: np.stack([P(x, y) * (k + 1) for k in range(4)], axis=1))
    bary = st["cell_bary"].reshape(-1, 2)
    DIFF = [(1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
    urec = np.zeros((m["n_cell"], 36))
    for i in range(m["n_cell"]):
        X = m["node_xy"][m["cell2node"][m["cell_ptr"][i]:m["cell_ptr"][i + 1]]]
        sc = (X.max(0) - X.min(0)) / 2
        if not s["anisotropic_lengths"]:
            sc = np.full(2, np.linalg.norm(sc))
        for k, (dx, dy) in enumerate(DIFF):
            t = sum(c * comb(a, dx) * comb(b, dy) * bary[i, 0] ** (a - dx) * bary[i, 1] ** (b - dy)
                    for (a, b), c in coef.items() if a >= dx and b >= dy)
            for v in range(4):
                urec[i, k + 9 * v] = t * sc[0] ** dx * sc[1] ** dy * (v + 1)
    orc = oracle.Oracle(m, st)
    new = orc.jacobi_step(means, urec)
    assert rel_err(new, urec) <= 1e-12


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_ainv_is_symmetric_and_moments_consistent(kind):
    mesh, vfv = mesh_and_setup(kind=kind, nx=7, ny=6, jitter=0.2, seed=4, bc=(2, 3, 4, 3))
    st = vfv.arrays()
    Ai = st["ainv"].reshape(-1, 9, 9)
    assert rel_err(Ai, np.transpose(Ai, (0, 2, 1))) <= 1e-10
    # A_ii is positive definite: its inverse too
    assert np.all(np.linalg.eigvalsh(0.5 * (Ai + np.transpose(Ai, (0, 2, 1)))) > 0)
    mom = st["base_moment"].reshape(-1, 10)
    assert np.all(mom[:, 0] == 1.0)
    # linear moments vanish about the barycentre
    assert np.abs(mom[:, 1:3]).max() < 1e-13


def test_periodic_reconstruction_converges_at_fourth_order():
    """Static reconstruction of the reference's test function
    (1 - cos 2 pi x)(1 - cos 2 pi y) (test/staticReconstructionTest.cpp:38-57)
    on periodic split-quad meshes: Jacobi to convergence, L1 error of the
    point values at the cell quadrature points; p = 3 gives order ~4."""
    import oracle
    from dnds_amd import cases

    def f(x, y):
        return (1 - np.cos(2 * np.pi * x)) * (1 - np.cos(2 * np.pi * y))

    errs = []
    for n in (8, 16, 32):
        mesh, vfv = mesh_and_setup(kind=0, nx=n, ny=n, periodic=True, jitter=0.0, seed=5)
        m, st = mesh.arrays(), vfv.arrays()
        # high-order cell means: 4 x 4 sub-triangles of TRI_4 points
        means = np.zeros((m["n_cell"], 4))
        sub = []
        k = 4
        for a in range(k):
            for b in range(k - a):
                sub.append(((a, b), (a + 1, b), (a, b + 1)))
                if a + b < k - 1:
                    sub.append(((a + 1, b), (a + 1, b + 1), (a, b + 1)))
        xy, cp, c2n = m["node_xy"], m["cell_ptr"], m["cell2node"]
        X = xy[c2n.reshape(-1, 3)]
        for tri in sub:
            P = [np.array(p, float) / k for p in tri]
            for (s1, s2, w) in cases.TRI4:
                lam = P[0] + s1 * (P[1] - P[0]) + s2 * (P[2] - P[0])
                pt = X[:, 0] + lam[0] * (X[:, 1] - X[:, 0]) + lam[1] * (X[:, 2] - X[:, 0])
                means[:, 0] += w * 2 / k ** 2 * f(pt[:, 0], pt[:, 1])
        orc = oracle.Oracle(m, st)
        urec = np.zeros((m["n_cell"], 36))
        for _ in range(300):
            urec = orc.jacobi_step(means, urec)
        bary = st["cell_bary"].reshape(-1, 2)
        mom = st["base_moment"].reshape(-1, 10)
        err, vol = 0.0, st["volume"]
        for (s1, s2, w) in cases.TRI4:
            pt = X[:, 0] + s1 * (X[:, 1] - X[:, 0]) + s2 * (X[:, 2] - X[:, 0])
            sc = (X.max(1) - X.min(1)) / 2
            dx, dy = (pt[:, 0] - bary[:, 0]) / sc[:, 0], (pt[:, 1] - bary[:, 1]) / sc[:, 1]
            phi = np.stack([dx, dy, dx * dx, dx * dy, dy * dy, dx ** 3, dx * dx * dy, dx * dy * dy, dy ** 3], 1)
            val = means[:, 0] + np.einsum("ik,ik->i", phi - mom[:, 1:], urec[:, :9])
            err += np.sum(np.abs(val - f(pt[:, 0], pt[:, 1])) * w * 2 * vol)
        errs.append(err)
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 3.5), (errs, orders)
