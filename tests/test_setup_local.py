"""The local-inertia-coordinate basis (useLocalCoord, DNDS_FV_VR.hpp:358-1018;
the frame from the cell's second-moment tensor, DNDS_FV_VR.cpp:118-139):
the host setup against the numpy restatement (whose chain rule is a
polynomial expansion of its own, not the product's tensor contraction), and
against what the mathematics fixes -- the local monomials span the same
cubics as the global ones, so a global cubic is a fixed point of the sweep
and the converged reconstruction's point values are those of the global
basis, whatever the frame's orientation.  CPU only."""
import numpy as np
import pytest

from conftest import mesh_and_setup, rel_err

SETTINGS = {
    # the reference's isentropic-vortex vfvSetting (data/euler_config_IV.json:60-80)
    "iv": dict(anisotropic_lengths=1, tang_weight=0.0, tang_weight_direction=0, wall_weight=0.5, far_weight=0.0,
               weight_len_scale=1.0, weight_scheme_geom=3, weight_scheme_dir=1),
    # isotropic lengths, barycentric tangential direction
    "iso": dict(anisotropic_lengths=0, tang_weight=1.0, tang_weight_direction=1, wall_weight=1.0, far_weight=1.0,
                weight_len_scale=1.0, weight_scheme_geom=0, weight_scheme_dir=0),
}


def local(kind, s, **kw):
    return mesh_and_setup(kind=kind, limiter=False, use_local_coord=1, **dict(s, **kw))


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("sname", list(SETTINGS))
@pytest.mark.parametrize("bcset", ["periodic", "mixed"])
def test_local_setup_matches_numpy_restatement(kind, sname, bcset):
    from oracle import setup_np
    s = SETTINGS[sname]
    periodic = bcset == "periodic"
    mesh, vfv = local(kind, s, nx=6, ny=5, periodic=periodic, jitter=0.0 if periodic else 0.2, seed=7,
                      bc=None if periodic else (2, 3, 4, 1))
    got = vfv.arrays()
    ref = setup_np.build(mesh.arrays(), dict(s, use_local_coord=1))
    for k in ("base_moment", "face_d0", "face_gauss_xy", "cell_frame"):
        assert rel_err(got[k], ref[k]) <= 1e-12, (k, rel_err(got[k], ref[k]))
    for k in ("ainv", "ainv_b", "ainv_bvec"):
        assert rel_err(got[k], ref[k]) <= 1e-11, (k, rel_err(got[k], ref[k]))


def test_frame_is_the_scaled_principal_axes():
    """A = invPJacobi: the columns of A^-1 are orthogonal, ordered by the
    second moments along them, each as long as the cell's half extent along
    it (anisotropic lengths), and a stretched quad's first axis is its long
    side."""
    mesh, vfv = local(1, SETTINGS["iv"], nx=5, ny=4, jitter=0.25, seed=2, bc=(2, 2, 2, 2))
    m, st = mesh.arrays(), vfv.arrays()
    A = st["cell_frame"].reshape(-1, 2, 2)
    bary = st["cell_bary"].reshape(-1, 2)
    for i in range(m["n_cell"]):
        P = np.linalg.inv(A[i])
        u0, u1 = P[:, 0] / np.linalg.norm(P[:, 0]), P[:, 1] / np.linalg.norm(P[:, 1])
        assert abs(u0 @ u1) < 1e-12
        X = m["node_xy"][m["cell2node"][m["cell_ptr"][i]:m["cell_ptr"][i + 1]]] - bary[i]
        for k, u in enumerate((u0, u1)):
            ext = X @ u
            assert abs(np.linalg.norm(P[:, k]) - 0.5 * (ext.max() - ext.min())) < 1e-12
    # [0,4] x [0,1] in 4 x 4 quads: cells 4 times longer than high
    mesh, vfv = local(1, SETTINGS["iv"], nx=4, ny=4, x1=4.0, jitter=0.0, seed=2, bc=(2, 2, 2, 2))
    P = np.linalg.inv(vfv.arrays()["cell_frame"].reshape(-1, 2, 2))
    assert np.allclose(np.abs(P[:, :, 0]), [[0.5, 0.0]] * 1, atol=1e-12)  # half the long side along x
    assert np.allclose(np.abs(P[:, :, 1]), [[0.0, 0.125]] * 1, atol=1e-12)


def test_local_setup_refuses_limiter_data():
    from dnds_amd.host import DNDSError
    with pytest.raises(DNDSError):
        mesh_and_setup(kind=0, nx=4, ny=4, limiter=True, use_local_coord=1, **SETTINGS["iv"])


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_exact_cubic_is_a_fixed_point_in_local_coordinates(kind):
    """Cell means of a global cubic; the exact coefficients in the local
    basis come from the cubic's values at the cell's face points (a least
    squares fit on 3 points of each face, which the cubic satisfies
    exactly); one sweep returns them."""
    import oracle
    from dnds_amd import cases
    s = SETTINGS["iv"]
    mesh, vfv = local(kind, s, nx=7, ny=6, jitter=0.2, seed=3, bc=(2, 2, 2, 2))
    m, st = mesh.arrays(), vfv.arrays()
    coef = {(0, 0): 1.0, (1, 0): 2.0, (0, 1): -1.0, (2, 0): 0.5, (1, 1): 0.3, (0, 2): -0.7, (3, 0): 0.2,
            (2, 1): -0.1, (1, 2): 0.4, (0, 3): 0.3}

    def P(x, y):
        return sum(c * x ** a * y ** b for (a, b), c in coef.items())

    vol = st["volume"]
    means = cases.cell_means(mesh, vol, lambda x, y: np.stack([P(x, y) * (k + 1) for k in range(4)], axis=1))
    A = st["cell_frame"].reshape(-1, 2, 2)
    bary = st["cell_bary"].reshape(-1, 2)
    mom = st["base_moment"].reshape(-1, 10)
    DIFF = [(1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
    rng = np.random.default_rng(0)
    urec = np.zeros((m["n_cell"], 36))
    for i in range(m["n_cell"]):
        # points in the cell's neighbourhood; the cubic minus its mean is
        # exactly a combination of the 9 zero-mean local monomials
        pts = bary[i] + rng.uniform(-0.05, 0.05, (30, 2))
        xi = (A[i] @ (pts - bary[i]).T).T
        phi = np.stack([xi[:, 0] ** a * xi[:, 1] ** b for a, b in DIFF], 1) - mom[i, 1:]
        rhs = P(pts[:, 0], pts[:, 1]) - means[i, 0]
        c, *_ = np.linalg.lstsq(phi, rhs, rcond=None)
        for v in range(4):
            urec[i, 9 * v:9 * v + 9] = c * (v + 1)
    orc = oracle.Oracle(m, st)
    new = orc.jacobi_step(means, urec)
    assert rel_err(new, urec) <= 1e-10


@pytest.mark.parametrize("kind", [0, 2])
def test_converged_point_values_do_not_depend_on_the_basis(kind):
    """Converge the sweep from zero on the same cell means in global and in
    local coordinates: the reconstruction at every face Gauss point agrees
    (the two bases span the same polynomials and the functional is written in
    physical derivatives)."""
    import oracle
    s = SETTINGS["iv"]
    out = []
    for lc in (0, 1):
        mesh, vfv = mesh_and_setup(kind=kind, nx=8, ny=7, jitter=0.2, seed=9, bc=(2, 3, 2, 3), limiter=False,
                                   use_local_coord=lc, **s)
        m, st = mesh.arrays(), vfv.arrays()
        bary = st["cell_bary"].reshape(-1, 2)
        rng = np.random.default_rng(1)
        means = np.stack([1 + 0.3 * np.sin(3 * bary[:, 0] + 2 * bary[:, 1]) + 0.01 * rng.standard_normal(len(bary)),
                          0.2 + 0.1 * bary[:, 0] * bary[:, 1], -0.1 + 0.2 * bary[:, 1] ** 2,
                          2.5 + 0.1 * np.cos(4 * bary[:, 0])], 1)
        orc = oracle.Oracle(m, st, far_field=(1.0, 0.2, -0.1, 2.5))
        urec = np.zeros((m["n_cell"], 36))
        for _ in range(400):
            urec = orc.jacobi_step(means, urec)
        d0 = st["face_d0"].reshape(-1, 3, 2, 9)
        f2c = m["face2cell"]
        vals = np.zeros((m["n_face"], 3, 4))
        for g in range(3):
            vals[:, g] = np.einsum("fk,fkv->fv", d0[:, g, 0], urec[f2c[:, 0]].reshape(-1, 4, 9).transpose(0, 2, 1))
        out.append(vals)
    assert rel_err(out[1], out[0]) <= 1e-9, rel_err(out[1], out[0])
    assert np.abs(out[0]).max() > 1e-3
