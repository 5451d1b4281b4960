"""Mesh topology and the reference's conventions (DNDS_Mesh.hpp:238-350,
1330-1361; normal orientation DNDS_FV_VR.cpp:579).  CPU only."""
import numpy as np
import pytest

from conftest import mesh_and_setup

KINDS = [0, 1, 2]


def gen(kind, periodic=False, nx=7, ny=5, jitter=0.2, seed=9, bc=None):
    from dnds_amd import _capi as A
    from dnds_amd.host import CompactFacedMeshSerialRW
    bc = bc or (A.BC_FARFIELD,) * 4
    return CompactFacedMeshSerialRW.generate(kind, nx, ny, 0.0, 2.0, 0.0, 1.0, periodic=(periodic, periodic),
                                             jitter=0.0 if periodic else jitter, seed=seed, bc=bc)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("periodic", [False, True])
def test_faces_have_one_or_two_cells_and_refs_agree(kind, periodic):
    m = gen(kind, periodic).arrays()
    f2c, c2f, cp, ref = m["face2cell"], m["cell2face"], m["cell_ptr"], m["face2cell_ref"]
    from dnds_amd import _capi as A
    counts = np.zeros(m["n_face"], int)
    for i in range(m["n_cell"]):
        for lf, f in enumerate(c2f[cp[i]:cp[i + 1]]):
            counts[f] += 1
            side = 0 if f2c[f, 0] == i else 1
            assert f2c[f, side] == i
            assert ref[f, side] == lf
    inner = f2c[:, 1] != A.NO_CELL
    assert np.all(counts[inner] == 2) and np.all(counts[~inner] == 1)
    if periodic:
        assert inner.all()
    assert np.all(m["face_bc"][inner] == A.BC_INNER)
    assert np.all(m["face_bc"][~inner] == A.BC_FARFIELD)
    # Euler characteristic: torus V - E + F = 0, disc = 1
    nodes_used = np.unique(m["cell2node"]).size
    if not periodic:
        assert nodes_used - m["n_face"] + m["n_cell"] == 1


@pytest.mark.parametrize("kind", KINDS)
def test_faces_numbered_by_first_appearance(kind):
    """Faces appear in the order the cells, and each cell's local faces, are
    walked; the face's nodes are its first cell's local face nodes."""
    m = gen(kind).arrays()
    f2c, c2f, cp, c2n, f2n = m["face2cell"], m["cell2face"], m["cell_ptr"], m["cell2node"], m["face2node"]
    seen = 0
    for i in range(m["n_cell"]):
        nodes = c2n[cp[i]:cp[i + 1]]
        nv = len(nodes)
        for lf, f in enumerate(c2f[cp[i]:cp[i + 1]]):
            if f2c[f, 0] == i:
                assert f == seen
                seen += 1
                assert tuple(f2n[f]) == (nodes[lf], nodes[(lf + 1) % nv])
    assert seen == m["n_face"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("periodic", [False, True])
def test_normals_point_out_of_first_cell_and_cells_close(kind, periodic):
    mesh, vfv = mesh_and_setup(kind=kind, nx=7, ny=5, periodic=periodic, jitter=0.0 if periodic else 0.2, x1=2.0)
    m, s = mesh.arrays(), vfv.arrays()
    n = s["face_norm_center"].reshape(-1, 2)
    ctr = s["face_center"].reshape(-1, 2)
    bary = s["cell_bary"].reshape(-1, 2)
    f2c, cp, c2f = m["face2cell"], m["cell_ptr"], m["cell2face"]
    assert np.all(np.einsum("ij,ij->i", n, bary[f2c[:, 0]] - ctr) < 0)
    # sum over faces of the outward normal times the face length is zero
    for i in range(m["n_cell"]):
        acc = np.zeros(2)
        for f in c2f[cp[i]:cp[i + 1]]:
            sgn = 1.0 if f2c[f, 0] == i else -1.0
            acc += sgn * 2.0 * n[f]  # |face_norm| is half the face length (Line2 Jacobian)
        assert np.abs(acc).max() < 1e-13
    area = 2.0 * 1.0
    assert abs(s["volume"].sum() - area) < 1e-12 * area
    assert np.all(s["volume"] > 0)
    assert np.allclose(s["face_area"], 2.0 * np.linalg.norm(n, axis=1), rtol=1e-14)


def test_periodic_shift_moves_second_cell_next_to_the_face():
    m = gen(0, periodic=True, nx=6, ny=4).arrays()
    xy, f2n, f2c, sh, cp, c2n, ref = (m["node_xy"], m["face2node"], m["face2cell"], m["face_shift"],
                                      m["cell_ptr"], m["cell2node"], m["face2cell_ref"])
    n_shift = 0
    for f in range(m["n_face"]):
        R = f2c[f, 1]
        nodes = c2n[cp[R]:cp[R + 1]]
        lf = ref[f, 1]
        rn = xy[[nodes[lf], nodes[(lf + 1) % len(nodes)]]] + sh[f]
        fx = xy[f2n[f]]
        # R walks the shared edge in the opposite direction
        assert np.allclose(rn[::-1], fx, atol=1e-14)
        n_shift += np.any(sh[f] != 0)
    assert n_shift == 6 + 4  # one column and one row of periodic faces


def test_generator_is_deterministic_and_seeded():
    a = gen(0, seed=3).arrays()
    b = gen(0, seed=3).arrays()
    c = gen(0, seed=4).arrays()
    for k in ("node_xy", "cell2node", "face2cell"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["cell2node"], c["cell2node"])


@pytest.mark.parametrize("kind", KINDS)
def test_from_cells_rebuilds_the_same_topology(kind):
    """The reader path (raw cells + boundary edges) builds the generator's faces."""
    from dnds_amd.host import CompactFacedMeshSerialRW
    from dnds_amd import _capi as A
    g = gen(kind).arrays()
    bnd = g["face2cell"][:, 1] == A.NO_CELL
    r = CompactFacedMeshSerialRW.from_cells(g["node_xy"], g["cell_ptr"], g["cell2node"], g["face2node"][bnd],
                                            g["face_bc"][bnd]).arrays()
    for k in ("cell2face", "face2node", "face2cell", "face2cell_ref", "face_bc"):
        assert np.array_equal(g[k], r[k]), k


def test_from_cells_rejects_boundary_without_kind():
    from dnds_amd.host import CompactFacedMeshSerialRW, DNDSError
    xy = np.array([[0, 0], [1, 0], [0, 1.0]])
    with pytest.raises(DNDSError):
        CompactFacedMeshSerialRW.from_cells(xy, [0, 3], [0, 1, 2], np.zeros((0, 2), np.int64), np.zeros(0, np.int32))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("periodic", [False, True])
def test_mesh_view_check_accepts_generated_meshes(kind, periodic):
    from dnds_amd import _capi as A
    from dnds_amd.host import ExternalMesh
    m = gen(kind, periodic)
    assert A.load().dndsmi_mesh_view_check(m.view, None, 0) == A.OK
    ExternalMesh(m.arrays())


def _corrupt(fn):
    def apply(a):
        a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
        fn(a)
        return a
    return apply


def _inner_face(a):
    from dnds_amd import _capi as A
    return int(np.nonzero(a["face2cell"][:, 1] != A.NO_CELL)[0][0])


def _boundary_face(a):
    from dnds_amd import _capi as A
    return int(np.nonzero(a["face2cell"][:, 1] == A.NO_CELL)[0][0])


CORRUPTIONS = [
    pytest.param(_corrupt(lambda a: a["cell2face"].__setitem__(0, a["n_face"])), "face index", id="face-index"),
    pytest.param(_corrupt(lambda a: a["cell2node"].__setitem__(1, -1)), "node index -1 out of range",
                 id="node-index"),
    pytest.param(_corrupt(lambda a: a["face2node"].__setitem__((2, 0), a["n_node"])), "face 2: node index",
                 id="face-node"),
    pytest.param(_corrupt(lambda a: a["face2cell_ref"].__setitem__((0, 0), (a["face2cell_ref"][0, 0] + 1) % 3)),
                 "do not name it there", id="ref"),
    pytest.param(_corrupt(lambda a: a["face2cell"].__setitem__((_inner_face(a), 1), a["n_cell"] + 5)),
                 "do not name it there", id="cell-index"),
    pytest.param(_corrupt(lambda a: a["face_bc"].__setitem__(_boundary_face(a), 7)), "unknown boundary kind 7",
                 id="bc-kind"),
    pytest.param(_corrupt(lambda a: a["cell_ptr"].__setitem__(1, 5)), "cell 0 has 5 faces", id="five-faces"),
]


@pytest.mark.parametrize("corrupt,message", CORRUPTIONS)
def test_mesh_view_check_names_the_broken_index(corrupt, message):
    """A caller-owned mesh view is checked before any array built from it
    reaches the device: each broken index is refused with its description,
    by the check itself and by dndsmi_vr_create."""
    import ctypes as C
    from dnds_amd import _capi as A
    from dnds_amd.host import DNDSError, ExternalMesh, vr_setting
    good = gen(0).arrays()
    bad = corrupt(good)
    with pytest.raises(DNDSError) as e:
        ExternalMesh(bad)
    assert e.value.code == A.ERR_INVALID_ARGUMENT and message in str(e.value), str(e.value)
    # the same view handed straight to the setup
    m = ExternalMesh(good)
    for k, v in bad.items():
        if isinstance(v, np.ndarray):
            m._keep[k][:] = v.reshape(-1)
    out = C.c_void_p()
    s = vr_setting()
    assert A.load().dndsmi_vr_create(C.byref(m.view), C.byref(s), 1, C.byref(out)) == A.ERR_INVALID_ARGUMENT
    assert not out.value


@pytest.mark.parametrize("field", ["volume", "face_norm", "ainv", "ainv_b", "ainv_bvec", "face_gauss_xy"])
def test_context_refuses_a_setup_view_without_a_required_array(field):
    """dndsmi_ctx_create checks the caller's setup view before anything is
    read from it or allocated on the device: a missing array it needs is an
    invalid argument (matrix_secondary and face_dibj_g0 alone may be absent).
    Host-only: the check returns before any device call."""
    import ctypes as C
    from conftest import mesh_and_setup
    from dnds_amd import _capi as A
    mesh, vfv = mesh_and_setup(nx=4, ny=3)
    vv = A.VRView()
    C.memmove(C.byref(vv), C.byref(vfv.view), C.sizeof(A.VRView))
    setattr(vv, field, A.PD())
    es = A.EulerSetting(riemann_solver=A.RS_ROE, gamma=1.4, far_field=(C.c_double * 4)(1.0, 0.3, 0.0, 2.6))
    out = C.c_void_p()
    rc = A.load().dndsmi_ctx_create(C.byref(mesh.view), C.byref(vv), C.byref(vfv.setting), C.byref(es), None,
                                    C.byref(out))
    assert rc == A.ERR_INVALID_ARGUMENT and not out.value
