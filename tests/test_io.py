"""Mesh and solution files: the Gmsh 2.2 reader (DNDS_Mesh.hpp:87-350,
760-835) and the Tecplot binary writer (DNDS_Mesh.hpp:1634-1814).  CPU only.

Fixtures: tests/golden/quad_6x4.msh, tri_delaunay_a.msh and
tri_delaunay_b.msh, written by tests/golden/make_meshes.py in an external
exporter's layout (comment blocks, aligned columns, two physical groups of
one name, boundary lines against the boundary's direction).  The reader is
checked against a plain Python parse of the same files fed to the topology
builder, against the generator through a written file, and on its error
paths; the writer against a parse of the Tecplot binary format (Tecplot 360
data format guide, version 112).
"""
import os
import struct

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
BC_NAMES = {1: "bc-1", 2: "bc-2", 3: "bc-3", 4: "bc-4", -2: "bc-DMRFar", -3: "bc-RTFar", -4: "bc-IVFar",
            -5: "bc-2DRiemannFar"}


def parse_gmsh(path):
    """Gmsh 2.2 ASCII: nodes by id, elements by id, physical names."""
    lines = open(path).read().splitlines()
    nodes, elems, names = {}, {}, {}
    k = 0
    while k < len(lines):
        s = lines[k].strip()
        k += 1
        if s == "$Nodes":
            n = int(lines[k])
            for ln in lines[k + 1:k + 1 + n]:
                f = ln.split()
                nodes[int(f[0])] = (float(f[1]), float(f[2]))
            k += 1 + n
        elif s == "$Elements":
            n = int(lines[k])
            for ln in lines[k + 1:k + 1 + n]:
                f = [int(x) for x in ln.split()]
                elems[f[0]] = (f[1], f[3], [x - 1 for x in f[3 + f[2]:]])
            k += 1 + n
        elif s == "$PhysicalNames":
            n = int(lines[k])
            for ln in lines[k + 1:k + 1 + n]:
                f = ln.split(maxsplit=2)
                names[int(f[1])] = (int(f[0]), f[2].strip('"'))
            k += 1 + n
    return nodes, elems, names


def from_parse(path):
    from dnds_amd.host import CompactFacedMeshSerialRW
    nodes, elems, names = parse_gmsh(path)
    xy = np.array([nodes[i + 1] for i in range(len(nodes))])
    cells, bnd, bc = [], [], []
    inv = {v: k for k, v in BC_NAMES.items()}
    for i in sorted(elems):
        t, phys, nl = elems[i]
        dim, name = names[phys]
        if dim == 2:
            cells.append(nl)
        elif dim == 1:
            bnd.append(nl)
            bc.append(inv[name])
    cp = np.cumsum([0] + [len(c) for c in cells])
    return CompactFacedMeshSerialRW.from_cells(xy, cp, np.concatenate(cells), np.array(bnd), np.array(bc))


def same_topology(a, b):
    A, B = a.arrays(), b.arrays()
    for key in ("node_xy", "cell_ptr", "cell2node", "cell2face", "face2node", "face2cell", "face2cell_ref",
                "face_bc"):
        assert np.array_equal(A[key], B[key]), key


@pytest.mark.parametrize("name,n_cell,n_face,bcs", [("quad_6x4.msh", 24, 58, {2: 20}),
                                                    ("tri_delaunay_a.msh", 58, 97, {1: 20}),
                                                    ("tri_delaunay_b.msh", 238, 377, {1: 40})])
def test_reads_the_mesh_files(name, n_cell, n_face, bcs):
    from dnds_amd.host import CompactFacedMeshSerialRW
    path = os.path.join(GOLD, name)
    m = CompactFacedMeshSerialRW.read_gmsh(path)
    assert (m.n_cell, m.n_face) == (n_cell, n_face)
    a = m.arrays()
    bc, cnt = np.unique(a["face_bc"][a["face_bc"] != -1], return_counts=True)
    assert dict(zip(bc.tolist(), cnt.tolist())) == bcs
    same_topology(m, from_parse(path))
    # the unit square: Euler's formula, total area, outward normals of f2c[0]
    xy = a["node_xy"].reshape(-1, 2)
    assert len(xy) - m.n_face + m.n_cell == 1
    cp, c2n = a["cell_ptr"], a["cell2node"]
    area, cen = 0.0, np.zeros((m.n_cell, 2))
    for i in range(m.n_cell):
        p = xy[c2n[cp[i]:cp[i + 1]]]
        q = np.roll(p, -1, axis=0)
        area += 0.5 * np.sum(p[:, 0] * q[:, 1] - q[:, 0] * p[:, 1])
        cen[i] = p.mean(0)
    assert abs(area - 1.0) < 1e-12
    f2n, f2c = a["face2node"].reshape(-1, 2), a["face2cell"].reshape(-1, 2)
    d = xy[f2n[:, 1]] - xy[f2n[:, 0]]
    normal = np.stack([d[:, 1], -d[:, 0]], 1)
    mid = 0.5 * (xy[f2n[:, 0]] + xy[f2n[:, 1]])
    assert np.all(np.sum(normal * (mid - cen[f2c[:, 0]]), 1) > 0)
    assert np.all((f2c[:, 1] >= 0) == (a["face_bc"] == -1))


def write_gmsh(path, mesh, point_group=False):
    a = mesh.arrays()
    xy = a["node_xy"].reshape(-1, 2)
    cp, c2n = a["cell_ptr"], a["cell2node"]
    f2n, bc = a["face2node"].reshape(-1, 2), a["face_bc"]
    kinds = sorted(set(bc[bc != -1].tolist()))
    phys = {k: 2 + j for j, k in enumerate(kinds)}
    rows = []
    for i in range(mesh.n_cell):
        nl = c2n[cp[i]:cp[i + 1]]
        rows.append((2 if len(nl) == 3 else 3, 1, nl))
    for f in np.nonzero(bc != -1)[0]:
        rows.append((1, phys[int(bc[f])], f2n[f]))
    if point_group:
        rows.append((15, 99, [0]))
    with open(path, "w") as fo:
        fo.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Comments\nNumNodes\n$EndComments\n")
        fo.write(f"$Nodes\n{len(xy)}\n")
        for i, (x, y) in enumerate(xy):
            fo.write(f"{i + 1} {float(x)!r} {float(y)!r} 0\n")
        fo.write(f"$EndNodes\n$Elements\n{len(rows)}\n")
        for i, (t, p, nl) in enumerate(rows):
            fo.write(f"{i + 1} {t} 2 {p} {p} " + " ".join(str(int(v) + 1) for v in nl) + "\n")
        fo.write("$EndElements\n$PhysicalNames\n")
        fo.write(f"{len(kinds) + 1 + int(point_group)}\n2 1 \"fluid\"\n")
        for k in kinds:
            fo.write(f"1 {phys[k]} \"{BC_NAMES[k]}\"\n")
        if point_group:
            fo.write("0 99 \"corner\"\n")
        fo.write("$EndPhysicalNames\n")


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_generated_mesh_round_trips_through_a_gmsh_file(tmp_path, kind):
    from dnds_amd.host import CompactFacedMeshSerialRW
    m = CompactFacedMeshSerialRW.generate(kind, 7, 5, 0.0, 2.0, 0.0, 1.0, jitter=0.2, seed=11, bc=(2, 3, 4, -5))
    path = str(tmp_path / "m.msh")
    write_gmsh(path, m, point_group=True)
    same_topology(CompactFacedMeshSerialRW.read_gmsh(path), m)


def edit(src, dst, old, new):
    s = open(src).read()
    assert old in s
    open(dst, "w").write(s.replace(old, new, 1))


@pytest.mark.parametrize("old,new,msg", [
    ('"bc-2"', '"bc-7"', "boundary group"),
    ("2.2 0 8", "2.2 1 8", "ASCII"),
    ("2.2 0 8", "4.1 0 8", "2.x"),
    ("1  3 2 1 1  1  2  9  8", "1  9 2 1 1  1  2  9  8 3 4 5", "not supported"),
    ("1  3 2 1 1  1  2  9  8", "1  3 2 1 1  1  8  9  2", "counter-clockwise"),
    ("1  3 2 1 1  1  2  9  8", "1  3 2 5 1  1  2  9  8", "no $PhysicalNames"),
    ("25  1 2 2 2  1  2", "25  1 2 2 2  9 10", "not on a boundary face"),
    ("25  1 2 2 2  1  2", "25  1 2 2 2  1  99", "outside"),
    ("$EndNodes", "$EndNodez", "not concluded"),
    ("$Nodes\n35\n", "$Nodes\n99999999999999\n", "bad node count"),
    ("$Elements\n", "$Elements\n-3\n", "bad element count"),
])
def test_reader_errors(tmp_path, old, new, msg):
    from dnds_amd import _capi as A
    from dnds_amd.host import CompactFacedMeshSerialRW, DNDSError
    path = str(tmp_path / "bad.msh")
    edit(os.path.join(GOLD, "quad_6x4.msh"), path, old, new)
    with pytest.raises(DNDSError) as e:
        CompactFacedMeshSerialRW.read_gmsh(path)
    assert msg in str(e.value) and e.value.code == A.ERR_INVALID_ARGUMENT
    with pytest.raises(DNDSError) as e:
        CompactFacedMeshSerialRW.read_gmsh(str(tmp_path / "missing.msh"))
    assert e.value.code == A.ERR_IO


def read_plt(path):
    """The subset of the Tecplot binary format (v112) the writer uses: one
    finite-element zone with nodal coordinates and cell-centred data."""
    b = open(path, "rb").read()
    o = 0

    def i32():
        nonlocal o
        v = struct.unpack_from("<i", b, o)[0]
        o += 4
        return v

    def f32():
        nonlocal o
        v = struct.unpack_from("<f", b, o)[0]
        o += 4
        return v

    def f64(n=1):
        nonlocal o
        v = np.frombuffer(b, "<f8", n, o)
        o += 8 * n
        return v

    def text():
        s = []
        while True:
            c = i32()
            if c == 0:
                return "".join(s)
            s.append(chr(c))

    out = {"magic": b[:8].decode()}
    o = 8
    out["byte_order"], out["file_type"], out["title"] = i32(), i32(), text()
    nv = i32()
    out["names"] = [text() for _ in range(nv)]
    assert f32() == 299.0
    out["zone"], _, _, out["time"] = text(), i32(), i32(), f64()[0]
    _, out["zone_type"], specified = i32(), i32(), i32()
    out["location"] = [i32() for _ in range(nv)] if specified else [0] * nv
    assert i32() == 0 and i32() == 0
    out["n_node"], out["n_elem"] = i32(), i32()
    assert (i32(), i32(), i32(), i32()) == (0, 0, 0, 0)
    assert f32() == 357.0 and f32() == 299.0
    out["format"] = [i32() for _ in range(nv)]
    assert (i32(), i32(), i32()) == (0, 0, -1)
    out["minmax"] = f64(2 * nv).reshape(nv, 2)
    out["values"] = [f64(out["n_node"] if loc == 0 else out["n_elem"]) for loc in out["location"]]
    out["conn"] = np.frombuffer(b, "<i4", 4 * out["n_elem"], o).reshape(-1, 4)
    o += 16 * out["n_elem"]
    assert o == len(b)
    return out


def test_tecplot_file_holds_the_mesh_and_the_cell_data(tmp_path):
    from dnds_amd.host import CompactFacedMeshSerialRW
    m = CompactFacedMeshSerialRW.generate(2, 6, 4, 0.0, 3.0, 0.0, 1.0, jitter=0.2, seed=5)
    a = m.arrays()
    rng = np.random.default_rng(0)
    data = rng.standard_normal((3, m.n_cell))
    path = str(tmp_path / "out.plt")
    m.PrintSerialPartPltBinaryDataArray(path, ["R", "U", "P"], data, t=0.25)
    r = read_plt(path)
    assert r["magic"] == "#!TDV112" and (r["byte_order"], r["file_type"]) == (1, 0)
    assert r["names"] == ["X", "Y", "Z", "R", "U", "P", "iPart"]
    assert r["zone_type"] == 3 and r["time"] == 0.25
    assert r["location"] == [0, 0, 0, 1, 1, 1, 1] and r["format"] == [2] * 7
    assert (r["n_node"], r["n_elem"]) == (a["node_xy"].size // 2, m.n_cell)
    xy = a["node_xy"].reshape(-1, 2)
    assert np.array_equal(r["values"][0], xy[:, 0]) and np.array_equal(r["values"][1], xy[:, 1])
    assert np.all(r["values"][2] == 0)
    for k in range(3):
        assert np.array_equal(r["values"][3 + k], data[k])
        assert tuple(r["minmax"][3 + k]) == (data[k].min(), data[k].max())
    assert tuple(r["minmax"][0]) == (xy[:, 0].min(), xy[:, 0].max())
    assert np.all(r["values"][6] == 0) and tuple(r["minmax"][6]) == (0.0, 1.0)
    cp, c2n = a["cell_ptr"], a["cell2node"]
    for i in range(m.n_cell):
        nl = list(c2n[cp[i]:cp[i + 1]])
        assert list(r["conn"][i]) == (nl if len(nl) == 4 else nl + [nl[2]])


def test_cpp_host_classes_read_and_write(tmp_path):
    """tests/cpp/io_run.cpp through the C++ host classes (host code only)."""
    import subprocess
    lib = os.path.join(ROOT, "dnds_amd")
    exe = str(tmp_path / "io_run")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "io_run.cpp"), "-o", exe, "-L" + lib, "-ldndsmi",
                        "-Wl,-rpath," + lib], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "o.plt")
    r = subprocess.run([exe, os.path.join(GOLD, "quad_6x4.msh"), out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["24", "58"]
    p = read_plt(out)
    assert p["names"] == ["X", "Y", "Z", "A", "B", "iPart"] and p["time"] == 1.5
    assert np.array_equal(p["values"][4], 1000.0 + np.arange(24))
    r = subprocess.run([exe, str(tmp_path / "none.msh"), out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_solver_output_fields(tmp_path):
    """EulerSolver::PrintData's eight NS_2D slots, written and read back."""
    from dnds_amd.host import CompactFacedMeshSerialRW, solver_output
    m = CompactFacedMeshSerialRW.read_gmsh(os.path.join(GOLD, "tri_delaunay_a.msh"))
    n = m.n_cell
    rng = np.random.default_rng(3)
    rho, vx, vy, p = (rng.uniform(0.5, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(0.5, 2, n))
    u = np.stack([rho, rho * vx, rho * vy, p / 0.4 + 0.5 * rho * (vx * vx + vy * vy)], 1)
    ind, rhs = rng.uniform(0, 1, n), rng.standard_normal((n, 4))
    names, data = solver_output(u, ind, rhs, gamma=1.4, Rgas=2.0, smooth_indicator_scale=0.5)
    assert names == ["R", "U", "V", "W", "P", "T", "M", "ifUseLimiter"]
    ref = [rho, vx, vy, p, p / rho / 2.0, np.sqrt((vx * vx + vy * vy) / (1.4 * p / rho)), ind / 0.5, rhs[:, 0]]
    for k in range(8):
        assert np.allclose(data[k], ref[k], rtol=1e-13, atol=1e-15), names[k]
    path = str(tmp_path / "s.plt")
    m.PrintSerialPartPltBinaryDataArray(path, names, data, 2.0)
    r = read_plt(path)
    assert r["names"][3:-1] == names
    for k in range(8):
        assert np.array_equal(r["values"][3 + k], data[k])
