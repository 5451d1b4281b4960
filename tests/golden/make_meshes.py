"""Writes the Gmsh 2.2 ASCII mesh fixtures of the reader tests:

* quad_6x4.msh      -- 6 x 4 quadrangles on the unit square, boundary lines
                       in two physical groups that share the name "bc-2"
                       (Farfield), some written against the boundary's
                       direction;
* tri_delaunay_a.msh -- a Delaunay triangulation of the unit square with 20
                       boundary nodes and 20 seeded interior nodes, boundary
                       "bc-1" (Wall);
* tri_delaunay_b.msh -- the same with 40 boundary and 100 interior nodes.

The files use the layout of an external mesh exporter (comment blocks
between sections, right-aligned columns, a $PhysicalNames block at the end)
so that the reader meets formatting it did not write itself.  Run from any
directory: python tests/golden/make_meshes.py
"""
import os

import numpy as np
from scipy.spatial import Delaunay

HERE = os.path.dirname(os.path.abspath(__file__))

RULE = "-" * 55


def _comment(lines):
    return ["$Comments", RULE, *lines, RULE, "$EndComments"]


def write(path, xy, cells, lines, names):
    """xy [n][2]; cells: lists of 0-based CCW node ids (3 or 4); lines:
    (physical id, node a, node b); names: (dim, id, name)."""
    out = _comment(["Gmsh 2.2 ASCII mesh written by tests/golden/make_meshes.py"])
    out += ["$MeshFormat", "2.2 0 8", "$EndMeshFormat"]
    out += _comment(["NumNodes", "Id x y z"])
    out += ["$Nodes", str(len(xy))]
    w = len(str(len(xy)))
    out += [f"{i + 1:>{w}} {x!r} {y!r} 0" for i, (x, y) in enumerate(xy.tolist())]
    out += ["$EndNodes"]
    out += _comment(["Types: line(1) triangle(2) quadrangle(3)", "NumElements",
                     "Id Type NumTags PhysGrp ElemGrp IndexList"])
    n_el = len(cells) + len(lines)
    we, wn = len(str(n_el)), len(str(len(xy)))
    out += ["$Elements", str(n_el)]
    k = 1
    for c in cells:
        t = 2 if len(c) == 3 else 3
        out.append(f"{k:>{we}}  {t} 2 1 1 " + " ".join(f"{n + 1:>{wn}}" for n in c))
        k += 1
    for phys, a, b in lines:
        out.append(f"{k:>{we}}  1 2 {phys} {phys} {a + 1:>{wn}} {b + 1:>{wn}}")
        k += 1
    out += ["$EndElements"]
    out += _comment(["NumNames", "Dim Id Name"])
    out += ["$PhysicalNames", str(len(names))] + [f'{d} {i} "{s}"' for d, i, s in names] + ["$EndPhysicalNames"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def quad_grid(nx, ny):
    xs, ys = np.linspace(0, 1, nx + 1), np.linspace(0, 1, ny + 1)
    xy = np.array([(x, y) for y in ys for x in xs])
    nid = lambda i, j: j * (nx + 1) + i  # noqa: E731
    cells = [[nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)] for j in range(ny) for i in range(nx)]
    lines = []
    lines += [(2, nid(i, 0), nid(i + 1, 0)) for i in range(nx)]              # bottom, along the boundary
    lines += [(2, nid(i + 1, ny), nid(i, ny)) for i in reversed(range(nx))]  # top, along the boundary
    lines += [(3, nid(nx, j), nid(nx, j + 1)) for j in range(ny)]            # right
    lines += [(3, nid(0, j), nid(0, j + 1)) for j in reversed(range(ny))]    # left, against the boundary
    return xy, cells, lines


def delaunay_square(n_side, n_inner, seed):
    """n_side boundary nodes per side (corners once), n_inner seeded interior
    nodes kept at least 0.3 h from the boundary and from each other."""
    h = 1.0 / n_side
    t = np.arange(n_side) * h
    bnd = np.concatenate([np.stack([t, 0 * t], 1), np.stack([1 + 0 * t, t], 1), np.stack([1 - t, 1 + 0 * t], 1),
                          np.stack([0 * t, 1 - t], 1)])
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n_inner:
        p = rng.uniform(0.3 * h, 1 - 0.3 * h, 2)
        if all(np.hypot(*(p - q)) > 0.3 * h for q in pts):
            pts.append(p)
    xy = np.concatenate([bnd, np.array(pts)])
    tri = Delaunay(xy).simplices
    cells = []
    for a, b, c in tri.tolist():
        (xa, ya), (xb, yb), (xc, yc) = xy[a], xy[b], xy[c]
        if (xb - xa) * (yc - ya) - (xc - xa) * (yb - ya) < 0:
            b, c = c, b
        cells.append([a, b, c])
    nb = len(bnd)
    lines = [(2, k, (k + 1) % nb) for k in range(nb)]
    return xy, cells, lines


def main():
    xy, cells, lines = quad_grid(6, 4)
    write(os.path.join(HERE, "quad_6x4.msh"), xy, cells, lines,
          [(2, 1, "fluid"), (1, 2, "bc-2"), (1, 3, "bc-2")])
    for name, (ns, ni, seed) in {"tri_delaunay_a.msh": (5, 20, 1), "tri_delaunay_b.msh": (10, 100, 2)}.items():
        xy, cells, lines = delaunay_square(ns, ni, seed)
        write(os.path.join(HERE, name), xy, cells, lines, [(2, 1, "fluid"), (1, 2, "bc-1")])


if __name__ == "__main__":
    main()
