// Mesh topology and the seeded structured-as-unstructured mesh generator.
//
// Topology follows the reference's SerialGmshReader2d::InterpolateTopology
// (DNDS_Mesh.hpp:238-350): cells are walked in order, each cell's faces in
// local order (FaceNodeList, DNDS_Elements.hpp:248-253); a face not seen
// before is created with the walking cell's local node order and that cell as
// f2c[0], a face seen before gets the walking cell as f2c[1].  face2cellRef is
// the local index of the face in each of its cells (DNDS_Mesh.hpp:1330-1361).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <unordered_map>

#include "internal.hpp"

namespace {

// splitmix64 seeding + xoshiro256** (Blackman & Vigna); fixed draw order
// documented at the generator.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9E3779B97F4A7C15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; } // [0, 1)
    bool coin() { return (next() >> 63) != 0; }
};

struct PairHash {
    size_t operator()(const std::pair<int64_t, int64_t> &p) const {
        uint64_t h = uint64_t(p.first) * 0x9E3779B97F4A7C15ULL;
        h ^= uint64_t(p.second) + 0x7F4A7C159E3779B9ULL + (h << 6) + (h >> 2);
        return size_t(h);
    }
};

inline std::pair<int64_t, int64_t> edge_key(int64_t a, int64_t b) { return a < b ? std::make_pair(a, b) : std::make_pair(b, a); }

// local face f of a cell with nv vertices: nodes (f, f+1 mod nv)
inline int face_local_node(int nv, int f, int k) { return (f + k) % nv; }

// Fills cell2face, face2node, face2cell, face2cell_ref, face_shift; boundary
// faces get the kind returned by bc_of(face nodes) (DNDSMI_BC_INNER = none).
template <class FBc>
int build_topology(dndsmi_mesh &m, const std::vector<int64_t> &canon, FBc &&bc_of, std::string &err) {
    const int64_t nc = m.n_cell;
    m.cell2face.assign(m.cell_ptr[nc], -1);
    m.face2node.clear();
    m.face2cell.clear();
    m.face2node.reserve(size_t(nc) * 4);
    m.face2cell.reserve(size_t(nc) * 4);
    std::unordered_map<std::pair<int64_t, int64_t>, int64_t, PairHash> faces;
    faces.reserve(size_t(nc) * 2);
    for (int64_t ic = 0; ic < nc; ic++) {
        const int64_t b = m.cell_ptr[ic];
        const int nv = int(m.cell_ptr[ic + 1] - b);
        if (nv != 3 && nv != 4) {
            err = "cell " + std::to_string(ic) + " has " + std::to_string(nv) + " nodes (3 or 4 expected)";
            return DNDSMI_ERR_INVALID_ARGUMENT;
        }
        for (int f = 0; f < nv; f++) {
            const int64_t n0 = m.cell2node[b + face_local_node(nv, f, 0)];
            const int64_t n1 = m.cell2node[b + face_local_node(nv, f, 1)];
            auto key = edge_key(canon[n0], canon[n1]);
            auto it = faces.find(key);
            if (it == faces.end()) {
                const int64_t iface = int64_t(m.face2cell.size() / 2);
                faces.emplace(key, iface);
                m.face2node.push_back(n0);
                m.face2node.push_back(n1);
                m.face2cell.push_back(ic);
                m.face2cell.push_back(DNDSMI_NO_CELL);
                m.cell2face[b + f] = iface;
            } else {
                const int64_t iface = it->second;
                if (m.face2cell[2 * iface + 1] != DNDSMI_NO_CELL) {
                    err = "face shared by more than two cells";
                    return DNDSMI_ERR_INVALID_ARGUMENT;
                }
                if (m.face2cell[2 * iface] == ic) {
                    err = "cell " + std::to_string(ic) + " meets itself across a face";
                    return DNDSMI_ERR_INVALID_ARGUMENT;
                }
                m.face2cell[2 * iface + 1] = ic;
                m.cell2face[b + f] = iface;
            }
        }
    }
    const int64_t nf = int64_t(m.face2cell.size() / 2);
    m.n_face = nf;
    m.face2cell_ref.assign(size_t(nf) * 2, -1);
    m.face_bc.assign(size_t(nf), DNDSMI_BC_INNER);
    m.face_shift.assign(size_t(nf) * 2, 0.0);
    for (int64_t ic = 0; ic < nc; ic++) {
        const int64_t b = m.cell_ptr[ic];
        const int nv = int(m.cell_ptr[ic + 1] - b);
        for (int f = 0; f < nv; f++) {
            const int64_t iface = m.cell2face[b + f];
            const int side = m.face2cell[2 * iface] == ic ? 0 : 1;
            m.face2cell_ref[2 * iface + side] = f;
        }
    }
    for (int64_t iface = 0; iface < nf; iface++) {
        const int64_t n0 = m.face2node[2 * iface], n1 = m.face2node[2 * iface + 1];
        const int64_t R = m.face2cell[2 * iface + 1];
        if (R == DNDSMI_NO_CELL) {
            int32_t bc = bc_of(n0, n1);
            if (bc == DNDSMI_BC_INNER) {
                err = "boundary face " + std::to_string(iface) + " has no boundary kind";
                return DNDSMI_ERR_INVALID_ARGUMENT;
            }
            m.face_bc[iface] = bc;
            continue;
        }
        // node of R's local face that is (periodically) the face's node 0
        const int64_t b = m.cell_ptr[R];
        const int nv = int(m.cell_ptr[R + 1] - b);
        const int fr = m.face2cell_ref[2 * iface + 1];
        const int64_t r0 = m.cell2node[b + face_local_node(nv, fr, 0)];
        const int64_t r1 = m.cell2node[b + face_local_node(nv, fr, 1)];
        const int64_t match = canon[r0] == canon[n0] ? r0 : r1;
        if (match != n0) {
            m.face_shift[2 * iface + 0] = m.node_xy[2 * n0 + 0] - m.node_xy[2 * match + 0];
            m.face_shift[2 * iface + 1] = m.node_xy[2 * n0 + 1] - m.node_xy[2 * match + 1];
        }
    }
    return DNDSMI_OK;
}

} // namespace

extern "C" int dndsmi_mesh_generate(const dndsmi_mesh_spec *spec, dndsmi_mesh **out) {
    if (!spec || !out)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const int64_t nx = spec->nx, ny = spec->ny;
    if (nx < 1 || ny < 1 || (spec->periodic_x && nx < 2) || (spec->periodic_y && ny < 2) || !(spec->x1 > spec->x0) ||
        !(spec->y1 > spec->y0) || spec->jitter < 0 || spec->jitter >= 0.5 || !(spec->y_ratio > 0) ||
        spec->kind < DNDSMI_MESH_TRI_SPLIT || spec->kind > DNDSMI_MESH_MIXED)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    dndsmi_mesh *m = new (std::nothrow) dndsmi_mesh();
    if (!m)
        return DNDSMI_ERR_OUT_OF_MEMORY;
    try {
        const int64_t nnx = nx + 1, nny = ny + 1;
        m->n_node = nnx * nny;
        m->node_xy.resize(size_t(m->n_node) * 2);
        // y lines: geometric spacing with ratio y_ratio (uniform if 1)
        std::vector<double> yl(nny);
        const double r = spec->y_ratio;
        for (int64_t j = 0; j < nny; j++) {
            double frac = (r == 1.0) ? double(j) / double(ny) : (std::pow(r, double(j)) - 1.0) / (std::pow(r, double(ny)) - 1.0);
            yl[j] = j == ny ? spec->y1 : spec->y0 + (spec->y1 - spec->y0) * frac;
        }
        const double hx = (spec->x1 - spec->x0) / double(nx);
        for (int64_t j = 0; j < nny; j++)
            for (int64_t i = 0; i < nnx; i++) {
                int64_t n = j * nnx + i;
                m->node_xy[2 * n + 0] = i == nx ? spec->x1 : spec->x0 + (spec->x1 - spec->x0) * double(i) / double(nx);
                m->node_xy[2 * n + 1] = yl[j];
            }
        Rng rng(spec->seed);
        // cells: quads row by row; draw order per quad: [keep (MIXED)] [diagonal]
        m->n_cell = 0;
        m->cell_ptr.assign(1, 0);
        m->cell2node.clear();
        m->cell2node.reserve(size_t(nx * ny) * 6);
        auto push_cell = [&](std::initializer_list<int64_t> nodes) {
            for (int64_t n : nodes)
                m->cell2node.push_back(n);
            m->cell_ptr.push_back(int64_t(m->cell2node.size()));
            m->n_cell++;
        };
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                const int64_t n00 = j * nnx + i, n10 = n00 + 1, n01 = n00 + nnx, n11 = n01 + 1;
                bool quad = spec->kind == DNDSMI_MESH_QUAD;
                if (spec->kind == DNDSMI_MESH_MIXED)
                    quad = rng.uniform() < 0.5;
                if (quad) {
                    push_cell({n00, n10, n11, n01});
                } else if (!rng.coin()) { // diagonal n00-n11
                    push_cell({n00, n10, n11});
                    push_cell({n00, n11, n01});
                } else { // diagonal n10-n01
                    push_cell({n00, n10, n01});
                    push_cell({n10, n11, n01});
                }
            }
        // jitter of strictly interior nodes, row by row, x draw then y draw
        if (spec->jitter > 0) {
            for (int64_t j = 1; j < ny; j++) {
                const double hy = std::min(yl[j] - yl[j - 1], yl[j + 1] - yl[j]);
                for (int64_t i = 1; i < nx; i++) {
                    int64_t n = j * nnx + i;
                    double dx = (2.0 * rng.uniform() - 1.0) * spec->jitter * hx;
                    double dy = (2.0 * rng.uniform() - 1.0) * spec->jitter * hy;
                    m->node_xy[2 * n + 0] += dx;
                    m->node_xy[2 * n + 1] += dy;
                }
            }
        }
        std::vector<int64_t> canon(size_t(m->n_node));
        for (int64_t j = 0; j < nny; j++)
            for (int64_t i = 0; i < nnx; i++) {
                int64_t ci = spec->periodic_x ? i % nx : i;
                int64_t cj = spec->periodic_y ? j % ny : j;
                canon[size_t(j * nnx + i)] = cj * nnx + ci;
            }
        auto bc_of = [&](int64_t a, int64_t b) -> int32_t {
            const int64_t ia = a % nnx, ja = a / nnx, ib = b % nnx, jb = b / nnx;
            if (ia == 0 && ib == 0)
                return spec->bc_xmin;
            if (ia == nx && ib == nx)
                return spec->bc_xmax;
            if (ja == 0 && jb == 0)
                return spec->bc_ymin;
            if (ja == ny && jb == ny)
                return spec->bc_ymax;
            return DNDSMI_BC_INNER;
        };
        std::string err;
        int rc = build_topology(*m, canon, bc_of, err);
        if (rc != DNDSMI_OK) {
            delete m;
            return rc;
        }
    } catch (const std::bad_alloc &) {
        delete m;
        return DNDSMI_ERR_OUT_OF_MEMORY;
    }
    *out = m;
    return DNDSMI_OK;
}

int mesh_from_cells(int64_t n_node, const double *node_xy, int64_t n_cell, const int64_t *cell_ptr,
                    const int64_t *cell2node, int64_t n_bnd, const int64_t *bnd2node, const int32_t *bnd_bc,
                    dndsmi_mesh **out, std::string &err) {
    if (!out || n_node < 3 || n_cell < 1 || !node_xy || !cell_ptr || !cell2node || n_bnd < 0 ||
        (n_bnd > 0 && (!bnd2node || !bnd_bc))) {
        err = "invalid mesh arguments";
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    dndsmi_mesh *m = new (std::nothrow) dndsmi_mesh();
    if (!m)
        return DNDSMI_ERR_OUT_OF_MEMORY;
    try {
        m->n_node = n_node;
        m->n_cell = n_cell;
        m->node_xy.assign(node_xy, node_xy + 2 * n_node);
        m->cell_ptr.assign(cell_ptr, cell_ptr + n_cell + 1);
        if (m->cell_ptr[0] != 0)
            throw std::invalid_argument("cell_ptr[0] != 0");
        for (int64_t i = 0; i < n_cell; i++)
            if (m->cell_ptr[i + 1] <= m->cell_ptr[i])
                throw std::invalid_argument("cell_ptr not increasing");
        m->cell2node.assign(cell2node, cell2node + m->cell_ptr[n_cell]);
        for (int64_t n : m->cell2node)
            if (n < 0 || n >= n_node)
                throw std::invalid_argument("node index out of range");
        std::unordered_map<std::pair<int64_t, int64_t>, int32_t, PairHash> bmap;
        for (int64_t k = 0; k < n_bnd; k++)
            bmap[edge_key(bnd2node[2 * k], bnd2node[2 * k + 1])] = bnd_bc[k];
        std::vector<int64_t> canon(static_cast<size_t>(n_node));
        for (int64_t i = 0; i < n_node; i++)
            canon[size_t(i)] = i;
        auto bc_of = [&](int64_t a, int64_t b) -> int32_t {
            auto it = bmap.find(edge_key(a, b));
            return it == bmap.end() ? DNDSMI_BC_INNER : it->second;
        };
        int rc = build_topology(*m, canon, bc_of, err);
        if (rc != DNDSMI_OK) {
            delete m;
            return rc;
        }
    } catch (const std::bad_alloc &) {
        delete m;
        err = "out of memory";
        return DNDSMI_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        delete m;
        err = e.what();
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    *out = m;
    return DNDSMI_OK;
}

extern "C" int dndsmi_mesh_from_cells(int64_t n_node, const double *node_xy, int64_t n_cell, const int64_t *cell_ptr,
                                      const int64_t *cell2node, int64_t n_bnd, const int64_t *bnd2node,
                                      const int32_t *bnd_bc, dndsmi_mesh **out) {
    std::string err;
    return mesh_from_cells(n_node, node_xy, n_cell, cell_ptr, cell2node, n_bnd, bnd2node, bnd_bc, out, err);
}

int check_mesh_view(const dndsmi_mesh_view *m, std::string &err) {
    auto bad = [&](const std::string &why) {
        err = why;
        return DNDSMI_ERR_INVALID_ARGUMENT;
    };
    if (!m)
        return bad("null mesh view");
    if (!m->node_xy || !m->cell_ptr || !m->cell2node || !m->cell2face || !m->face2node || !m->face2cell ||
        !m->face2cell_ref || !m->face_bc || !m->face_shift)
        return bad("null array in the mesh view");
    const int64_t nn = m->n_node, nc = m->n_cell, nf = m->n_face;
    if (nn < 3 || nc < 1 || nf < 3)
        return bad("mesh view without cells");
    if (m->cell_ptr[0] != 0)
        return bad("cell_ptr[0] is not 0");
    auto cell = [](int64_t i) { return "cell " + std::to_string(i); };
    auto face = [](int64_t f) { return "face " + std::to_string(f); };
    for (int64_t i = 0; i < nc; i++) {
        const int64_t b = m->cell_ptr[i], n = m->cell_ptr[i + 1] - b;
        if (n != 3 && n != 4)
            return bad(cell(i) + " has " + std::to_string(n) + " faces (triangles and quadrilaterals only)");
        for (int64_t s = 0; s < n; s++) {
            const int64_t v = m->cell2node[b + s], f = m->cell2face[b + s];
            if (v < 0 || v >= nn)
                return bad(cell(i) + ": node index " + std::to_string(v) + " out of range");
            if (f < 0 || f >= nf)
                return bad(cell(i) + ": face index " + std::to_string(f) + " out of range");
            const bool left = m->face2cell[2 * f] == i && m->face2cell_ref[2 * f] == s;
            const bool right = m->face2cell[2 * f + 1] == i && m->face2cell_ref[2 * f + 1] == s;
            if (!left && !right)
                return bad(cell(i) + " lists " + face(f) + " as its local face " + std::to_string(s) +
                           ", but the face's face2cell / face2cell_ref do not name it there");
        }
    }
    for (int64_t f = 0; f < nf; f++) {
        for (int k = 0; k < 2; k++) {
            const int64_t v = m->face2node[2 * f + k];
            if (v < 0 || v >= nn)
                return bad(face(f) + ": node index " + std::to_string(v) + " out of range");
        }
        for (int side = 0; side < 2; side++) {
            const int64_t c = m->face2cell[2 * f + side];
            if (side == 1 && c == DNDSMI_NO_CELL)
                break;
            if (c < 0 || c >= nc)
                return bad(face(f) + ": cell index " + std::to_string(c) + " out of range");
            const int64_t ref = m->face2cell_ref[2 * f + side];
            if (ref < 0 || ref >= m->cell_ptr[c + 1] - m->cell_ptr[c] || m->cell2face[m->cell_ptr[c] + ref] != f)
                return bad(face(f) + ": face2cell_ref does not point back to the face in " + cell(c));
        }
        const int32_t bc = m->face_bc[f];
        if (m->face2cell[2 * f + 1] == DNDSMI_NO_CELL) {
            if (bc != DNDSMI_BC_WALL && bc != DNDSMI_BC_FARFIELD && bc != DNDSMI_BC_WALL_EULER &&
                bc != DNDSMI_BC_WALL_NOSLIP && bc != DNDSMI_BC_SPECIAL_DMRFAR && bc != DNDSMI_BC_SPECIAL_RTFAR &&
                bc != DNDSMI_BC_SPECIAL_IVFAR && bc != DNDSMI_BC_SPECIAL_2DRIEMANNFAR)
                return bad(face(f) + ": boundary face with unknown boundary kind " + std::to_string(bc));
        }
    }
    return DNDSMI_OK;
}

extern "C" int dndsmi_mesh_view_check(const dndsmi_mesh_view *view, char *err, size_t err_len) {
    std::string why;
    const int rc = check_mesh_view(view, why);
    if (rc != DNDSMI_OK && err && err_len)
        std::snprintf(err, err_len, "%s", why.c_str());
    return rc;
}

extern "C" int dndsmi_mesh_get_view(const dndsmi_mesh *m, dndsmi_mesh_view *v) {
    if (!m || !v)
        return DNDSMI_ERR_INVALID_ARGUMENT;
    v->n_node = m->n_node;
    v->n_cell = m->n_cell;
    v->n_face = m->n_face;
    v->node_xy = m->node_xy.data();
    v->cell_ptr = m->cell_ptr.data();
    v->cell2node = m->cell2node.data();
    v->cell2face = m->cell2face.data();
    v->face2node = m->face2node.data();
    v->face2cell = m->face2cell.data();
    v->face2cell_ref = m->face2cell_ref.data();
    v->face_bc = m->face_bc.data();
    v->face_shift = m->face_shift.data();
    return DNDSMI_OK;
}

extern "C" void dndsmi_mesh_destroy(dndsmi_mesh *m) { delete m; }
