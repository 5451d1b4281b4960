// Mesh and solution files on either side of the hot path: the Gmsh 2.2 ASCII
// reader (GmshReader::FileRead + InterpolateTopology + the boundary mapping of
// CompactFacedMeshSerialRWBuild, DNDS_Mesh.hpp:87-350, 760-835) and the
// Tecplot binary writer (PrintSerialPartPltBinaryDataArray,
// DNDS_Mesh.hpp:1634-1814; format: Tecplot 360 data format guide, #!TDV112).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace {

void put_err(char *buf, size_t n, const std::string &msg) {
    if (buf && n) {
        std::snprintf(buf, n, "%s", msg.c_str());
    }
}

// Nodes per Gmsh element type for the first-order 2D types (and points,
// which 2D meshes carry for their corner groups).
int gmsh_nodes(int type) {
    switch (type) {
    case 1: return 2;  // 2-node line
    case 2: return 3;  // 3-node triangle
    case 3: return 4;  // 4-node quadrangle
    case 15: return 1; // 1-node point
    default: return -1;
    }
}

// Physical-group names the reference maps to boundary kinds
// (DNDS_Mesh.hpp:803-824); any other name on a boundary group is an error.
bool bc_of_name(const std::string &name, int32_t &bc) {
    static const std::map<std::string, int32_t> kinds = {
        {"bc-1", DNDSMI_BC_WALL},
        {"bc-2", DNDSMI_BC_FARFIELD},
        {"bc-3", DNDSMI_BC_WALL_EULER},
        {"bc-4", DNDSMI_BC_WALL_NOSLIP},
        {"bc-DMRFar", DNDSMI_BC_SPECIAL_DMRFAR},
        {"bc-RTFar", DNDSMI_BC_SPECIAL_RTFAR},
        {"bc-IVFar", DNDSMI_BC_SPECIAL_IVFAR},
        {"bc-2DRiemannFar", DNDSMI_BC_SPECIAL_2DRIEMANNFAR},
    };
    auto it = kinds.find(name);
    if (it == kinds.end())
        return false;
    bc = it->second;
    return true;
}

struct Element {
    int type = 0;
    int64_t phys = 0;
    std::vector<int64_t> nodes; // 0-based
    bool seen = false;
};

struct PhysGroup {
    int dim = -1;
    std::string name;
};

// Reads the integers / reals of one line without allocating.
struct LineReader {
    const char *p;
    bool ok = true;
    explicit LineReader(const std::string &s) : p(s.c_str()) {}
    int64_t i() {
        char *e = nullptr;
        long long v = std::strtoll(p, &e, 10);
        if (e == p)
            ok = false;
        p = e;
        return v;
    }
    double d() {
        char *e = nullptr;
        double v = std::strtod(p, &e);
        if (e == p)
            ok = false;
        p = e;
        return v;
    }
};

bool starts_with(const std::string &s, const char *pre) { return s.compare(0, std::strlen(pre), pre) == 0; }

bool blank(const std::string &s) { return s.find_first_not_of(" \t\r") == std::string::npos; }

} // namespace

// node and element counts above this are refused before anything is sized
// by them (2^40: 16 TiB of node coordinates)
constexpr int64_t kMaxCount = int64_t(1) << 40;

extern "C" int dndsmi_mesh_read_gmsh(const char *path, dndsmi_mesh **out, char *err, size_t err_len) {
    if (!path || !out) {
        put_err(err, err_len, "null argument");
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::ifstream fin(path);
    if (!fin) {
        put_err(err, err_len, std::string("cannot open ") + path);
        return DNDSMI_ERR_IO;
    }
    auto fail = [&](const std::string &msg) {
        put_err(err, err_len, std::string(path) + ": " + msg);
        return DNDSMI_ERR_INVALID_ARGUMENT;
    };
    std::vector<double> xy;
    std::vector<char> node_seen;
    std::vector<Element> elems;
    std::map<int64_t, PhysGroup> groups;
    bool have_nodes = false, have_elems = false;
    std::string line;
    try {
        while (std::getline(fin, line)) {
            if (line.empty() || line[0] != '$')
                continue;
            if (starts_with(line, "$MeshFormat")) {
                if (!std::getline(fin, line))
                    return fail("$MeshFormat not concluded");
                LineReader r(line);
                const double version = r.d();
                const int64_t file_type = r.i();
                if (!r.ok || version < 2.0 || version >= 3.0)
                    return fail("only the Gmsh 2.x format is supported");
                if (file_type != 0)
                    return fail("only ASCII Gmsh files are supported");
                while (std::getline(fin, line) && !starts_with(line, "$EndMeshFormat"))
                    ;
            } else if (starts_with(line, "$PhysicalNames")) {
                if (!std::getline(fin, line))
                    return fail("$PhysicalNames not concluded");
                bool closed = false;
                while (std::getline(fin, line)) {
                    if (starts_with(line, "$End")) {
                        closed = true;
                        break;
                    }
                    if (blank(line))
                        continue;
                    LineReader r(line);
                    const int64_t dim = r.i(), id = r.i();
                    const size_t q0 = line.find('"'), q1 = line.rfind('"');
                    if (!r.ok || q0 == std::string::npos || q1 == q0)
                        return fail("bad $PhysicalNames line: " + line);
                    groups[id] = PhysGroup{int(dim), line.substr(q0 + 1, q1 - q0 - 1)};
                }
                if (!closed)
                    return fail("$PhysicalNames not concluded");
            } else if (starts_with(line, "$Nodes")) {
                if (!std::getline(fin, line))
                    return fail("$Nodes not concluded");
                LineReader h(line);
                const int64_t n = h.i();
                if (!h.ok || n < 0 || n > kMaxCount)
                    return fail("bad node count");
                xy.assign(size_t(2 * n), 0.0);
                node_seen.assign(size_t(n), 0);
                bool closed = false;
                while (std::getline(fin, line)) {
                    if (!line.empty() && line[0] == '$') {
                        closed = starts_with(line, "$EndNodes");
                        break;
                    }
                    if (blank(line))
                        continue;
                    LineReader r(line);
                    const int64_t id = r.i();
                    const double x = r.d(), y = r.d();
                    if (!r.ok)
                        return fail("bad node line: " + line);
                    if (id < 1 || id > n)
                        return fail("node id " + std::to_string(id) + " outside 1.." + std::to_string(n));
                    xy[size_t(2 * (id - 1))] = x, xy[size_t(2 * (id - 1) + 1)] = y; // z is dropped
                    node_seen[size_t(id - 1)] = 1;
                }
                if (!closed)
                    return fail("$Nodes not concluded");
                have_nodes = true;
            } else if (starts_with(line, "$Elements")) {
                if (!std::getline(fin, line))
                    return fail("$Elements not concluded");
                LineReader h(line);
                const int64_t n = h.i();
                if (!h.ok || n < 0 || n > kMaxCount)
                    return fail("bad element count");
                elems.assign(size_t(n), Element());
                bool closed = false;
                while (std::getline(fin, line)) {
                    if (!line.empty() && line[0] == '$') {
                        closed = starts_with(line, "$EndElements");
                        break;
                    }
                    if (blank(line))
                        continue;
                    LineReader r(line);
                    const int64_t id = r.i(), type = r.i(), ntag = r.i();
                    if (!r.ok || ntag < 2)
                        return fail("bad element line (two or more tags expected): " + line);
                    const int64_t phys = r.i();
                    for (int64_t k = 1; k < ntag; k++)
                        r.i();
                    if (id < 1 || id > n)
                        return fail("element id " + std::to_string(id) + " outside 1.." + std::to_string(n));
                    const int nn = gmsh_nodes(int(type));
                    if (nn < 0)
                        return fail("element type " + std::to_string(type) +
                                    " is not supported (first-order lines, triangles and quadrangles only)");
                    Element &e = elems[size_t(id - 1)];
                    e.type = int(type), e.phys = phys, e.seen = true;
                    e.nodes.resize(size_t(nn));
                    for (int k = 0; k < nn; k++)
                        e.nodes[size_t(k)] = r.i() - 1;
                    if (!r.ok)
                        return fail("bad element line: " + line);
                }
                if (!closed)
                    return fail("$Elements not concluded");
                have_elems = true;
            } else if (starts_with(line, "$End")) {
                continue;
            } else {
                // $Comments and sections this reader does not use
                bool closed = false;
                while (std::getline(fin, line))
                    if (!line.empty() && line[0] == '$' && line.find("End") != std::string::npos) {
                        closed = true;
                        break;
                    }
                if (!closed)
                    return fail("section not concluded");
            }
        }
        if (!have_nodes || !have_elems)
            return fail("no $Nodes or no $Elements section");
        const int64_t n_node = int64_t(node_seen.size());
        for (int64_t k = 0; k < n_node; k++)
            if (!node_seen[size_t(k)])
                return fail("node " + std::to_string(k + 1) + " missing");
        // cells and boundary edges in element-id order, classified by the
        // dimension of their physical group (DNDS_Mesh.hpp:243-263)
        std::vector<int64_t> cell_ptr{0}, cell2node, bnd2node;
        std::vector<int32_t> bnd_bc;
        std::vector<int64_t> bnd_id;
        for (size_t k = 0; k < elems.size(); k++) {
            const Element &e = elems[k];
            if (!e.seen)
                return fail("element " + std::to_string(k + 1) + " missing");
            for (int64_t v : e.nodes)
                if (v < 0 || v >= n_node)
                    return fail("element " + std::to_string(k + 1) + " names a node outside 1.." +
                                std::to_string(n_node));
            auto g = groups.find(e.phys);
            if (g == groups.end())
                return fail("physical group " + std::to_string(e.phys) + " of element " + std::to_string(k + 1) +
                            " has no $PhysicalNames entry");
            const int dim = g->second.dim;
            if (dim == 2) {
                if (e.type != 2 && e.type != 3)
                    return fail("element " + std::to_string(k + 1) + " in a 2D group is not a triangle or quadrangle");
                // the normal of each face must point out of the cell that
                // creates it (asserted at DNDS_FV_VR.cpp:579): cells counter-clockwise
                double a2 = 0;
                const size_t nv = e.nodes.size();
                for (size_t q = 0; q < nv; q++) {
                    const int64_t p0 = e.nodes[q], p1 = e.nodes[(q + 1) % nv];
                    a2 += xy[size_t(2 * p0)] * xy[size_t(2 * p1 + 1)] - xy[size_t(2 * p1)] * xy[size_t(2 * p0 + 1)];
                }
                if (!(a2 > 0))
                    return fail("cell element " + std::to_string(k + 1) + " is not counter-clockwise");
                cell2node.insert(cell2node.end(), e.nodes.begin(), e.nodes.end());
                cell_ptr.push_back(int64_t(cell2node.size()));
            } else if (dim == 1) {
                if (e.type != 1)
                    return fail("element " + std::to_string(k + 1) + " in a 1D group is not a line");
                int32_t bc;
                if (!bc_of_name(g->second.name, bc))
                    return fail("boundary group \"" + g->second.name + "\" is not one of bc-1..bc-4, bc-DMRFar, "
                                "bc-RTFar, bc-IVFar, bc-2DRiemannFar");
                bnd2node.push_back(e.nodes[0]), bnd2node.push_back(e.nodes[1]);
                bnd_bc.push_back(bc);
                bnd_id.push_back(int64_t(k + 1));
            } else if (dim != 0) {
                return fail("physical group " + std::to_string(e.phys) + " has dimension " + std::to_string(dim));
            }
        }
        const int64_t n_cell = int64_t(cell_ptr.size()) - 1;
        if (n_cell < 1)
            return fail("no cells");
        // each boundary element must lie on a face with one cell (DNDS_Mesh.hpp:321-333)
        std::unordered_map<uint64_t, int> edge_count;
        auto key = [&](int64_t a, int64_t b) {
            if (a > b)
                std::swap(a, b);
            return uint64_t(a) * uint64_t(n_node) + uint64_t(b);
        };
        for (int64_t c = 0; c < n_cell; c++) {
            const int64_t b0 = cell_ptr[size_t(c)], nv = cell_ptr[size_t(c + 1)] - b0;
            for (int64_t q = 0; q < nv; q++)
                edge_count[key(cell2node[size_t(b0 + q)], cell2node[size_t(b0 + (q + 1) % nv)])]++;
        }
        for (size_t k = 0; k < bnd_bc.size(); k++) {
            auto it = edge_count.find(key(bnd2node[2 * k], bnd2node[2 * k + 1]));
            if (it == edge_count.end() || it->second != 1)
                return fail("boundary element " + std::to_string(bnd_id[k]) + " is not on a boundary face");
        }
        std::string why;
        const int rc = mesh_from_cells(n_node, xy.data(), n_cell, cell_ptr.data(), cell2node.data(),
                                       int64_t(bnd_bc.size()), bnd2node.data(), bnd_bc.data(), out, why);
        if (rc != DNDSMI_OK) {
            put_err(err, err_len, std::string(path) + ": " + why);
            return rc;
        }
    } catch (const std::bad_alloc &) {
        put_err(err, err_len, "out of memory");
        return DNDSMI_ERR_OUT_OF_MEMORY;
    } catch (const std::length_error &) {
        return fail("a node or element count too large to hold");
    }
    return DNDSMI_OK;
}

extern "C" int dndsmi_mesh_write_plt(const dndsmi_mesh *m, const char *path, int32_t n_data,
                                     const char *const *names, const double *data, double t, char *err,
                                     size_t err_len) {
    if (!m || !path || n_data < 0 || (n_data > 0 && (!names || !data))) {
        put_err(err, err_len, "null argument");
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    if (m->n_node > INT32_MAX || m->n_cell > INT32_MAX) {
        put_err(err, err_len, "the format's counts are 32-bit");
        return DNDSMI_ERR_INVALID_ARGUMENT;
    }
    std::ofstream fout(path, std::ios::binary);
    if (!fout) {
        put_err(err, err_len, std::string("cannot open ") + path);
        return DNDSMI_ERR_IO;
    }
    const int64_t nn = m->n_node, nc = m->n_cell;
    auto w_int = [&](int32_t v) { fout.write(reinterpret_cast<const char *>(&v), sizeof v); };
    auto w_float = [&](float v) { fout.write(reinterpret_cast<const char *>(&v), sizeof v); };
    auto w_double = [&](double v) { fout.write(reinterpret_cast<const char *>(&v), sizeof v); };
    auto w_string = [&](const std::string &s) { // one 32-bit word per character, then 0
        for (char ch : s)
            w_int(int32_t(ch));
        w_int(0);
    };
    // ---- header section
    fout.write("#!TDV112", 8);
    w_int(1);            // byte order
    w_int(0);            // file type: grid and solution
    w_string("TitleHahaha"); // the reference's title, kept so the files compare equal
    w_int(n_data + 4);   // X, Y, Z, the data, iPart
    w_string("X"), w_string("Y"), w_string("Z");
    for (int32_t k = 0; k < n_data; k++)
        w_string(names[k] ? names[k] : "");
    w_string("iPart");
    w_float(299.0f);     // zone marker
    w_string("zone_0");
    w_int(-1);           // parent zone
    w_int(-1);           // strand id
    w_double(t);         // solution time
    w_int(-1);           // zone colour
    w_int(3);            // FEQUADRILATERAL (triangles repeat their last node)
    w_int(1);            // variable locations specified
    for (int k = 0; k < 3; k++)
        w_int(0);        // coordinates at the nodes
    for (int32_t k = 0; k < n_data; k++)
        w_int(1);        // data at the cell centres
    w_int(1);            // iPart at the cell centres
    w_int(0);            // no raw face neighbours
    w_int(0);            // no user-defined face connections
    w_int(int32_t(nn));
    w_int(int32_t(nc));
    w_int(0), w_int(0), w_int(0); // I, J, K cell dims
    w_int(0);            // no auxiliary data
    w_float(357.0f);     // end of header
    // ---- data section
    w_float(299.0f);
    for (int32_t k = 0; k < n_data + 4; k++)
        w_int(2);        // every variable as double
    w_int(0);            // no passive variables
    w_int(0);            // no variable sharing
    w_int(-1);           // no connectivity sharing
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int64_t i = 0; i < nn; i++)
        for (int d = 0; d < 2; d++) {
            lo[d] = std::min(lo[d], m->node_xy[size_t(2 * i + d)]);
            hi[d] = std::max(hi[d], m->node_xy[size_t(2 * i + d)]);
        }
    w_double(lo[0]), w_double(hi[0]), w_double(lo[1]), w_double(hi[1]);
    w_double(0.0), w_double(0.0); // z
    for (int32_t k = 0; k < n_data; k++) {
        double a = INFINITY, b = -INFINITY;
        for (int64_t i = 0; i < nc; i++) {
            a = std::min(a, data[size_t(k) * size_t(nc) + size_t(i)]);
            b = std::max(b, data[size_t(k) * size_t(nc) + size_t(i)]);
        }
        w_double(a), w_double(b);
    }
    w_double(0.0), w_double(1.0); // iPart: ranks 0 .. number of ranks
    for (int d = 0; d < 2; d++)
        for (int64_t i = 0; i < nn; i++)
            w_double(m->node_xy[size_t(2 * i + d)]);
    for (int64_t i = 0; i < nn; i++)
        w_double(0.0);
    for (int32_t k = 0; k < n_data; k++)
        fout.write(reinterpret_cast<const char *>(data + size_t(k) * size_t(nc)), std::streamsize(sizeof(double) * nc));
    for (int64_t i = 0; i < nc; i++)
        w_double(0.0);   // one rank
    for (int64_t i = 0; i < nc; i++) {
        const int64_t b0 = m->cell_ptr[size_t(i)], nv = m->cell_ptr[size_t(i + 1)] - b0;
        for (int q = 0; q < 4; q++)
            w_int(int32_t(m->cell2node[size_t(b0 + (q < nv ? q : nv - 1))])); // zero-based
    }
    fout.close();
    if (!fout) {
        put_err(err, err_len, std::string("write failed: ") + path);
        return DNDSMI_ERR_IO;
    }
    return DNDSMI_OK;
}
