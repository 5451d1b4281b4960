// Internal host-side types shared by the mesh builder, the VR setup and the
// device context.  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dndsmi.h"

struct dndsmi_mesh {
    int64_t n_node = 0, n_cell = 0, n_face = 0;
    std::vector<double> node_xy;      // [n_node][2]
    std::vector<int64_t> cell_ptr;    // [n_cell+1]
    std::vector<int64_t> cell2node;
    std::vector<int64_t> cell2face;
    std::vector<int64_t> face2node;   // [n_face][2]
    std::vector<int64_t> face2cell;   // [n_face][2]
    std::vector<int32_t> face2cell_ref;
    std::vector<int32_t> face_bc;
    std::vector<double> face_shift;   // [n_face][2]
};

// dndsmi_mesh_from_cells with the failure's description in err.
int mesh_from_cells(int64_t n_node, const double *node_xy, int64_t n_cell, const int64_t *cell_ptr,
                    const int64_t *cell2node, int64_t n_bnd, const int64_t *bnd2node, const int32_t *bnd_bc,
                    dndsmi_mesh **out, std::string &err);

// dndsmi_mesh_view_check with the failure's description in err: every
// index in range, cell2face and face2cell / face2cell_ref naming each other,
// triangles and quadrilaterals only, a known kind on every boundary face.
int check_mesh_view(const dndsmi_mesh_view *m, std::string &err);

struct dndsmi_vr {
    int64_t n_cell = 0, n_face = 0, n_c2f = 0;
    std::vector<double> volume, face_area, cell_bary, base_moment;
    std::vector<double> face_center, face_norm_center, face_norm, face_weight;
    std::vector<double> face_d0, face_dibj_g0, matrix_secondary;
    std::vector<double> face_gauss_xy, cell_scale, cell_frame;
    std::vector<double> ainv, ainv_b, ainv_bvec;
};
