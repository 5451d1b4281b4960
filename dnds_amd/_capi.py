"""ctypes declarations of the C ABI in include/dndsmi.h.

The structures mirror the C structs field by field; `load()` opens the
in-tree shared library built by `dnds_amd.build` and declares every entry
point's signature.  Nothing here falls back to another implementation: a
missing library is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdndsmi.so")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_NAN = 3
ERR_NONPOSITIVE_STATE = 4
ERR_UNSUPPORTED = 5
ERR_OUT_OF_MEMORY = 6
ERR_SETUP = 7
ERR_IO = 8

BC_INNER = -1
BC_WALL = 1
BC_FARFIELD = 2
BC_WALL_EULER = 3
BC_WALL_NOSLIP = 4
BC_SPECIAL_DMRFAR = -2
BC_SPECIAL_RTFAR = -3
BC_SPECIAL_IVFAR = -4
BC_SPECIAL_2DRIEMANNFAR = -5

NO_CELL = -(2 ** 63)

MESH_TRI_SPLIT = 0
MESH_QUAD = 1
MESH_MIXED = 2

RS_ROE = 1
RS_HLLC = 2
RS_HLLEP = 3
RS_ROE_M = {1: 11, 2: 12, 3: 13, 4: 14, 5: 15}

ARRAY_DOF = 0
ARRAY_REC = 1
ARRAY_SCALAR = 2

PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)
PF = C.POINTER(C.c_float)


class MeshView(C.Structure):
    _fields_ = [
        ("n_node", C.c_int64), ("n_cell", C.c_int64), ("n_face", C.c_int64),
        ("node_xy", PD), ("cell_ptr", PI64), ("cell2node", PI64), ("cell2face", PI64),
        ("face2node", PI64), ("face2cell", PI64), ("face2cell_ref", PI32), ("face_bc", PI32),
        ("face_shift", PD),
    ]


class MeshSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("periodic_x", C.c_int32), ("periodic_y", C.c_int32),
        ("nx", C.c_int64), ("ny", C.c_int64),
        ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
        ("jitter", C.c_double), ("y_ratio", C.c_double), ("seed", C.c_uint64),
        ("bc_xmin", C.c_int32), ("bc_xmax", C.c_int32), ("bc_ymin", C.c_int32), ("bc_ymax", C.c_int32),
    ]


class VRSetting(C.Structure):
    _fields_ = [
        ("rec_order", C.c_int32), ("jacobi_relax", C.c_double),
        ("wall_weight", C.c_double), ("far_weight", C.c_double), ("tang_weight", C.c_double),
        ("tang_weight_mod_min", C.c_double), ("tang_weight_direction", C.c_int32),
        ("anisotropic_lengths", C.c_int32), ("weight_scheme_geom", C.c_int32),
        ("weight_scheme_dir", C.c_int32), ("weight_len_scale", C.c_double),
        ("wbap_smooth_indicator_scale", C.c_double), ("wbap_n_std", C.c_double),
        ("use_local_coord", C.c_int32), ("orthogonalize_base", C.c_int32), ("norm_wbap", C.c_int32),
        ("build_limiter_data", C.c_int32), ("sor_instead", C.c_int32), ("sor_inverse_scanning", C.c_int32),
    ]


class VRView(C.Structure):
    _fields_ = [
        ("n_cell", C.c_int64), ("n_face", C.c_int64), ("n_c2f", C.c_int64),
        ("volume", PD), ("face_area", PD), ("cell_bary", PD), ("base_moment", PD),
        ("face_center", PD), ("face_norm_center", PD), ("face_norm", PD), ("face_weight", PD),
        ("face_d0", PD), ("face_dibj_g0", PD), ("matrix_secondary", PD),
        ("ainv", PD), ("ainv_b", PD), ("ainv_bvec", PD), ("face_gauss_xy", PD), ("cell_scale", PD),
        ("cell_frame", PD),
    ]


class EulerSetting(C.Structure):
    _fields_ = [
        ("riemann_solver", C.c_int32), ("gamma", C.c_double), ("far_field", C.c_double * 4),
        ("mu_gas", C.c_double), ("pr_gas", C.c_double), ("r_gas", C.c_double), ("t_ref", C.c_double),
        ("c_sutherland", C.c_double),
    ]


# (name, restype, argtypes) of every entry point declared in include/dndsmi.h
SIGNATURES = [
    ("dndsmi_mesh_generate", C.c_int, [C.POINTER(MeshSpec), C.POINTER(C.c_void_p)]),
    ("dndsmi_mesh_from_cells", C.c_int, [C.c_int64, PD, C.c_int64, PI64, PI64, C.c_int64, PI64, PI32,
                                         C.POINTER(C.c_void_p)]),
    ("dndsmi_mesh_read_gmsh", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]),
    ("dndsmi_mesh_write_plt", C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_char_p), PD, C.c_double,
                                        C.c_char_p, C.c_size_t]),
    ("dndsmi_mesh_get_view", C.c_int, [C.c_void_p, C.POINTER(MeshView)]),
    ("dndsmi_mesh_view_check", C.c_int, [C.POINTER(MeshView), C.c_char_p, C.c_size_t]),
    ("dndsmi_mesh_destroy", None, [C.c_void_p]),
    ("dndsmi_vr_setting_default", None, [C.POINTER(VRSetting)]),
    ("dndsmi_vr_create", C.c_int, [C.POINTER(MeshView), C.POINTER(VRSetting), C.c_int32, C.POINTER(C.c_void_p)]),
    ("dndsmi_vr_get_view", C.c_int, [C.c_void_p, C.POINTER(VRView)]),
    ("dndsmi_vr_destroy", None, [C.c_void_p]),
    ("dndsmi_ctx_create", C.c_int, [C.POINTER(MeshView), C.POINTER(VRView), C.POINTER(VRSetting),
                                    C.POINTER(EulerSetting), C.c_void_p, C.POINTER(C.c_void_p)]),
    ("dndsmi_ctx_destroy", None, [C.c_void_p]),
    ("dndsmi_last_error", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("dndsmi_sync", C.c_int, [C.c_void_p]),
    ("dndsmi_ctx_device_bytes", C.c_int64, [C.c_void_p]),
    ("dndsmi_array_create", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("dndsmi_array_destroy", None, [C.c_void_p]),
    ("dndsmi_array_upload", C.c_int, [C.c_void_p, PD]),
    ("dndsmi_array_download", C.c_int, [C.c_void_p, PD]),
    ("dndsmi_array_copy", C.c_int, [C.c_void_p, C.c_void_p]),
    ("dndsmi_array_set_constant", C.c_int, [C.c_void_p, C.c_double]),
    ("dndsmi_reconstruction_jacobi_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    ("dndsmi_evaluate_dt", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, PD, C.c_double, C.c_int32]),
    ("dndsmi_evaluate_rhs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    ("dndsmi_wbap_limit_facial_v3", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_int32]),
    ("dndsmi_wbap_limit_facial_v2", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_int32]),
    ("dndsmi_set_limiter_procedure", C.c_int, [C.c_void_p, C.c_int32]),
    ("dndsmi_ssprk3_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                     C.c_double, C.c_double, C.c_int32, C.c_double, PD]),
    ("dndsmi_ssprk3_step_mixed", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                           C.c_double, C.c_double, C.c_int32, C.c_double, PD]),
    ("dndsmi_lusgs_colouring", C.c_int, [C.c_void_p, PI32, PI32]),
    ("dndsmi_warmstart_create", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int64,
                                          C.POINTER(C.c_void_p)]),
    ("dndsmi_warmstart_destroy", None, [C.c_void_p]),
    ("dndsmi_warmstart_num_params", C.c_int64, []),
    ("dndsmi_warmstart_get_params", C.c_int, [C.c_void_p, PF]),
    ("dndsmi_warmstart_set_params", C.c_int, [C.c_void_p, PF]),
    ("dndsmi_warmstart_features", C.c_int, [C.c_void_p, C.c_void_p, PF, PF]),
    ("dndsmi_warmstart_add_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("dndsmi_warmstart_num_samples", C.c_int64, [C.c_void_p]),
    ("dndsmi_warmstart_add_rows", C.c_int, [C.c_void_p, PF, PF, C.c_int64]),
    ("dndsmi_warmstart_targets", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, PF]),
    ("dndsmi_warmstart_loss_grad", C.c_int, [C.c_void_p, PF, PF, C.c_int32, PF, PD]),
    ("dndsmi_warmstart_adam_step", C.c_int, [C.c_void_p, PF, C.c_double]),
    ("dndsmi_warmstart_train", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double, PD]),
    ("dndsmi_warmstart_predict", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("dndsmi_ssprk3_set_warm_start", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    ("dndsmi_read_probe", C.c_int, [C.c_void_p, C.c_int32]),
    ("dndsmi_set_limiter_if_all", C.c_int, [C.c_void_p, C.c_int32]),
    ("dndsmi_ssprk3_warm_start_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), PD]),
    ("dndsmi_reconstruct_to_tolerance", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                                  C.c_int32, C.c_double, PI32, PD]),
    ("dndsmi_lusgs_matrix_init", C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_int32, C.c_double]),
    ("dndsmi_update_lusgs_forward", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p]),
    ("dndsmi_update_lusgs_backward", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p]),
    ("dndsmi_evaluate_residual", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, PD]),
    ("dndsmi_evaluate_residual_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, PD]),
    ("dndsmi_implicit_euler_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                             C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double, PI32, PD]),
    ("dndsmi_lusgs_matrix_vec", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("dndsmi_set_linear_solver", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    ("dndsmi_sdirk4_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                     C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double, PI32, PD]),
    ("dndsmi_bdf_create", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("dndsmi_bdf_destroy", None, [C.c_void_p]),
    ("dndsmi_bdf_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double,
                                  C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double, PI32, PD]),
    ("dndsmi_profile_enable", C.c_int, [C.c_void_p, C.c_int32]),
    ("dndsmi_profile_query", C.c_int, [C.c_void_p, C.c_int32, PI64, PD]),
    ("dndsmi_kernel_bytes", C.c_int64, [C.c_void_p, C.c_int32]),
]

K_JACOBI = 0
K_WALL_CORR = 1
K_DT_FACE = 2
K_DT_CELL = 3
K_RHS_FACE = 4
K_RHS_CELL = 5
K_WBAP_FACE = 6
K_WBAP_CELL = 7
K_WBAP_LIM = 8
KERNEL_NAMES = ["k_jacobi", "k_wall_corr", "k_dt_face", "k_dt_cell", "k_rhs_face", "k_rhs_cell", "k_wbap_face",
                "k_wbap_cell", "k_wbap_lim", "k_lusgs_sweep", "k_delta32", "k_rhs_face32", "k_rhs_cell32",
                "k_lusgs_matvec", "k_rhs_bface", "k_read_probe"]

_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Open the in-tree library and declare all entry points (raises if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"dnds_amd: native library {p} is not built; run dnds_amd.build.build()")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def as_array(ptr, n, dtype=np.float64):
    """Copy n elements behind a ctypes pointer into a new numpy array."""
    if not ptr or n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PD)


def i64ptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PI64)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PF)


def i32ptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(PI32)
