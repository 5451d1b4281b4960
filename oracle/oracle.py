"""ctypes front end of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker or the baseline, never
as the product.  Build with `make -C oracle` (or `oracle.build()`).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the CPU-baseline copy: the same source compiled for speed on the machine
# that runs it (-O3 -march=native, contraction to FMA allowed); built where it
# is timed, since -march=native code need not run on another host
BASELINE_LIB = os.path.join(HERE, "_baseline", "liboracle_baseline.so")
BASELINE_FLAGS = ["-O3", "-march=native", "-ffp-contract=fast", "-std=c++17", "-fPIC", "-shared", "-pthread"]

PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)


class OMesh(C.Structure):
    _fields_ = [("n_cell", C.c_int64), ("n_face", C.c_int64), ("cell_ptr", PI64), ("cell2face", PI64),
                ("face2cell", PI64), ("face_bc", PI32)]


class OSetup(C.Structure):
    _fields_ = [("volume", PD), ("face_area", PD), ("face_center", PD), ("face_norm_center", PD),
                ("face_norm", PD), ("face_weight", PD), ("face_d0", PD), ("face_dibj_g0", PD),
                ("matrix_secondary", PD), ("ainv", PD), ("ainv_b", PD), ("ainv_bvec", PD), ("face_gauss_xy", PD),
                ("cell_scale", PD)]


class OSetting(C.Structure):
    _fields_ = [("riemann_solver", C.c_int32), ("gamma", C.c_double), ("far_field", C.c_double * 4),
                ("jacobi_relax", C.c_double), ("tang_weight", C.c_double), ("tang_weight_direction", C.c_int32),
                ("wbap_smooth_indicator_scale", C.c_double), ("wbap_n_std", C.c_double),
                ("gmres_code", C.c_int32), ("n_gmres_space", C.c_int32), ("n_gmres_iter", C.c_int32),
                ("limiter_procedure", C.c_int32), ("mu_gas", C.c_double), ("pr_gas", C.c_double),
                ("r_gas", C.c_double), ("t_ref", C.c_double), ("c_sutherland", C.c_double),
                ("sor_scan", C.POINTER(C.c_int64)), ("sor_inverse", C.c_int32)]


def build(quiet: bool = True) -> str:
    """Compile liboracle.so from oracle.cpp (g++, -O2 -ffp-contract=off)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    return LIB


def build_baseline() -> str:
    """Compile the CPU-baseline copy (BASELINE_FLAGS) into oracle/_baseline/
    on this host; it is timed by the benchmark, never used as the checker."""
    os.makedirs(os.path.dirname(BASELINE_LIB), exist_ok=True)
    tmp = BASELINE_LIB + ".tmp"
    out = subprocess.run(["g++", *BASELINE_FLAGS, os.path.join(HERE, "oracle.cpp"), "-o", tmp], capture_output=True,
                         text=True)
    if out.returncode != 0:
        raise RuntimeError("baseline oracle build failed:\n" + out.stdout + out.stderr)
    os.replace(tmp, BASELINE_LIB)
    return BASELINE_LIB


_libs = {}


def lib(variant: str = "parity") -> C.CDLL:
    """The parity build (liboracle.so, -O2 -ffp-contract=off: every product
    and sum rounded as written) or, with variant="baseline", the speed build
    of the CPU baseline (build_baseline first)."""
    if variant not in _libs:
        if variant == "parity":
            if not os.path.exists(LIB):
                build()
            path = LIB
        elif variant == "baseline":
            if not os.path.exists(BASELINE_LIB):
                raise RuntimeError("the baseline oracle is not built: call oracle.build_baseline() first")
            path = BASELINE_LIB
        else:
            raise ValueError(variant)
        L = C.CDLL(path)
        sig = {
            "oracle_jacobi_step": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, PD, PD, C.c_double],
            "oracle_sor_step": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, PD, PI64, C.c_double],
            "oracle_evaluate_dt": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, C.c_double,
                                   C.c_double, C.c_int32, PD, PD, PD, PD],
            "oracle_evaluate_rhs": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, PD, PD, C.c_double,
                                    PD],
            "oracle_ssprk3_steps": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), C.c_int32, C.c_int32,
                                    C.c_int32, C.c_double, C.c_double, C.c_int32, PD, PD, PD],
            "oracle_wbap_v3": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, PD, PD, PD, C.c_int32],
            "oracle_wbap_v2": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, PD, PD, PD, C.c_int32],
            "oracle_roe_flux": [PD, PD, C.c_double, C.c_double, PD],
            "oracle_hllc_flux": [PD, PD, C.c_double, C.c_double, PD],
            "oracle_hllep_flux": [PD, PD, C.c_double, C.c_double, PD],
            "oracle_roe_m_flux": [PD, PD, C.c_double, C.c_double, C.c_int, PD],
            "oracle_eigenvectors": [PD, C.c_double, PD, PD],
            "oracle_boundary_value": [C.POINTER(OSetting), PD, PD, PD, C.c_double, C.c_int32, C.c_int32, PD],
            "oracle_flux_jacobian_times_du": [PD, PD, PD, C.c_double, C.c_double, PD],
            "oracle_set_threads": [C.c_int32],
            "oracle_lusgs_matrix_init": [C.POINTER(OMesh), C.POINTER(OSetup), PD, C.c_double, C.c_double, PD, PD],
            "oracle_lusgs_forward": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PI64, PD, PD,
                                     C.c_double, PD, PD, PD, PD],
            "oracle_lusgs_backward": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PI64, PD, PD,
                                      C.c_double, PD, PD, PD, PD],
            "oracle_implicit_euler": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PI64, C.c_int32,
                                      C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double, PD, PD,
                                      PD, C.POINTER(C.c_int32)],
            "oracle_bdf_steps": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PI64, C.c_int32, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double, PD,
                                 PD, C.POINTER(C.c_int32)],
            "oracle_lusgs_matrix_vec": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PD, PD, C.c_double,
                                        PD, PD, PD],
            "oracle_fsolve": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PI64, PD, C.c_double,
                              C.c_double, PD, PD, PD, PD],
            "oracle_sdirk4_step": [C.POINTER(OMesh), C.POINTER(OSetup), C.POINTER(OSetting), PI64, C.c_int32,
                                   C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double, PD, PD,
                                   C.POINTER(C.c_int32)],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        _libs[variant] = L
    return _libs[variant]


def _d(a):
    if a is None:
        return PD()
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], a.dtype
    return a.ctypes.data_as(PD)


class Oracle:
    """Holds contiguous copies of the mesh and setup arrays and calls the oracle.

    `mesh` and `setup` are dicts of numpy arrays with the keys of the
    product's mesh and VR views (see dnds_amd.host), so the same data feeds
    both the oracle and the device.
    """

    def __init__(self, mesh: dict, setup: dict, riemann_solver: int = 1, gamma: float = 1.4,
                 far_field=(1.0, 0.0, 0.0, 2.5), jacobi_relax: float = 1.0, tang_weight: float = 1.0,
                 tang_weight_direction: int = 0, wbap_smooth_indicator_scale: float = 1e-10,
                 wbap_n_std: float = 10.0, gmres_code: int = 0, n_gmres_space: int = 10, n_gmres_iter: int = 2,
                 limiter_procedure: int = 0, mu_gas: float = 0.0, pr_gas: float = 0.72, r_gas: float = 1.0,
                 t_ref: float = 273.15, c_sutherland: float = 110.4, variant: str = "parity", sor_scan=None,
                 sor_inverse: bool = False):
        """sor_scan: a scan order (iScan -> iCell) makes the steps'
        reconstruction sweeps SOR_Instead sweeps in that order, forward then
        backward with sor_inverse (SOR_InverseScanning)."""
        self._keep = {}

        def keep(name, a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep[name] = a
            return a

        self.n_cell = int(mesh["n_cell"])
        self.n_face = int(mesh["n_face"])
        self.m = OMesh(self.n_cell, self.n_face,
                       keep("cell_ptr", mesh["cell_ptr"], np.int64).ctypes.data_as(PI64),
                       keep("cell2face", mesh["cell2face"], np.int64).ctypes.data_as(PI64),
                       keep("face2cell", mesh["face2cell"], np.int64).ctypes.data_as(PI64),
                       keep("face_bc", mesh["face_bc"], np.int32).ctypes.data_as(PI32))
        fields = {}
        for name, _ in OSetup._fields_:
            a = setup.get(name)
            fields[name] = PD() if a is None or a.size == 0 else keep(name, a, np.float64).ctypes.data_as(PD)
        self.s = OSetup(**fields)
        ff = (C.c_double * 4)(*far_field)
        self.cfg = OSetting(riemann_solver, gamma, ff, jacobi_relax, tang_weight, tang_weight_direction,
                            wbap_smooth_indicator_scale, wbap_n_std, gmres_code, n_gmres_space, n_gmres_iter,
                            limiter_procedure, mu_gas, pr_gas, r_gas, t_ref, c_sutherland,
                            PI64() if sor_scan is None else keep("sor_scan", sor_scan, np.int64).ctypes.data_as(PI64),
                            int(sor_inverse))
        if sor_scan is not None:
            assert np.array_equal(np.sort(self._keep["sor_scan"]), np.arange(self.n_cell))
        self.lib = lib(variant)

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed with code {rc}")

    def jacobi_step(self, u, urec, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        urec = np.array(urec, dtype=np.float64, order="C", copy=True)
        new = np.empty_like(urec)
        self._check(self.lib.oracle_jacobi_step(C.byref(self.m), C.byref(self.s), C.byref(self.cfg), _d(u),
                                                _d(urec), _d(new), t), "jacobi_step")
        return urec

    def sor_step(self, u, urec, scan=None, t=0.0):
        """one in-place SOR sweep (SOR_Instead) in the cell order `scan`
        (default: natural order); returns the new reconstruction"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        urec = np.array(urec, dtype=np.float64, order="C", copy=True)
        scan = np.ascontiguousarray(np.arange(self.n_cell) if scan is None else scan, dtype=np.int64)
        assert scan.shape == (self.n_cell,)
        self._check(self.lib.oracle_sor_step(C.byref(self.m), C.byref(self.s), C.byref(self.cfg), _d(u), _d(urec),
                                             scan.ctypes.data_as(PI64), t), "sor_step")
        return urec

    def evaluate_dt(self, u, cfl, max_dt=1e100, use_local_dt=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        dt = np.empty(self.n_cell)
        lf = np.empty(self.n_face)
        dlf = np.empty(self.n_face)
        dtmin = C.c_double()
        self._check(self.lib.oracle_evaluate_dt(C.byref(self.m), C.byref(self.s), C.byref(self.cfg), _d(u), cfl,
                                                max_dt, int(use_local_dt), _d(dt), C.byref(dtmin), _d(lf),
                                                _d(dlf)), "evaluate_dt")
        return dt, dtmin.value, lf, dlf

    def evaluate_rhs(self, u, urec, delta_lambda_face, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        urec = np.ascontiguousarray(urec, dtype=np.float64)
        dlf = np.ascontiguousarray(delta_lambda_face, dtype=np.float64)
        rhs = np.empty((self.n_cell, 4))
        self._check(self.lib.oracle_evaluate_rhs(C.byref(self.m), C.byref(self.s), C.byref(self.cfg), _d(dlf),
                                                 _d(u), _d(urec), t, _d(rhs)), "evaluate_rhs")
        return rhs

    def _wbap(self, fn, u, urec, if_all):
        u = np.ascontiguousarray(u, dtype=np.float64)
        urec = np.ascontiguousarray(urec, dtype=np.float64)
        out = np.empty_like(urec)
        ind = np.empty(self.n_cell)
        self._check(getattr(self.lib, "oracle_" + fn)(C.byref(self.m), C.byref(self.s), C.byref(self.cfg), _d(u),
                                                      _d(urec), _d(out), _d(ind), int(if_all)), fn)
        return out, ind

    def wbap_v3(self, u, urec, if_all=False):
        """Returns (limited reconstruction, smoothness indicator)."""
        return self._wbap("wbap_v3", u, urec, if_all)

    def wbap_v2(self, u, urec, if_all=False):
        """Returns (limited reconstruction, smoothness indicator)."""
        return self._wbap("wbap_v2", u, urec, if_all)

    def lusgs_matrix_init(self, dtau, dt, alpha, lambda_face):
        jinv = np.empty(self.n_cell)
        self._check(self.lib.oracle_lusgs_matrix_init(C.byref(self.m), C.byref(self.s), _d(np.ascontiguousarray(dtau,
                    float)), dt, alpha, _d(np.ascontiguousarray(lambda_face, float)), _d(jinv)), "lusgs init")
        return jinv

    def lusgs_sweep(self, forward, jinv, lambda_face, alpha, rhs, u, uinc, scan2cell=None, in_place=True):
        """One LU-SGS half sweep; returns uIncNew (a copy of uinc updated in
        place when in_place, else a separate array as in the reference's
        call with distinct buffers)."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        src = np.array(uinc, dtype=np.float64, order="C", copy=True)
        dst = src if in_place else np.array(uinc, dtype=np.float64, order="C", copy=True)
        order = None if scan2cell is None else np.ascontiguousarray(scan2cell, dtype=np.int64)
        fn = self.lib.oracle_lusgs_forward if forward else self.lib.oracle_lusgs_backward
        self._check(fn(C.byref(self.m), C.byref(self.s), C.byref(self.cfg),
                       order.ctypes.data_as(PI64) if order is not None else None, _d(np.ascontiguousarray(jinv, float)),
                       _d(np.ascontiguousarray(lambda_face, float)), alpha, _d(rhs), _d(u), _d(src), _d(dst)),
                    "lusgs sweep")
        return dst

    def implicit_euler(self, u, urec, max_iter, n_sweeps, cfl, dt=1e100, use_local_dt=True, threshold=1e-10,
                       t=0.0, scan2cell=None):
        """Returns (u, urec, residual history [iters][4])."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        urec = np.array(urec, dtype=np.float64, order="C", copy=True)
        hist = np.zeros((max_iter, 4))
        done = C.c_int32()
        order = None if scan2cell is None else np.ascontiguousarray(scan2cell, dtype=np.int64)
        self._check(self.lib.oracle_implicit_euler(C.byref(self.m), C.byref(self.s), C.byref(self.cfg),
                                                   order.ctypes.data_as(PI64) if order is not None else None,
                                                   max_iter, n_sweeps, cfl, dt, int(use_local_dt), threshold, t,
                                                   _d(u), _d(urec), _d(hist), C.byref(done)), "implicit euler")
        return u, urec, hist[:done.value]

    def bdf_steps(self, u, urec, k, n_steps, max_iter, n_sweeps, cfl, dt, use_local_dt=True, threshold=1e-10,
                  t=0.0, scan2cell=None):
        """n_steps BDF steps of order k (DNDS_ODE.hpp:392-512) from a fresh
        integrator; returns (u, urec, inner iterations per step)."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        urec = np.array(urec, dtype=np.float64, order="C", copy=True)
        its = np.zeros(n_steps, dtype=np.int32)
        order = None if scan2cell is None else np.ascontiguousarray(scan2cell, dtype=np.int64)
        self._check(self.lib.oracle_bdf_steps(C.byref(self.m), C.byref(self.s), C.byref(self.cfg),
                                              order.ctypes.data_as(PI64) if order is not None else None, k, n_steps,
                                              max_iter, n_sweeps, cfl, dt, int(use_local_dt), threshold, t, _d(u),
                                              _d(urec), its.ctypes.data_as(C.POINTER(C.c_int32))), "bdf steps")
        return u, urec, its

    def lusgs_matrix_vec(self, diag, lambda_face, alpha, u, x):
        """LUSGSMatrixVec: the implicit operator times x (diag = D per cell)."""
        out = np.empty((self.n_cell, 4))
        self._check(self.lib.oracle_lusgs_matrix_vec(C.byref(self.m), C.byref(self.s), C.byref(self.cfg),
                                                     _d(np.ascontiguousarray(diag, float)),
                                                     _d(np.ascontiguousarray(lambda_face, float)), alpha,
                                                     _d(np.ascontiguousarray(u, float)),
                                                     _d(np.ascontiguousarray(x, float)), _d(out)), "matrix vec")
        return out

    def fsolve(self, dtau, dt, alpha, lambda_face, u, rhs, scan2cell=None):
        """The solver's fsolve for this oracle's gmres_code: returns
        (xinc, the compressed rhs)."""
        rhs = np.array(rhs, dtype=np.float64, order="C", copy=True)
        xinc = np.empty((self.n_cell, 4))
        order = None if scan2cell is None else np.ascontiguousarray(scan2cell, dtype=np.int64)
        self._check(self.lib.oracle_fsolve(C.byref(self.m), C.byref(self.s), C.byref(self.cfg),
                                           order.ctypes.data_as(PI64) if order is not None else None,
                                           _d(np.ascontiguousarray(dtau, float)), dt, alpha,
                                           _d(np.ascontiguousarray(lambda_face, float)),
                                           _d(np.ascontiguousarray(u, float)), _d(rhs), _d(xinc)), "fsolve")
        return xinc, rhs

    def sdirk4_step(self, u, urec, max_iter, n_sweeps, cfl, dt, use_local_dt=True, threshold=1e-10, t=0.0,
                    scan2cell=None):
        """One SDIRK4 dual-time step (DNDS_ODE.hpp:270-372); returns (u, urec,
        inner iterations of the 3 stages)."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        urec = np.array(urec, dtype=np.float64, order="C", copy=True)
        its = np.zeros(3, dtype=np.int32)
        order = None if scan2cell is None else np.ascontiguousarray(scan2cell, dtype=np.int64)
        self._check(self.lib.oracle_sdirk4_step(C.byref(self.m), C.byref(self.s), C.byref(self.cfg),
                                                order.ctypes.data_as(PI64) if order is not None else None, max_iter,
                                                n_sweeps, cfl, dt, int(use_local_dt), threshold, t, _d(u), _d(urec),
                                                its.ctypes.data_as(C.POINTER(C.c_int32))), "sdirk4 step")
        return u, urec, its

    def ssprk3_steps(self, u, urec, n_steps, n_sweeps, cfl, t=0.0, max_dt=1e100, use_local_dt=False,
                     use_limiter=False):
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        urec = np.array(urec, dtype=np.float64, order="C", copy=True)
        tt = C.c_double(t)
        self._check(self.lib.oracle_ssprk3_steps(C.byref(self.m), C.byref(self.s), C.byref(self.cfg), n_steps,
                                                 n_sweeps, int(use_limiter), cfl, max_dt, int(use_local_dt), _d(u),
                                                 _d(urec), C.byref(tt)), "ssprk3_steps")
        return u, urec, tt.value


def roe_flux(UL, UR, gamma=1.4, dlambda=0.0):
    F = np.empty(4)
    rc = lib().oracle_roe_flux(_d(np.ascontiguousarray(UL, float)), _d(np.ascontiguousarray(UR, float)), gamma,
                               dlambda, _d(F))
    if rc:
        raise RuntimeError(f"roe flux failed {rc}")
    return F


def hllc_flux(UL, UR, gamma=1.4, dlambda=0.0):
    F = np.empty(4)
    rc = lib().oracle_hllc_flux(_d(np.ascontiguousarray(UL, float)), _d(np.ascontiguousarray(UR, float)), gamma,
                                dlambda, _d(F))
    if rc:
        raise RuntimeError(f"hllc flux failed {rc}")
    return F


def roe_m_flux(UL, UR, scheme, gamma=1.4, dlambda=0.0):
    """Roe flux with eigenvalue scheme `scheme` (0 Harten-Yee, 1..5 Roe_M1..M5)"""
    F = np.empty(4)
    rc = lib().oracle_roe_m_flux(_d(np.ascontiguousarray(UL, float)), _d(np.ascontiguousarray(UR, float)), gamma,
                                 dlambda, scheme, _d(F))
    if rc:
        raise RuntimeError(f"roe_m flux failed {rc}")
    return F


def hllep_flux(UL, UR, gamma=1.4, dlambda=0.0):
    F = np.empty(4)
    rc = lib().oracle_hllep_flux(_d(np.ascontiguousarray(UL, float)), _d(np.ascontiguousarray(UR, float)), gamma,
                                dlambda, _d(F))
    if rc:
        raise RuntimeError(f"hllep flux failed {rc}")
    return F


def eigenvectors(U, gamma=1.4):
    L = np.empty((4, 4))
    R = np.empty((4, 4))
    rc = lib().oracle_eigenvectors(_d(np.ascontiguousarray(U, float)), gamma, _d(L), _d(R))
    if rc:
        raise RuntimeError(f"eigenvectors failed {rc}")
    return L, R


def boundary_value(UL, unit_norm, point, t, bc, far_field=(1.0, 0.0, 0.0, 2.5), gamma=1.4):
    """generateBoundaryValue for one Gauss point; raises on a non-OK status."""
    cfg = OSetting(riemann_solver=1, gamma=gamma, far_field=(C.c_double * 4)(*far_field), jacobi_relax=1.0,
                   tang_weight=1.0, tang_weight_direction=0, wbap_smooth_indicator_scale=1e100, wbap_n_std=10.0)
    UL = np.array(UL, dtype=np.float64)
    UR = np.empty(4)
    rc = lib().oracle_boundary_value(C.byref(cfg), _d(UL), _d(np.ascontiguousarray(unit_norm, float)),
                                     _d(np.ascontiguousarray(point, float)), t, bc, 0, _d(UR))
    if rc:
        raise RuntimeError(f"boundary value failed {rc}")
    return UR


def flux_jacobian_times_du(U, unit_norm, dU, lambda_main=0.0, gamma=1.4):
    dF = np.empty(4)
    lib().oracle_flux_jacobian_times_du(_d(np.ascontiguousarray(U, float)), _d(np.ascontiguousarray(unit_norm, float)),
                                        _d(np.ascontiguousarray(dU, float)), lambda_main, gamma, _d(dF))
    return dF


def set_threads(n: int, variant: str = "parity"):
    """Worker threads of the oracle's per-cell / per-face loops."""
    lib(variant).oracle_set_threads(int(n))
