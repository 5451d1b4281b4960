"""Host-side mirror of the reference's hot-path interfaces over the C ABI.

Names, argument order and meaning follow the reference:

* `CompactFacedMeshSerialRW`  -- DNDS_Mesh.hpp:634-709 (topology and its
  conventions; built by the seeded generator or from raw cells).
* `VRFiniteVolume2D`          -- DNDS_FV_VR.hpp:125-1713: `Initialization`
  (DNDS_FV_VR.cpp:3-13), `BuildRec`, `ReconstructionJacobiStep`.
* `EulerEvaluator`            -- DNDS_FV_EulerEvaluator.hpp:71-1462: `BuildMean`,
  `EvaluateDt`, `EvaluateRHS`.
* `ExplicitSSPRK3LocalDt`     -- DNDS_ODE.hpp:95-179: `Step`.

Errors: the reference aborts through DNDS_assert (DNDS_Defines.h:31-44); here
every failing call raises `DNDSError` carrying the ABI status code and the
library's message.  Device arrays (`ArrayDOFV`, `ArrayRecV`, `ArrayScalar`)
live on the GPU; `upload` / `download` convert to and from the reference's
host layouts.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _capi as A

# vfvSetting of the reference's 2D Riemann configuration
# (data/euler_config_2DRiemann.json: anisotropicLengths, tangWeight 1 along
# the face normal, S geometric weights, weightLenScale 0.5, farWeight 0,
# wallWeight 0.5, WBAP_SmoothIndicatorScale 0.1, global coordinates).  That
# file also sets SOR_Instead; the preset keeps the Jacobi sweep that the
# workloads here are defined with, and sor_instead=1 adds the in-place sweep.
RIEMANN_VR_SETTING = dict(anisotropic_lengths=1, tang_weight=1.0, tang_weight_direction=0, wall_weight=0.5,
                          far_weight=0.0, weight_len_scale=0.5, weight_scheme_geom=2, weight_scheme_dir=0,
                          wbap_smooth_indicator_scale=0.1, jacobi_relax=1.0)


class DNDSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dndsmi error {code}: {msg}")
        self.code = code


def _lib():
    return A.load()


def _check(rc: int, what: str, ctx=None):
    if rc != A.OK:
        msg = what
        if ctx:
            buf = C.create_string_buffer(512)
            _lib().dndsmi_last_error(ctx, buf, 512)
            if buf.value:
                msg = f"{what}: {buf.value.decode()}"
        raise DNDSError(rc, msg)


class CompactFacedMeshSerialRW:
    """A 2D mesh with the reference's topology conventions (owns the C handle)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self.view = A.MeshView()
        _check(_lib().dndsmi_mesh_get_view(self._h, C.byref(self.view)), "mesh view")
        self._arrays = None

    @classmethod
    def generate(cls, kind: int, nx: int, ny: int, x0=0.0, x1=1.0, y0=0.0, y1=1.0, periodic=(False, False),
                 jitter=0.0, y_ratio=1.0, seed=0, bc=(A.BC_FARFIELD,) * 4):
        """Seeded structured-as-unstructured mesh; bc = (xmin, xmax, ymin, ymax)."""
        spec = A.MeshSpec(kind=kind, periodic_x=int(periodic[0]), periodic_y=int(periodic[1]), nx=nx, ny=ny,
                          x0=x0, x1=x1, y0=y0, y1=y1, jitter=jitter, y_ratio=y_ratio, seed=seed,
                          bc_xmin=bc[0], bc_xmax=bc[1], bc_ymin=bc[2], bc_ymax=bc[3])
        h = C.c_void_p()
        _check(_lib().dndsmi_mesh_generate(C.byref(spec), C.byref(h)), "mesh generate")
        return cls(h)

    @classmethod
    def from_cells(cls, node_xy, cell_ptr, cell2node, bnd2node, bnd_bc):
        """Topology from raw cells and boundary edges (the Gmsh-reader path)."""
        node_xy = np.ascontiguousarray(node_xy, dtype=np.float64)
        cell_ptr = np.ascontiguousarray(cell_ptr, dtype=np.int64)
        cell2node = np.ascontiguousarray(cell2node, dtype=np.int64)
        bnd2node = np.ascontiguousarray(bnd2node, dtype=np.int64).reshape(-1)
        bnd_bc = np.ascontiguousarray(bnd_bc, dtype=np.int32)
        h = C.c_void_p()
        _check(_lib().dndsmi_mesh_from_cells(node_xy.shape[0], A.dptr(node_xy), cell_ptr.shape[0] - 1,
                                             A.i64ptr(cell_ptr), A.i64ptr(cell2node), bnd_bc.shape[0],
                                             A.i64ptr(bnd2node), A.i32ptr(bnd_bc), C.byref(h)), "mesh from cells")
        return cls(h)

    @classmethod
    def read_gmsh(cls, fname: str):
        """A Gmsh 2.2 ASCII mesh as CompactFacedMeshSerialRWBuild reads it
        (GmshReader::FileRead + InterpolateTopology, DNDS_Mesh.hpp:87-350,
        760-835); boundary groups named bc-1..bc-4, bc-DMRFar, bc-RTFar,
        bc-IVFar, bc-2DRiemannFar."""
        h = C.c_void_p()
        buf = C.create_string_buffer(1024)
        rc = _lib().dndsmi_mesh_read_gmsh(os.fsencode(fname), C.byref(h), buf, 1024)
        if rc != A.OK:
            raise DNDSError(rc, f"Gmsh read: {buf.value.decode(errors='replace')}")
        return cls(h)

    def PrintSerialPartPltBinaryDataArray(self, fname: str, names, data, t: float = 0.0):
        """Tecplot binary file of the mesh with cell-centred data
        (DNDS_Mesh.hpp:1634-1814): names[k] and data[k][n_cell]."""
        names = list(names)
        data = np.ascontiguousarray(data, dtype=np.float64).reshape(len(names), self.n_cell) if names else \
            np.zeros((0, self.n_cell))
        enc = [os.fsencode(n) for n in names]
        arr = (C.c_char_p * max(len(enc), 1))(*enc)
        buf = C.create_string_buffer(1024)
        rc = _lib().dndsmi_mesh_write_plt(self._h, os.fsencode(fname), len(names), arr, A.dptr(data), t, buf, 1024)
        if rc != A.OK:
            raise DNDSError(rc, f"Tecplot write: {buf.value.decode(errors='replace')}")

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().dndsmi_mesh_destroy(self._h)
            self._h = None

    @property
    def n_cell(self):
        return int(self.view.n_cell)

    @property
    def n_face(self):
        return int(self.view.n_face)

    def arrays(self) -> dict:
        """numpy copies of the topology (keys = dndsmi_mesh_view fields)."""
        if self._arrays is None:
            v = self.view
            nc, nf = int(v.n_cell), int(v.n_face)
            cp = A.as_array(v.cell_ptr, nc + 1, np.int64)
            self._arrays = dict(
                n_node=int(v.n_node), n_cell=nc, n_face=nf,
                node_xy=A.as_array(v.node_xy, 2 * v.n_node).reshape(-1, 2),
                cell_ptr=cp,
                cell2node=A.as_array(v.cell2node, cp[-1], np.int64),
                cell2face=A.as_array(v.cell2face, cp[-1], np.int64),
                face2node=A.as_array(v.face2node, 2 * nf, np.int64).reshape(-1, 2),
                face2cell=A.as_array(v.face2cell, 2 * nf, np.int64).reshape(-1, 2),
                face2cell_ref=A.as_array(v.face2cell_ref, 2 * nf, np.int32).reshape(-1, 2),
                face_bc=A.as_array(v.face_bc, nf, np.int32),
                face_shift=A.as_array(v.face_shift, 2 * nf).reshape(-1, 2),
            )
        return self._arrays


def vr_setting(**overrides) -> A.VRSetting:
    """VRFiniteVolume2D::Setting: the reference defaults, then `overrides`."""
    s = A.VRSetting()
    _lib().dndsmi_vr_setting_default(C.byref(s))
    for k, v in overrides.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


SETUP_SIZES = {  # name -> (per, count) where per is "cell", "face" or "c2f"
    "volume": ("cell", 1), "face_area": ("face", 1), "cell_bary": ("cell", 2), "base_moment": ("cell", 10),
    "face_center": ("face", 2), "face_norm_center": ("face", 2), "face_norm": ("face", 6),
    "face_weight": ("face", 10), "face_d0": ("face", 54), "face_dibj_g0": ("face", 180),
    "matrix_secondary": ("face", 162), "ainv": ("cell", 81), "ainv_b": ("c2f", 81), "ainv_bvec": ("c2f", 9),
    "face_gauss_xy": ("face", 12), "cell_scale": ("cell", 2), "cell_frame": ("cell", 4),
}


class ExternalMesh:
    """A mesh handed over as caller-owned arrays (keys = dndsmi_mesh_view
    fields), e.g. read from a file; exposes the same `view` as a built mesh."""

    def __init__(self, arrays: dict):
        self._keep = {}

        def d(name, dtype):
            a = np.ascontiguousarray(arrays[name], dtype=dtype).reshape(-1)
            self._keep[name] = a
            return a.ctypes.data_as(C.POINTER(C.c_double if dtype == np.float64 else
                                              (C.c_int64 if dtype == np.int64 else C.c_int32)))

        self.view = A.MeshView(int(arrays["n_node"]), int(arrays["n_cell"]), int(arrays["n_face"]),
                               d("node_xy", np.float64), d("cell_ptr", np.int64), d("cell2node", np.int64),
                               d("cell2face", np.int64), d("face2node", np.int64), d("face2cell", np.int64),
                               d("face2cell_ref", np.int32), d("face_bc", np.int32), d("face_shift", np.float64))
        sizes = {"node_xy": 2 * self.view.n_node, "cell_ptr": self.view.n_cell + 1, "face2node": 2 * self.view.n_face,
                 "face2cell": 2 * self.view.n_face, "face2cell_ref": 2 * self.view.n_face,
                 "face_bc": self.view.n_face, "face_shift": 2 * self.view.n_face}
        for name, n in sizes.items():
            if self._keep[name].size != n:
                raise DNDSError(A.ERR_INVALID_ARGUMENT, f"mesh array {name}: {self._keep[name].size} entries, "
                                                        f"expected {n}")
        nc2f = int(self._keep["cell_ptr"][-1])
        for name in ("cell2node", "cell2face"):
            if self._keep[name].size != nc2f:
                raise DNDSError(A.ERR_INVALID_ARGUMENT, f"mesh array {name}: {self._keep[name].size} entries, "
                                                        f"expected cell_ptr[-1] = {nc2f}")
        buf = C.create_string_buffer(512)
        rc = A.load().dndsmi_mesh_view_check(C.byref(self.view), buf, len(buf))
        if rc != A.OK:
            raise DNDSError(rc, f"mesh view: {buf.value.decode(errors='replace')}")
        self._arrays = dict(arrays)

    @property
    def n_cell(self):
        return int(self.view.n_cell)

    @property
    def n_face(self):
        return int(self.view.n_face)

    def arrays(self) -> dict:
        return self._arrays


class ExternalVR:
    """A VR setup handed over as caller-owned arrays (keys = dndsmi_vr_view
    fields) with its setting; stands in for VRFiniteVolume2D on the device."""

    def __init__(self, mesh: ExternalMesh, arrays: dict, setting: A.VRSetting):
        self.mesh = mesh
        self.setting = setting
        self.device = None
        self._keep = {}
        fields = {}
        for name, _ in A.VRView._fields_[3:]:
            a = arrays.get(name)
            if a is None or np.size(a) == 0:
                fields[name] = A.PD()
            else:
                a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                self._keep[name] = a
                fields[name] = a.ctypes.data_as(A.PD)
        m = mesh.view
        self.view = A.VRView(m.n_cell, m.n_face, int(mesh.arrays()["cell_ptr"][-1]), **fields)
        self._arrays = dict(arrays)

    def arrays(self, names=None) -> dict:
        return {k: v for k, v in self._arrays.items() if names is None or k in names}

    BuildRec = lambda self: ArrayRecV(self.device)  # noqa: E731

    def ReconstructionJacobiStep(self, u, uRec, uRecNewBuf, t=0.0):
        VRFiniteVolume2D.ReconstructionJacobiStep(self, u, uRec, uRecNewBuf, t)

    def ReconstructionWBAPLimitFacialV3(self, u, uRec, uRecNewBuf, uRecNewBuf1=None, ifUseLimiter=None, ifAll=False):
        VRFiniteVolume2D.ReconstructionWBAPLimitFacialV3(self, u, uRec, uRecNewBuf, uRecNewBuf1, ifUseLimiter, ifAll)

    def ReconstructionWBAPLimitFacialV2(self, u, uRec, uRecNewBuf, uRecNewBuf1=None, ifUseLimiter=None, ifAll=False):
        VRFiniteVolume2D.ReconstructionWBAPLimitFacialV2(self, u, uRec, uRecNewBuf, uRecNewBuf1, ifUseLimiter, ifAll)


class VRFiniteVolume2D:
    """Geometry + VR setup (host FP64) and the Jacobi sweep (device)."""

    def __init__(self, mesh: CompactFacedMeshSerialRW, setting: A.VRSetting | None = None, P_ORDER: int | None = None):
        """P_ORDER (DNDS_FV_VR.hpp:244): 3 or 1; None keeps the setting's
        rec_order (3 by default)."""
        self.mesh = mesh
        self.setting = setting if setting is not None else vr_setting()
        if P_ORDER is not None:
            self.setting.rec_order = P_ORDER
        self.P_ORDER = int(self.setting.rec_order)
        self._h = None
        self.view = None
        self.device = None  # set by EulerEvaluator

    def Initialization(self, nthreads: int = 0):
        h = C.c_void_p()
        _check(_lib().dndsmi_vr_create(C.byref(self.mesh.view), C.byref(self.setting), nthreads, C.byref(h)),
               "VR setup")
        self._h = h
        self.view = A.VRView()
        _check(_lib().dndsmi_vr_get_view(h, C.byref(self.view)), "VR view")
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().dndsmi_vr_destroy(self._h)
            self._h = None

    def arrays(self, names=None) -> dict:
        """numpy copies of setup arrays (keys = dndsmi_vr_view fields)."""
        v = self.view
        n = {"cell": int(v.n_cell), "face": int(v.n_face), "c2f": int(v.n_c2f)}
        out = {}
        for name, (per, cnt) in SETUP_SIZES.items():
            if names is not None and name not in names:
                continue
            ptr = getattr(v, name)
            out[name] = A.as_array(ptr, n[per] * cnt) if ptr else None
        return out

    def BuildRec(self):
        return ArrayRecV(self.device)

    def ReconstructionJacobiStep(self, u: "ArrayDOFV", uRec: "ArrayRecV", uRecNewBuf: "ArrayRecV", t: float = 0.0):
        """One reconstruction sweep (DNDS_FV_VR.hpp:1488-1712).  With
        SOR_Instead = false a Jacobi sweep into uRecNewBuf, then
        uRec = uRecNewBuf: both hold the new iterate.  With SOR_Instead an
        in-place sweep on uRec in the colour-major scan order of
        EulerEvaluator.LUSGSScanOrder (forward, then backward with
        SOR_InverseScanning); uRecNewBuf is not touched."""
        d = self.device
        _check(_lib().dndsmi_reconstruction_jacobi_step(d.ctx, u.h, uRec.h, uRecNewBuf.h, t), "Jacobi step", d.ctx)

    def ReconstructionWBAPLimitFacialV3(self, u: "ArrayDOFV", uRec: "ArrayRecV", uRecNewBuf: "ArrayRecV",
                                        uRecNewBuf1: "ArrayRecV | None" = None,
                                        ifUseLimiter: "ArrayScalar | None" = None, ifAll: bool = False):
        """WBAP V3 with the Euler characteristic matrices; uRecNewBuf receives
        the limited reconstruction, ifUseLimiter the smoothness indicator."""
        d = self.device
        _check(_lib().dndsmi_wbap_limit_facial_v3(d.ctx, u.h, uRec.h, uRecNewBuf.h,
                                                  uRecNewBuf1.h if uRecNewBuf1 is not None else None,
                                                  ifUseLimiter.h if ifUseLimiter is not None else None,
                                                  int(ifAll)), "WBAP V3", d.ctx)

    def ReconstructionWBAPLimitFacialV2(self, u: "ArrayDOFV", uRec: "ArrayRecV", uRecNewBuf: "ArrayRecV",
                                        uRecNewBuf1: "ArrayRecV | None" = None,
                                        ifUseLimiter: "ArrayScalar | None" = None, ifAll: bool = False):
        """WBAP V2 (the solver's default limiterProcedure 0) with the Euler
        characteristic matrices; arguments as for V3."""
        d = self.device
        _check(_lib().dndsmi_wbap_limit_facial_v2(d.ctx, u.h, uRec.h, uRecNewBuf.h,
                                                  uRecNewBuf1.h if uRecNewBuf1 is not None else None,
                                                  ifUseLimiter.h if ifUseLimiter is not None else None,
                                                  int(ifAll)), "WBAP V2", d.ctx)


def _reconstruct_to_tolerance(vfv, u, uRec, uRecNewBuf, tol, max_sweeps, t=0.0):
    d = vfv.device
    n, res = C.c_int32(), C.c_double()
    _check(_lib().dndsmi_reconstruct_to_tolerance(d.ctx, u.h, uRec.h, uRecNewBuf.h, tol, max_sweeps, t, C.byref(n),
                                                  C.byref(res)), "reconstruct to tolerance", d.ctx)
    return n.value, res.value


VRFiniteVolume2D.ReconstructToTolerance = _reconstruct_to_tolerance
VRFiniteVolume2D.ReconstructToTolerance.__doc__ = (
    "Jacobi sweeps until sum |uRec - uRecNew|^2 / n_cell < tol (the reference's rec increment, "
    "DNDS_FV_VR.hpp:1684-1705); returns (sweeps, last increment).")


class Device:
    """The device context: mesh + setup on the GPU (one per evaluator)."""

    def __init__(self, mesh: CompactFacedMeshSerialRW, vfv: VRFiniteVolume2D, esetting: A.EulerSetting,
                 stream: int | None = None):
        h = C.c_void_p()
        _check(_lib().dndsmi_ctx_create(C.byref(mesh.view), C.byref(vfv.view), C.byref(vfv.setting),
                                        C.byref(esetting), C.c_void_p(stream) if stream else None, C.byref(h)),
               "device context")
        self.ctx = h
        self.n_cell = mesh.n_cell

    def sync(self):
        _check(_lib().dndsmi_sync(self.ctx), "sync", self.ctx)

    def device_bytes(self) -> int:
        return int(_lib().dndsmi_ctx_device_bytes(self.ctx))

    def __del__(self):
        if getattr(self, "ctx", None):
            _lib().dndsmi_ctx_destroy(self.ctx)
            self.ctx = None


class _DeviceArray:
    KIND = None
    WIDTH = None

    def __init__(self, device: Device):
        self.device = device
        h = C.c_void_p()
        _check(_lib().dndsmi_array_create(device.ctx, self.KIND, C.byref(h)), "array create", device.ctx)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _lib().dndsmi_array_destroy(self.h)
            self.h = None

    def shape(self):
        return (self.device.n_cell, self.WIDTH) if self.WIDTH > 1 else (self.device.n_cell,)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape())
        _check(_lib().dndsmi_array_upload(self.h, A.dptr(a)), "upload", self.device.ctx)
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape(), dtype=np.float64)
        _check(_lib().dndsmi_array_download(self.h, A.dptr(out)), "download", self.device.ctx)
        return out

    def copy_from(self, other):
        _check(_lib().dndsmi_array_copy(self.h, other.h), "copy", self.device.ctx)
        return self

    def setConstant(self, value: float):
        _check(_lib().dndsmi_array_set_constant(self.h, float(value)), "set constant", self.device.ctx)
        return self


class ArrayDOFV(_DeviceArray):
    """ArrayDOFV<4>: 4 conservative variables per cell, host layout [n_cell][4]."""
    KIND = A.ARRAY_DOF
    WIDTH = 4


class ArrayRecV(_DeviceArray):
    """ArrayRecV: 9 x 4 rec coefficients per cell, host layout [n_cell][36]
    with element (r, v) at r + 9 v (column-major per cell)."""
    KIND = A.ARRAY_REC
    WIDTH = 36


class ArrayScalar(_DeviceArray):
    """One double per cell (dt, ifUseLimiter)."""
    KIND = A.ARRAY_SCALAR
    WIDTH = 1


RS = {"Roe": A.RS_ROE, "HLLC": A.RS_HLLC, "HLLEP": A.RS_HLLEP,
      **{f"Roe_M{k}": code for k, code in A.RS_ROE_M.items()}}


def solver_output(u, ifUseLimiter=None, rhs=None, gamma: float = 1.4, Rgas: float = 1.0,
                  smooth_indicator_scale: float = 1e-10):
    """(names, data[8][n_cell]) as EulerSolver::PrintData composes them for
    NS_2D (DNDS_FV_EulerSolver.hpp:1436-1492; nOUTS = nVars + 4, :55): rho,
    u, v, p, T = p / rho / Rgas, Mach, ifUseLimiter / (WBAP_SmoothIndicatorScale
    + 1e-200), rhs of rho.  The names are the first eight of the reference's
    list, R U V W P T M ifUseLimiter, as its 2D files carry them (the slots
    from the fourth on hold p, T, M, the indicator and the rhs)."""
    u = np.asarray(u, dtype=np.float64).reshape(-1, 4)
    n = u.shape[0]
    rho = u[:, 0]
    vel = u[:, 1:3] / rho[:, None]
    vsqr = np.sum(vel * vel, axis=1)
    p = (gamma - 1) * (u[:, 3] - 0.5 * rho * vsqr)
    asqr = gamma * p / rho
    ind = np.zeros(n) if ifUseLimiter is None else np.asarray(ifUseLimiter, dtype=np.float64).reshape(n)
    r0 = np.zeros(n) if rhs is None else np.asarray(rhs, dtype=np.float64).reshape(n, 4)[:, 0]
    data = np.stack([rho, vel[:, 0], vel[:, 1], p, p / rho / Rgas, np.sqrt(vsqr / asqr),
                     ind / (smooth_indicator_scale + 1e-200), r0])
    return ["R", "U", "V", "W", "P", "T", "M", "ifUseLimiter"], data


class EulerEvaluator:
    """EulerEvaluator<NS_2D>: the Euler model with muGas = 0 (the default
    here), the laminar Navier-Stokes terms with Sutherland's law for
    muGas > 0 (idealGasProperty, DNDS_FV_EulerEvaluator.hpp:151-160)."""

    def __init__(self, mesh: CompactFacedMeshSerialRW, vfv: VRFiniteVolume2D, riemannSolverType: str = "Roe",
                 gamma: float = 1.4, farFieldStaticValue=(1.0, 0.0, 0.0, 2.5), stream: int | None = None,
                 muGas: float = 0.0, prGas: float = 0.72, Rgas: float = 1.0, TRef: float = 273.15,
                 CSutherland: float = 110.4):
        self.mesh = mesh
        self.vfv = vfv
        self.settings = A.EulerSetting(riemann_solver=RS[riemannSolverType], gamma=gamma,
                                       far_field=(C.c_double * 4)(*farFieldStaticValue), mu_gas=muGas,
                                       pr_gas=prGas, r_gas=Rgas, t_ref=TRef, c_sutherland=CSutherland)
        self.device = Device(mesh, vfv, self.settings, stream)
        vfv.device = self.device

    def BuildMean(self):
        return ArrayDOFV(self.device)

    def BuildScalar(self):
        return ArrayScalar(self.device)

    def PrintData(self, fname: str, u: "ArrayDOFV", ifUseLimiter: "ArrayScalar | None" = None,
                  rhs: "ArrayDOFV | None" = None, t: float = 0.0, Rgas: float = 1.0):
        """EulerSolver::PrintData (DNDS_FV_EulerSolver.hpp:1436-1492): the
        solver's cell fields (solver_output) as a Tecplot binary file."""
        names, data = solver_output(u.download(), None if ifUseLimiter is None else ifUseLimiter.download(),
                                    None if rhs is None else rhs.download(), self.settings.gamma, Rgas,
                                    self.vfv.setting.wbap_smooth_indicator_scale)
        self.mesh.PrintSerialPartPltBinaryDataArray(fname, names, data, t)

    def EvaluateDt(self, dt: "ArrayScalar | list", u: ArrayDOFV, CFL: float, MaxDt: float = 1.0,
                   UseLocaldt: bool = False) -> float:
        """Returns dtMinall (the reference's output argument).  dt is a device
        ArrayScalar, or a Python list (the reference's std::vector<real>&,
        DNDS_FV_EulerEvaluator.hpp:1408-1411), which is refilled on the host."""
        if isinstance(dt, list):
            d = self.BuildScalar()
            dtmin = self.EvaluateDt(d, u, CFL, MaxDt, UseLocaldt)
            dt[:] = d.download().tolist()
            return dtmin
        dtmin = C.c_double()
        _check(_lib().dndsmi_evaluate_dt(self.device.ctx, dt.h, u.h, CFL, C.byref(dtmin), MaxDt, int(UseLocaldt)),
               "EvaluateDt", self.device.ctx)
        return dtmin.value

    def EvaluateRHS(self, rhs: ArrayDOFV, u: ArrayDOFV, uRec: ArrayRecV, t: float):
        _check(_lib().dndsmi_evaluate_rhs(self.device.ctx, rhs.h, u.h, uRec.h, t), "EvaluateRHS", self.device.ctx)

    def LUSGSColouring(self):
        """(colour of each cell, number of colours) of the multicolour sweeps."""
        n = C.c_int32()
        _check(_lib().dndsmi_lusgs_colouring(self.device.ctx, None, C.byref(n)), "LU-SGS colouring", self.device.ctx)
        col = np.empty(self.device.n_cell, dtype=np.int32)
        _check(_lib().dndsmi_lusgs_colouring(self.device.ctx, A.i32ptr(col), C.byref(n)), "LU-SGS colouring",
               self.device.ctx)
        return col, n.value

    def LUSGSScanOrder(self) -> np.ndarray:
        """The colour-major scan order of the device sweeps (iScan -> iCell)."""
        col, _ = self.LUSGSColouring()
        return np.argsort(col, kind="stable").astype(np.int64)

    def LUSGSMatrixInit(self, dTau: ArrayScalar, dt: float, alphaDiag: float, u: "ArrayDOFV | None" = None,
                        uRec: "ArrayRecV | None" = None, jacobianCode: int = 0, t: float = 0.0):
        """Diagonal from lambdaFace of the last EvaluateDt; u, uRec and t feed
        only the source-term Jacobian, absent in NS_2D; jacobianCode != 0 is
        unsupported (DNDSError).  dTau may be a host sequence (the reference's
        std::vector<real>&), uploaded for the call."""
        if not isinstance(dTau, ArrayScalar):
            dTau = self.BuildScalar().upload(np.asarray(dTau, dtype=np.float64))
        _check(_lib().dndsmi_lusgs_matrix_init(self.device.ctx, dTau.h, dt, alphaDiag,
                                               u.h if u is not None else None, uRec.h if uRec is not None else None,
                                               int(jacobianCode), t), "LUSGSMatrixInit", self.device.ctx)

    def UpdateLUSGSForward(self, alphaDiag: float, rhs: ArrayDOFV, u: ArrayDOFV, uInc: ArrayDOFV,
                           uIncNew: ArrayDOFV):
        _check(_lib().dndsmi_update_lusgs_forward(self.device.ctx, alphaDiag, rhs.h, u.h, uInc.h, uIncNew.h),
               "UpdateLUSGSForward", self.device.ctx)

    def UpdateLUSGSBackward(self, alphaDiag: float, rhs: ArrayDOFV, u: ArrayDOFV, uInc: ArrayDOFV,
                            uIncNew: ArrayDOFV):
        _check(_lib().dndsmi_update_lusgs_backward(self.device.ctx, alphaDiag, rhs.h, u.h, uInc.h, uIncNew.h),
               "UpdateLUSGSBackward", self.device.ctx)

    def LUSGSMatrixVec(self, alphaDiag: float, u: ArrayDOFV, uInc: ArrayDOFV, AuInc: ArrayDOFV):
        """EulerEvaluator::LUSGSMatrixVec (DNDS_FV_EulerEvaluator.cpp:1062-1145),
        with the diagonal of the last LUSGSMatrixInit."""
        _check(_lib().dndsmi_lusgs_matrix_vec(self.device.ctx, alphaDiag, u.h, uInc.h, AuInc.h), "LUSGSMatrixVec",
               self.device.ctx)

    def SetLinearSolver(self, gmresCode: int = 0, nGmresSpace: int = 10, nGmresIter: int = 2):
        """The dual-time steps' fsolve: 0 LU-SGS, 1 GMRES preconditioned by
        LU-SGS, 2 LU-SGS then GMRES (the solver's gmresCode, nGmresSpace,
        nGmresIter, DNDS_FV_EulerSolver.hpp:117-119)."""
        _check(_lib().dndsmi_set_linear_solver(self.device.ctx, gmresCode, nGmresSpace, nGmresIter),
               "linear solver", self.device.ctx)

    def EvaluateResidual(self, rhs: ArrayDOFV, P: int = 1, volWise: bool = False) -> np.ndarray:
        res = np.empty(4)
        _check(_lib().dndsmi_evaluate_residual(self.device.ctx, rhs.h, P, int(volWise), A.dptr(res)),
               "EvaluateResidual", self.device.ctx)
        return res


class ExplicitSSPRK3LocalDt:
    """ODE::ExplicitSSPRK3LocalDt with frhs = nInternalRecStep Jacobi sweeps
    (+ WBAP V2 or V3 by limiterProcedure, DNDS_FV_EulerSolver.hpp:796-806) +
    EvaluateRHS and fdt = EvaluateDt (DNDS_ODE.hpp:131-178); the stage times
    follow the reference's frhs arguments (t + dt, t + dt/4, t + 2dt/3)."""

    def __init__(self, evaluator: EulerEvaluator, nInternalRecStep: int, CFL: float, MaxDt: float = 1e100,
                 UseLocaldt: bool = False, useLimiter: bool = False, mixedPrecision: bool = False,
                 limiterProcedure: int = 0, warmStart: "WarmStart | None" = None, checkSteps: int = 5,
                 recheckInterval: int = 0, tolerance: float = 1e-6):
        """warmStart (mixedPrecision only; no reference counterpart): every
        stage adds the network's prediction on the change of the means to
        the previous stage's uRec; the plain method checks it over
        checkSteps steps (and again every recheckInterval steps) and the run
        falls back to the plain trajectory when the corrected state strays
        beyond `tolerance` (dndsmi_ssprk3_set_warm_start)."""
        self.ev = evaluator
        self.n_sweeps = nInternalRecStep
        self.CFL = CFL
        self.MaxDt = MaxDt
        self.UseLocaldt = UseLocaldt
        self.useLimiter = useLimiter
        self.mixedPrecision = mixedPrecision
        self.limiterProcedure = limiterProcedure
        self.warmStart = warmStart
        if warmStart is not None and not mixedPrecision:
            raise DNDSError(A.ERR_INVALID_ARGUMENT, "the learned warm start runs inside the mixed-precision step")
        if mixedPrecision:
            d = evaluator.device
            _check(_lib().dndsmi_ssprk3_set_warm_start(d.ctx, warmStart.h if warmStart is not None else None,
                                                       int(checkSteps), int(recheckInterval), float(tolerance)),
                   "warm start attach", d.ctx)

    def WarmStartStats(self) -> dict:
        """stages that added the correction, checks passed, checks failed
        (fallbacks), and the relative L2 deviation the last check found"""
        st = (C.c_int64 * 3)()
        dev = C.c_double()
        _check(_lib().dndsmi_ssprk3_warm_start_stats(self.ev.device.ctx, st, C.byref(dev)), "warm start stats")
        return {"corrected_stages": st[0], "checks_passed": st[1], "fallbacks": st[2],
                "last_check_deviation": dev.value}

    def Step(self, u: ArrayDOFV, uRec: ArrayRecV, uRecNew: ArrayRecV, t: float) -> float:
        """Advances u in place; returns the global minimum dt of the step.
        mixedPrecision: FP32 correction sweeps after one FP64 sweep per
        stage, FP32 face reconstruction and flux storage (no reference
        counterpart)."""
        dt = C.c_double()
        d = self.ev.device
        fn = _lib().dndsmi_ssprk3_step_mixed if self.mixedPrecision else _lib().dndsmi_ssprk3_step
        _check(_lib().dndsmi_set_limiter_procedure(d.ctx, self.limiterProcedure), "limiterProcedure", d.ctx)
        _check(fn(d.ctx, u.h, uRec.h, uRecNew.h, self.n_sweeps, int(self.useLimiter), self.CFL, self.MaxDt,
                  int(self.UseLocaldt), t, C.byref(dt)), "SSP-RK3 step", d.ctx)
        return dt.value


class ImplicitEulerDualTimeStep:
    """ODE::ImplicitEulerDualTimeStep (DNDS_ODE.hpp:201-265) with the solver's
    frhs (nInternalRecStep sweeps + EvaluateRHS), fdtau (EvaluateDt) and
    fsolve (LU-SGS forward + backward, jacobianTypeCode 0, gmresCode 0), and
    fstop on the density increment (rhsThresholdInternal), all on the device."""

    def __init__(self, evaluator: EulerEvaluator, nInternalRecStep: int, CFL: float, nTimeStepInternal: int,
                 rhsThresholdInternal: float, UseLocaldt: bool = True):
        self.ev = evaluator
        self.n_sweeps = nInternalRecStep
        self.CFL = CFL
        self.max_iter = nTimeStepInternal
        self.threshold = rhsThresholdInternal
        self.UseLocaldt = UseLocaldt

    def Step(self, u: ArrayDOFV, uRec: ArrayRecV, uRecNew: ArrayRecV, dt: float, t: float = 0.0):
        """One implicit-Euler step of size dt; returns the L1 increment norms
        of the inner iterations ([iterations][4])."""
        d = self.ev.device
        hist = np.zeros((self.max_iter, 4))
        n = C.c_int32()
        _check(_lib().dndsmi_implicit_euler_step(d.ctx, u.h, uRec.h, uRecNew.h, self.n_sweeps, self.CFL, dt,
                                                 int(self.UseLocaldt), self.max_iter, self.threshold, t,
                                                 C.byref(n), A.dptr(hist)), "implicit Euler step", d.ctx)
        return hist[:n.value]


class ImplicitSDIRK4DualTimeStep:
    """ODE::ImplicitSDIRK4DualTimeStep (DNDS_ODE.hpp:270-372), three stages
    with the same device frhs / fdtau / fsolve / fstop as
    ImplicitEulerDualTimeStep."""

    def __init__(self, evaluator: EulerEvaluator, nInternalRecStep: int, CFL: float, nTimeStepInternal: int,
                 rhsThresholdInternal: float, UseLocaldt: bool = True):
        self.ev = evaluator
        self.n_sweeps = nInternalRecStep
        self.CFL = CFL
        self.max_iter = nTimeStepInternal
        self.threshold = rhsThresholdInternal
        self.UseLocaldt = UseLocaldt

    def Step(self, u: ArrayDOFV, uRec: ArrayRecV, uRecNew: ArrayRecV, dt: float, t: float = 0.0):
        """One step of size dt; returns (iterations per stage [3], L1
        increment norms [3][max_iter][4], zero past each stage's end)."""
        d = self.ev.device
        hist = np.zeros((3, self.max_iter, 4))
        its = np.zeros(3, dtype=np.int32)
        _check(_lib().dndsmi_sdirk4_step(d.ctx, u.h, uRec.h, uRecNew.h, self.n_sweeps, self.CFL, dt,
                                         int(self.UseLocaldt), self.max_iter, self.threshold, t,
                                         its.ctypes.data_as(A.PI32), A.dptr(hist)), "SDIRK4 step", d.ctx)
        return its, hist


class ImplicitBDFDualTimeStep:
    """ODE::ImplicitBDFDualTimeStep (DNDS_ODE.hpp:392-512) of order k with the
    same device frhs / fdtau / fsolve / fstop as ImplicitEulerDualTimeStep;
    the k-1 previous states live on the device."""

    def __init__(self, evaluator: EulerEvaluator, nInternalRecStep: int, CFL: float, nTimeStepInternal: int,
                 rhsThresholdInternal: float, UseLocaldt: bool = True, k: int = 2):
        self.ev = evaluator
        self.n_sweeps = nInternalRecStep
        self.CFL = CFL
        self.max_iter = nTimeStepInternal
        self.threshold = rhsThresholdInternal
        self.UseLocaldt = UseLocaldt
        h = C.c_void_p()
        _check(_lib().dndsmi_bdf_create(evaluator.device.ctx, k, C.byref(h)), "BDF create", evaluator.device.ctx)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            _lib().dndsmi_bdf_destroy(self.h)
            self.h = None

    def Step(self, u: ArrayDOFV, uRec: ArrayRecV, uRecNew: ArrayRecV, dt: float, t: float = 0.0):
        """One BDF step of size dt; returns the L1 increment norms of the
        inner iterations ([iterations][4])."""
        d = self.ev.device
        hist = np.zeros((self.max_iter, 4))
        n = C.c_int32()
        _check(_lib().dndsmi_bdf_step(self.h, u.h, uRec.h, uRecNew.h, self.n_sweeps, self.CFL, dt,
                                      int(self.UseLocaldt), self.max_iter, self.threshold, t, C.byref(n),
                                      A.dptr(hist)), "BDF step", d.ctx)
        return hist[:n.value]


class WarmStart:
    """The learned warm start of the reconstruction (no reference
    counterpart): features, training set, training, inference on the device."""

    def __init__(self, evaluator: EulerEvaluator, seed: int = 0, capacity: int = 0):
        self.ev = evaluator
        h = C.c_void_p()
        d = evaluator.device
        _check(_lib().dndsmi_warmstart_create(d.ctx, C.byref(evaluator.mesh.view), C.byref(evaluator.vfv.view), seed,
                                              capacity, C.byref(h)), "warm start create", d.ctx)
        self.h = h
        self.n_param = int(_lib().dndsmi_warmstart_num_params())

    def __del__(self):
        if getattr(self, "h", None):
            _lib().dndsmi_warmstart_destroy(self.h)
            self.h = None

    def _ctx(self):
        return self.ev.device.ctx

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_param, dtype=np.float32)
        _check(_lib().dndsmi_warmstart_get_params(self.h, A.fptr(p)), "get params", self._ctx())
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib().dndsmi_warmstart_set_params(self.h, A.fptr(p)), "set params", self._ctx())

    def features(self, u: ArrayDOFV):
        n = self.ev.device.n_cell
        X = np.empty((n, 58), dtype=np.float32)
        S = np.empty((n, 4), dtype=np.float32)
        _check(_lib().dndsmi_warmstart_features(self.h, u.h, A.fptr(X), A.fptr(S)), "features", self._ctx())
        return X, S

    def add_samples(self, u: ArrayDOFV, uRec: ArrayRecV):
        _check(_lib().dndsmi_warmstart_add_samples(self.h, u.h, uRec.h), "add samples", self._ctx())

    def add_rows(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        _check(_lib().dndsmi_warmstart_add_rows(self.h, A.fptr(X), A.fptr(T), X.shape[0]), "add rows", self._ctx())

    def targets(self, u: ArrayDOFV, uRec: ArrayRecV) -> np.ndarray:
        T = np.empty((self.ev.device.n_cell, 36), dtype=np.float32)
        _check(_lib().dndsmi_warmstart_targets(self.h, u.h, uRec.h, A.fptr(T)), "targets", self._ctx())
        return T

    @property
    def n_samples(self) -> int:
        return int(_lib().dndsmi_warmstart_num_samples(self.h))

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.n_param, dtype=np.float32)
        loss = C.c_double()
        _check(_lib().dndsmi_warmstart_loss_grad(self.h, A.fptr(X), A.fptr(T), X.shape[0], A.fptr(g),
                                                 C.byref(loss)), "loss/grad", self._ctx())
        return loss.value, g

    def adam_step(self, grad, lr):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        _check(_lib().dndsmi_warmstart_adam_step(self.h, A.fptr(g), lr), "adam step", self._ctx())

    def train(self, epochs: int, batch: int = 4096, lr: float = 2e-3, lr_final: float = 1e-5) -> np.ndarray:
        hist = np.zeros(epochs)
        _check(_lib().dndsmi_warmstart_train(self.h, epochs, batch, lr, lr_final, A.dptr(hist)), "train", self._ctx())
        return hist

    def predict(self, u: ArrayDOFV, uRec: ArrayRecV):
        _check(_lib().dndsmi_warmstart_predict(self.h, u.h, uRec.h), "predict", self._ctx())
