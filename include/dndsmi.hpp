/*
 * dndsmi.hpp -- C++ host classes over the C ABI of dndsmi.h, with the
 * reference's class and method names and argument meaning, so that a caller
 * of DNDS's VRFiniteVolume2D / EulerEvaluator / ODE classes can switch to the
 * GPU path by changing types, not call sites:
 *
 *   reference (DNDS::...)                          here (dndsmi::...)
 *   CompactFacedMeshSerialRW   DNDS_Mesh.hpp:634   CompactFacedMeshSerialRW
 *   VRFiniteVolume2D           DNDS_FV_VR.hpp:143  VRFiniteVolume2D
 *   EulerEvaluator<NS_2D>      DNDS_FV_EulerEvaluator.hpp:71   EulerEvaluator
 *   ArrayDOFV<4>, ArrayRecV    DNDS_DerivedArray.hpp:302, 446  ArrayDOFV, ArrayRecV
 *   std::vector<real> dt       (EvaluateDt, LUSGSMatrixInit)   ArrayScalar, or std::vector<real>
 *                              through an overload that moves it across PCIe
 *   ODE::ExplicitSSPRK3LocalDt DNDS_ODE.hpp:95-179 ExplicitSSPRK3LocalDt
 *   ODE::ImplicitEulerDualTimeStep DNDS_ODE.hpp:201-265 ImplicitEulerDualTimeStep
 *
 * Differences, all forced by the ABI or the device:
 *  - the reference's lambdas (FBoundary, fML/fMR, frhs/fdt/fsolve) are fixed
 *    device functions selected by the settings (BC kinds, gas, Riemann solver);
 *  - arrays live on the device; upload/download convert from/to the
 *    reference's host layouts (u [cell][4]; uRec [cell][9x4 column-major]);
 *  - the reference aborts on a failed assertion (DNDS_assert,
 *    DNDS_Defines.h:31-44); here every failure throws dndsmi::Error with the
 *    library's status and message.
 * Header-only; link with libdndsmi.so.
 */
#ifndef DNDSMI_HPP
#define DNDSMI_HPP

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dndsmi.h"

namespace dndsmi {

using real = double;
using index = int64_t;

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

namespace detail {
inline void check(int rc, const char *what, const dndsmi_ctx *ctx = nullptr) {
    if (rc == DNDSMI_OK)
        return;
    std::string msg = std::string(what) + " failed with status " + std::to_string(rc);
    if (ctx) {
        char buf[512];
        if (dndsmi_last_error(ctx, buf, sizeof(buf)) == DNDSMI_OK && buf[0])
            msg += ": " + std::string(buf);
    }
    throw Error(rc, msg);
}
} // namespace detail

class CompactFacedMeshSerialRW {
  public:
    static CompactFacedMeshSerialRW generate(const dndsmi_mesh_spec &spec) {
        dndsmi_mesh *m = nullptr;
        detail::check(dndsmi_mesh_generate(&spec, &m), "mesh generate");
        return CompactFacedMeshSerialRW(m);
    }
    /* the Gmsh-reader path: raw cells and boundary edges */
    static CompactFacedMeshSerialRW from_cells(const std::vector<double> &node_xy, const std::vector<int64_t> &cell_ptr,
                                               const std::vector<int64_t> &cell2node,
                                               const std::vector<int64_t> &bnd2node,
                                               const std::vector<int32_t> &bnd_bc) {
        dndsmi_mesh *m = nullptr;
        detail::check(dndsmi_mesh_from_cells(int64_t(node_xy.size() / 2), node_xy.data(), int64_t(cell_ptr.size()) - 1,
                                             cell_ptr.data(), cell2node.data(), int64_t(bnd_bc.size()),
                                             bnd2node.data(), bnd_bc.data(), &m),
                      "mesh from cells");
        return CompactFacedMeshSerialRW(m);
    }
    /* CompactFacedMeshSerialRWBuild's reading step: a Gmsh 2.2 ASCII file */
    static CompactFacedMeshSerialRW read_gmsh(const std::string &fname) {
        dndsmi_mesh *m = nullptr;
        char buf[1024] = {0};
        const int rc = dndsmi_mesh_read_gmsh(fname.c_str(), &m, buf, sizeof(buf));
        if (rc != DNDSMI_OK)
            throw Error(rc, "Gmsh read failed: " + std::string(buf));
        return CompactFacedMeshSerialRW(m);
    }
    /* PrintSerialPartPltBinaryDataArray(fname, 0, arraySiz, names, data, t)
     * on one rank: names(k) and data(k, iCell) as in the reference */
    template <class FNAMES, class FDATA>
    void PrintSerialPartPltBinaryDataArray(const std::string &fname, int arraySiz, FNAMES &&names, FDATA &&data,
                                           double t) const {
        std::vector<std::string> nm(static_cast<size_t>(arraySiz));
        std::vector<const char *> np(static_cast<size_t>(arraySiz));
        std::vector<double> d(static_cast<size_t>(arraySiz) * static_cast<size_t>(NumCell()));
        for (int k = 0; k < arraySiz; k++) {
            nm[size_t(k)] = names(k);
            np[size_t(k)] = nm[size_t(k)].c_str();
            for (index i = 0; i < NumCell(); i++)
                d[size_t(k) * size_t(NumCell()) + size_t(i)] = data(k, i);
        }
        char buf[1024] = {0};
        const int rc = dndsmi_mesh_write_plt(h_.get(), fname.c_str(), arraySiz, np.data(), d.data(), t, buf,
                                             sizeof(buf));
        if (rc != DNDSMI_OK)
            throw Error(rc, "Tecplot write failed: " + std::string(buf));
    }
    const dndsmi_mesh_view &view() const { return view_; }
    index NumCell() const { return view_.n_cell; }
    index NumFace() const { return view_.n_face; }

  private:
    explicit CompactFacedMeshSerialRW(dndsmi_mesh *m) : h_(m, &dndsmi_mesh_destroy) {
        detail::check(dndsmi_mesh_get_view(m, &view_), "mesh view");
    }
    std::shared_ptr<dndsmi_mesh> h_;
    dndsmi_mesh_view view_{};
};

class Device;

/* device arrays in the reference's logical layouts */
class DeviceArray {
  public:
    DeviceArray(const std::shared_ptr<Device> &dev, int32_t kind, int width);
    void upload(const std::vector<real> &host) {
        if (host.size() != size())
            throw Error(DNDSMI_ERR_INVALID_ARGUMENT, "upload: size mismatch");
        detail::check(dndsmi_array_upload(h_.get(), host.data()), "upload");
    }
    std::vector<real> download() const {
        std::vector<real> out(size());
        detail::check(dndsmi_array_download(h_.get(), out.data()), "download");
        return out;
    }
    void setConstant(real v) { detail::check(dndsmi_array_set_constant(h_.get(), v), "setConstant"); }
    void copyFrom(const DeviceArray &o) { detail::check(dndsmi_array_copy(h_.get(), o.h_.get()), "copy"); }
    size_t size() const;
    dndsmi_array *h() const { return h_.get(); }

  private:
    std::shared_ptr<Device> dev_;
    std::shared_ptr<dndsmi_array> h_;
    int width_;
};

class ArrayDOFV : public DeviceArray { /* ArrayDOFV<4>: [cell][4] */
  public:
    explicit ArrayDOFV(const std::shared_ptr<Device> &d) : DeviceArray(d, DNDSMI_ARRAY_DOF, 4) {}
};
class ArrayRecV : public DeviceArray { /* ArrayRecV: [cell][r + 9 v] */
  public:
    explicit ArrayRecV(const std::shared_ptr<Device> &d) : DeviceArray(d, DNDSMI_ARRAY_REC, 36) {}
};
class ArrayScalar : public DeviceArray { /* one real per cell (dt, ifUseLimiter) */
  public:
    explicit ArrayScalar(const std::shared_ptr<Device> &d) : DeviceArray(d, DNDSMI_ARRAY_SCALAR, 1) {}
};

class VRFiniteVolume2D {
  public:
    using Setting = dndsmi_vr_setting;
    static Setting DefaultSetting() {
        Setting s;
        dndsmi_vr_setting_default(&s);
        return s;
    }
    VRFiniteVolume2D(const CompactFacedMeshSerialRW &mesh, const Setting &setting)
        : mesh_(mesh), setting_(setting) {}
    /* geometry + VR setup on the host (nthreads 0: hardware concurrency) */
    void Initialization(int nthreads = 0) {
        dndsmi_vr *v = nullptr;
        detail::check(dndsmi_vr_create(&mesh_.view(), &setting_, nthreads, &v), "VR setup");
        h_.reset(v, &dndsmi_vr_destroy);
        detail::check(dndsmi_vr_get_view(v, &view_), "VR view");
    }
    const dndsmi_vr_view &view() const { return view_; }
    const Setting &setting() const { return setting_; }
    const CompactFacedMeshSerialRW &mesh() const { return mesh_; }

    /* after the sweep uRec holds the new iterate (DNDS_FV_VR.hpp:1707-1711) */
    void ReconstructionJacobiStep(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNewBuf, real t = 0);
    void ReconstructionWBAPLimitFacialV3(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNewBuf,
                                         ArrayRecV *uRecNewBuf1, ArrayScalar *ifUseLimiter, bool ifAll);
    void ReconstructionWBAPLimitFacialV2(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNewBuf,
                                         ArrayRecV *uRecNewBuf1, ArrayScalar *ifUseLimiter, bool ifAll);
    ArrayRecV BuildRec() const;

  private:
    friend class EulerEvaluator;
    CompactFacedMeshSerialRW mesh_;
    Setting setting_;
    std::shared_ptr<dndsmi_vr> h_;
    dndsmi_vr_view view_{};
    std::shared_ptr<Device> dev_;
};

class Device {
  public:
    Device(const CompactFacedMeshSerialRW &mesh, const VRFiniteVolume2D &vfv, const dndsmi_euler_setting &es,
           void *stream = nullptr) : n_cell(mesh.NumCell()) {
        dndsmi_ctx *c = nullptr;
        detail::check(dndsmi_ctx_create(&mesh.view(), &vfv.view(), &vfv.setting(), &es, stream, &c), "device context");
        ctx.reset(c, &dndsmi_ctx_destroy);
    }
    void sync() { detail::check(dndsmi_sync(ctx.get()), "sync", ctx.get()); }
    std::shared_ptr<dndsmi_ctx> ctx;
    index n_cell;
};

inline DeviceArray::DeviceArray(const std::shared_ptr<Device> &dev, int32_t kind, int width)
    : dev_(dev), width_(width) {
    dndsmi_array *a = nullptr;
    detail::check(dndsmi_array_create(dev->ctx.get(), kind, &a), "array create", dev->ctx.get());
    h_.reset(a, &dndsmi_array_destroy);
}
inline size_t DeviceArray::size() const { return size_t(dev_->n_cell) * size_t(width_); }

inline void VRFiniteVolume2D::ReconstructionJacobiStep(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNewBuf, real t) {
    detail::check(dndsmi_reconstruction_jacobi_step(dev_->ctx.get(), u.h(), uRec.h(), uRecNewBuf.h(), t),
                  "ReconstructionJacobiStep", dev_->ctx.get());
}
inline void VRFiniteVolume2D::ReconstructionWBAPLimitFacialV3(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNewBuf,
                                                              ArrayRecV *uRecNewBuf1, ArrayScalar *ifUseLimiter,
                                                              bool ifAll) {
    detail::check(dndsmi_wbap_limit_facial_v3(dev_->ctx.get(), u.h(), uRec.h(), uRecNewBuf.h(),
                                              uRecNewBuf1 ? uRecNewBuf1->h() : nullptr,
                                              ifUseLimiter ? ifUseLimiter->h() : nullptr, ifAll),
                  "ReconstructionWBAPLimitFacialV3", dev_->ctx.get());
}
inline void VRFiniteVolume2D::ReconstructionWBAPLimitFacialV2(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNewBuf,
                                                              ArrayRecV *uRecNewBuf1, ArrayScalar *ifUseLimiter,
                                                              bool ifAll) {
    detail::check(dndsmi_wbap_limit_facial_v2(dev_->ctx.get(), u.h(), uRec.h(), uRecNewBuf.h(),
                                              uRecNewBuf1 ? uRecNewBuf1->h() : nullptr,
                                              ifUseLimiter ? ifUseLimiter->h() : nullptr, ifAll),
                  "ReconstructionWBAPLimitFacialV2", dev_->ctx.get());
}
inline ArrayRecV VRFiniteVolume2D::BuildRec() const { return ArrayRecV(dev_); }

class EulerEvaluator {
  public:
    /* muGas = 0: the Euler model; muGas > 0: laminar NS_2D with Sutherland's law */
    struct Setting {
        int riemannSolverType = DNDSMI_RS_ROE;
        real gamma = 1.4;
        std::array<real, 4> farFieldStaticValue{1.0, 0.0, 0.0, 2.5};
        real muGas = 0, prGas = 0.72, Rgas = 1, TRef = 273.15, CSutherland = 110.4;
    };
    EulerEvaluator(const CompactFacedMeshSerialRW &mesh, VRFiniteVolume2D &vfv, const Setting &s,
                   void *stream = nullptr) {
        dndsmi_euler_setting es{};
        es.riemann_solver = s.riemannSolverType;
        es.gamma = s.gamma;
        for (int k = 0; k < 4; k++)
            es.far_field[k] = s.farFieldStaticValue[k];
        es.mu_gas = s.muGas, es.pr_gas = s.prGas, es.r_gas = s.Rgas, es.t_ref = s.TRef;
        es.c_sutherland = s.CSutherland;
        dev_ = std::make_shared<Device>(mesh, vfv, es, stream);
        vfv.dev_ = dev_;
    }
    ArrayDOFV BuildMean() const { return ArrayDOFV(dev_); }
    ArrayScalar BuildScalar() const { return ArrayScalar(dev_); }
    const std::shared_ptr<Device> &device() const { return dev_; }

    void EvaluateDt(ArrayScalar &dt, ArrayDOFV &u, real CFL, real &dtMinall, real MaxDt = 1, bool UseLocaldt = false) {
        detail::check(dndsmi_evaluate_dt(ctx(), dt.h(), u.h(), CFL, &dtMinall, MaxDt, UseLocaldt), "EvaluateDt", ctx());
    }
    /* the reference's std::vector<real> form (DNDS_FV_EulerEvaluator.hpp:1408-1411): dt is resized and
     * filled on the host */
    void EvaluateDt(std::vector<real> &dt, ArrayDOFV &u, real CFL, real &dtMinall, real MaxDt = 1,
                    bool UseLocaldt = false) {
        ArrayScalar d(dev_);
        EvaluateDt(d, u, CFL, dtMinall, MaxDt, UseLocaldt);
        dt = d.download();
    }
    void EvaluateRHS(ArrayDOFV &rhs, ArrayDOFV &u, ArrayRecV &uRec, real t) {
        detail::check(dndsmi_evaluate_rhs(ctx(), rhs.h(), u.h(), uRec.h(), t), "EvaluateRHS", ctx());
    }
    /* DNDS_FV_EulerEvaluator.hpp:1429-1433; jacobianCode != 0 throws (unsupported) */
    void LUSGSMatrixInit(ArrayScalar &dTau, real dt, real alphaDiag, ArrayDOFV &u, ArrayRecV &uRec, int jacobianCode,
                         real t) {
        detail::check(dndsmi_lusgs_matrix_init(ctx(), dTau.h(), dt, alphaDiag, u.h(), uRec.h(), jacobianCode, t),
                      "LUSGSMatrixInit", ctx());
    }
    /* the reference's std::vector<real> form: dTau uploaded from the host */
    void LUSGSMatrixInit(std::vector<real> &dTau, real dt, real alphaDiag, ArrayDOFV &u, ArrayRecV &uRec,
                         int jacobianCode, real t) {
        ArrayScalar d(dev_);
        d.upload(dTau);
        LUSGSMatrixInit(d, dt, alphaDiag, u, uRec, jacobianCode, t);
    }
    void UpdateLUSGSForward(real alphaDiag, ArrayDOFV &rhs, ArrayDOFV &u, ArrayDOFV &uInc, ArrayDOFV &uIncNew) {
        detail::check(dndsmi_update_lusgs_forward(ctx(), alphaDiag, rhs.h(), u.h(), uInc.h(), uIncNew.h()),
                      "UpdateLUSGSForward", ctx());
    }
    void UpdateLUSGSBackward(real alphaDiag, ArrayDOFV &rhs, ArrayDOFV &u, ArrayDOFV &uInc, ArrayDOFV &uIncNew) {
        detail::check(dndsmi_update_lusgs_backward(ctx(), alphaDiag, rhs.h(), u.h(), uInc.h(), uIncNew.h()),
                      "UpdateLUSGSBackward", ctx());
    }
    void LUSGSMatrixVec(real alphaDiag, ArrayDOFV &u, ArrayDOFV &uInc, ArrayDOFV &AuInc) {
        detail::check(dndsmi_lusgs_matrix_vec(ctx(), alphaDiag, u.h(), uInc.h(), AuInc.h()), "LUSGSMatrixVec", ctx());
    }
    /* the dual-time steps' fsolve: gmresCode 0 LU-SGS, 1 GMRES, 2 LU-SGS then GMRES */
    void SetLinearSolver(int gmresCode, int nGmresSpace = 10, int nGmresIter = 2) {
        detail::check(dndsmi_set_linear_solver(ctx(), gmresCode, nGmresSpace, nGmresIter), "SetLinearSolver", ctx());
    }
    void EvaluateResidual(std::array<real, 4> &res, ArrayDOFV &rhs, index P = 1, bool volWise = false) {
        detail::check(dndsmi_evaluate_residual(ctx(), rhs.h(), int32_t(P), volWise, res.data()), "EvaluateResidual",
                      ctx());
    }

  private:
    dndsmi_ctx *ctx() const { return dev_->ctx.get(); }
    std::shared_ptr<Device> dev_;
};

/* ODE::ExplicitSSPRK3LocalDt with frhs = nInternalRecStep sweeps (+ WBAP V2
 * or V3 by limiterProcedure) + EvaluateRHS and fdt = EvaluateDt */
class ExplicitSSPRK3LocalDt {
  public:
    /* mixedPrecision: FP32 correction sweeps after one FP64 sweep per stage and
     * FP32 face reconstruction (dndsmi_ssprk3_step_mixed; no reference counterpart) */
    ExplicitSSPRK3LocalDt(EulerEvaluator &ev, int nInternalRecStep, real CFL, real MaxDt = 1e100,
                          bool UseLocaldt = false, bool useLimiter = false, bool mixedPrecision = false,
                          int limiterProcedure = 0)
        : ev_(ev), n_(nInternalRecStep), cfl_(CFL), maxdt_(MaxDt), local_(UseLocaldt), lim_(useLimiter),
          mixed_(mixedPrecision), proc_(limiterProcedure) {}
    /* advances u in place; returns the global minimum dt of the step */
    real Step(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNew, real t) {
        real dt = 0;
        dndsmi_ctx *c = ev_.device()->ctx.get();
        detail::check(dndsmi_set_limiter_procedure(c, proc_), "limiterProcedure", c);
        auto fn = mixed_ ? &dndsmi_ssprk3_step_mixed : &dndsmi_ssprk3_step;
        detail::check(fn(c, u.h(), uRec.h(), uRecNew.h(), n_, lim_, cfl_, maxdt_, local_, t, &dt), "SSP-RK3 step", c);
        return dt;
    }

  private:
    EulerEvaluator &ev_;
    int n_;
    real cfl_, maxdt_;
    bool local_, lim_, mixed_;
    int proc_;
};

/* ODE::ImplicitEulerDualTimeStep with LU-SGS (jacobianTypeCode 0, gmresCode 0) */
class ImplicitEulerDualTimeStep {
  public:
    ImplicitEulerDualTimeStep(EulerEvaluator &ev, int nInternalRecStep, real CFL, int nTimeStepInternal,
                              real rhsThresholdInternal, bool UseLocaldt = true)
        : ev_(ev), n_(nInternalRecStep), cfl_(CFL), maxit_(nTimeStepInternal), thr_(rhsThresholdInternal),
          local_(UseLocaldt) {}
    /* one step of size dt; returns the inner L1 increment norms (4 per iteration) */
    std::vector<real> Step(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNew, real dt, real t = 0) {
        std::vector<real> hist(size_t(maxit_) * 4, 0.0);
        int32_t n = 0;
        dndsmi_ctx *c = ev_.device()->ctx.get();
        detail::check(dndsmi_implicit_euler_step(c, u.h(), uRec.h(), uRecNew.h(), n_, cfl_, dt, local_, maxit_, thr_,
                                                 t, &n, hist.data()),
                      "implicit Euler step", c);
        hist.resize(size_t(n) * 4);
        return hist;
    }

  private:
    EulerEvaluator &ev_;
    int n_;
    real cfl_;
    int maxit_;
    real thr_;
    bool local_;
};

/* ODE::ImplicitSDIRK4DualTimeStep (DNDS_ODE.hpp:270-372): three stages with
 * the frhs / fdtau / fsolve / fstop of ImplicitEulerDualTimeStep. */
class ImplicitSDIRK4DualTimeStep {
  public:
    ImplicitSDIRK4DualTimeStep(EulerEvaluator &ev, int nInternalRecStep, real CFL, int nTimeStepInternal,
                               real rhsThresholdInternal, bool UseLocaldt = true)
        : ev_(ev), n_(nInternalRecStep), cfl_(CFL), maxit_(nTimeStepInternal), thr_(rhsThresholdInternal),
          local_(UseLocaldt) {}
    /* one step of size dt; returns the iterations of the three stages */
    std::array<int32_t, 3> Step(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNew, real dt, real t = 0) {
        std::array<int32_t, 3> its{};
        dndsmi_ctx *c = ev_.device()->ctx.get();
        detail::check(dndsmi_sdirk4_step(c, u.h(), uRec.h(), uRecNew.h(), n_, cfl_, dt, local_, maxit_, thr_, t,
                                         its.data(), nullptr),
                      "SDIRK4 step", c);
        return its;
    }

  private:
    EulerEvaluator &ev_;
    int n_;
    real cfl_;
    int maxit_;
    real thr_;
    bool local_;
};

/* ODE::ImplicitBDFDualTimeStep (DNDS_ODE.hpp:392-512): order k, the k-1
 * previous states on the device; frhs / fdtau / fsolve / fstop as
 * ImplicitEulerDualTimeStep.  Must not outlive the evaluator's device. */
class ImplicitBDFDualTimeStep {
  public:
    ImplicitBDFDualTimeStep(EulerEvaluator &ev, int nInternalRecStep, real CFL, int nTimeStepInternal,
                            real rhsThresholdInternal, bool UseLocaldt = true, int k = 2)
        : ev_(ev), n_(nInternalRecStep), cfl_(CFL), maxit_(nTimeStepInternal), thr_(rhsThresholdInternal),
          local_(UseLocaldt) {
        dndsmi_ctx *c = ev_.device()->ctx.get();
        detail::check(dndsmi_bdf_create(c, k, &bdf_), "BDF create", c);
    }
    ~ImplicitBDFDualTimeStep() { dndsmi_bdf_destroy(bdf_); }
    ImplicitBDFDualTimeStep(const ImplicitBDFDualTimeStep &) = delete;
    ImplicitBDFDualTimeStep &operator=(const ImplicitBDFDualTimeStep &) = delete;
    /* one step of size dt; returns the inner L1 increment norms (4 per iteration) */
    std::vector<real> Step(ArrayDOFV &u, ArrayRecV &uRec, ArrayRecV &uRecNew, real dt, real t = 0) {
        std::vector<real> hist(size_t(maxit_) * 4, 0.0);
        int32_t n = 0;
        dndsmi_ctx *c = ev_.device()->ctx.get();
        detail::check(dndsmi_bdf_step(bdf_, u.h(), uRec.h(), uRecNew.h(), n_, cfl_, dt, local_, maxit_, thr_, t, &n,
                                      hist.data()),
                      "BDF step", c);
        hist.resize(size_t(n) * 4);
        return hist;
    }

  private:
    EulerEvaluator &ev_;
    int n_;
    real cfl_;
    int maxit_;
    real thr_;
    bool local_;
    dndsmi_bdf *bdf_ = nullptr;
};

} // namespace dndsmi

#endif
