// Drives the C++ host classes of dndsmi.hpp the way a DNDS solver drives its
// own: mesh, VR setup, evaluator, explicit SSP-RK3 steps with WBAP, then an
// implicit-Euler LU-SGS step.  Writes u then uRec (host layouts, doubles) to
// argv[1]; the test compares them with the same run through the Python host.
#include <cmath>
#include <cstdio>
#include <vector>

#include "dndsmi.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: facade_run OUT\n");
        return 2;
    }
    try {
        dndsmi_mesh_spec spec{};
        spec.kind = DNDSMI_MESH_TRI_SPLIT;
        spec.nx = 10, spec.ny = 8;
        spec.x0 = 0, spec.x1 = 1, spec.y0 = 0, spec.y1 = 1;
        spec.jitter = 0.2, spec.y_ratio = 1.0, spec.seed = 5;
        spec.bc_xmin = spec.bc_xmax = spec.bc_ymin = spec.bc_ymax = DNDSMI_BC_FARFIELD;
        auto mesh = dndsmi::CompactFacedMeshSerialRW::generate(spec);
        auto setting = dndsmi::VRFiniteVolume2D::DefaultSetting();
        setting.build_limiter_data = 1;
        setting.far_weight = 1.0;
        dndsmi::VRFiniteVolume2D vfv(mesh, setting);
        vfv.Initialization(2);
        dndsmi::EulerEvaluator::Setting es;
        es.riemannSolverType = DNDSMI_RS_HLLC;
        es.farFieldStaticValue = {1.0, 0.3, -0.2, 2.6};
        dndsmi::EulerEvaluator ev(mesh, vfv, es);
        const auto n = mesh.NumCell();
        std::vector<double> u0(size_t(n) * 4);
        for (int64_t i = 0; i < n; i++) {
            const double x = vfv.view().cell_bary[2 * i], y = vfv.view().cell_bary[2 * i + 1];
            const double s = 1 + 0.1 * std::sin(2 * M_PI * x) * std::cos(2 * M_PI * y);
            for (int v = 0; v < 4; v++)
                u0[4 * i + v] = es.farFieldStaticValue[v] * (v == 0 || v == 3 ? s : 1.0);
        }
        auto u = ev.BuildMean();
        u.upload(u0);
        auto uRec = vfv.BuildRec(), uRecNew = vfv.BuildRec();
        uRec.setConstant(0.0);
        dndsmi::ExplicitSSPRK3LocalDt rk(ev, 2, 0.5, 1e100, false, true);
        double t = 0;
        for (int step = 0; step < 3; step++)
            t += rk.Step(u, uRec, uRecNew, t);
        dndsmi::ImplicitEulerDualTimeStep imp(ev, 1, 5.0, 4, 0.0, true);
        auto hist = imp.Step(u, uRec, uRecNew, 0.05, t);
        t += 0.05;
        dndsmi::ImplicitBDFDualTimeStep bdf(ev, 1, 5.0, 3, 0.0, true, 2);
        for (int step = 0; step < 2; step++, t += 0.05)
            bdf.Step(u, uRec, uRecNew, 0.05, t);
        dndsmi::ImplicitSDIRK4DualTimeStep sdirk(ev, 1, 5.0, 2, 0.0, true);
        sdirk.Step(u, uRec, uRecNew, 0.05, t);
        t += 0.05;
        std::array<double, 4> res{}, resInc{};
        auto rhs = ev.BuildMean();
        ev.EvaluateRHS(rhs, u, uRec, t);
        ev.EvaluateResidual(res, rhs, 1, false);
        // one LU-SGS solve through the reference's host-vector forms of
        // EvaluateDt and LUSGSMatrixInit
        std::vector<double> dTau;
        double dtMin = 0;
        ev.EvaluateDt(dTau, u, 5.0, dtMin, 1e100, true);
        auto inc = ev.BuildMean();
        inc.setConstant(0.0);
        ev.LUSGSMatrixInit(dTau, 1e100, 1.0, u, uRec, 0, t);
        ev.UpdateLUSGSForward(1.0, rhs, u, inc, inc);
        ev.UpdateLUSGSBackward(1.0, rhs, u, inc, inc);
        ev.EvaluateResidual(resInc, inc, 1, false);
        bool refused = false;
        try {
            ev.LUSGSMatrixInit(dTau, 1e100, 1.0, u, uRec, 1, t);
        } catch (const dndsmi::Error &e) {
            refused = e.code == DNDSMI_ERR_UNSUPPORTED;
        }
        if (!refused) {
            std::fprintf(stderr, "facade_run: jacobianCode 1 was not refused\n");
            return 4;
        }
        const auto hu = u.download(), hr = uRec.download();
        FILE *f = std::fopen(argv[1], "wb");
        if (!f)
            return 3;
        std::fwrite(hu.data(), sizeof(double), hu.size(), f);
        std::fwrite(hr.data(), sizeof(double), hr.size(), f);
        std::fwrite(res.data(), sizeof(double), 4, f);
        std::fwrite(resInc.data(), sizeof(double), 4, f);
        std::fwrite(dTau.data(), sizeof(double), dTau.size(), f);
        std::fclose(f);
        std::printf("facade_run: %lld cells, t = %.6f, %zu implicit iterations, |rhs|_1 = %.6e\n", (long long)n, t,
                    hist.size() / 4, res[0]);
    } catch (const dndsmi::Error &e) {
        std::fprintf(stderr, "facade_run: %s\n", e.what());
        return 1;
    }
    return 0;
}
