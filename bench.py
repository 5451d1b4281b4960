"""Benchmark: p = 3 VR + RHS cell-updates/s of the SSP-RK3 hot path on one GPU.

Workload (default): the four-quadrant 2D Riemann problem on [0,1]^2,
1024 x 1024 quads cut along seeded diagonals (2,097,152 triangles), interior
nodes jittered by U(-0.2h, 0.2h); p = 3 VR, HLLC, WBAP limiter, SSP-RK3 at
CFL 0.5 with a global time step, 5 Jacobi sweeps per stage, FP64.  One
"step" is one SSP-RK3 step: EvaluateDt once, then three stages of (5 sweeps +
WBAP + EvaluateRHS + stage update), all on the device with the state resident
in HBM.

value = cells x 3 stages x steps / wall time of the timed steps.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(the Jacobi sweep: algorithmic bytes per launch / mean launch time from HIP
events recorded by the library on its stream) and a CPU baseline (the FP64
oracle, single thread, one SSP-RK3 step on the same mesh).  The default
workload carries the WBAP limiter in every stage: V2, the solver's default
limiterProcedure 0, or V3 with --limiter-procedure 1.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


PROFILE = os.path.join(ROOT, "profiles", "riemann_2M_kernels_traffic.md")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary (None if absent)."""
    try:
        with open(PROFILE) as f:
            for line in f:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) == 7 and cols[0] == f"`{kernel}`":
                    return float(cols[4]) * 1e9, os.path.relpath(PROFILE, ROOT)
    except OSError:
        pass
    return None, None


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024, help="quads per direction")
    ap.add_argument("--sweeps", type=int, default=5)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1), help="host threads for the setup")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the oracle on one step (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=10,
                    help="after the timed steps, N device copies of one rec array (a plain copy of known size "
                         "for scaling profiler byte counters), timed: the copy_probe object of the JSON line "
                         "(0 = skip)")
    ap.add_argument("--mixed", type=int, default=1,
                    help="also time the mixed-precision step (FP32 correction sweeps + FP64 refinement) "
                         "from the same initial state and report it beside the FP64 line")
    ap.add_argument("--implicit", type=int, default=0,
                    help="also time N inner iterations of the implicit-Euler dual-time step with multicolour LU-SGS "
                         "(1 sweep, CFL 10, dt = 5 explicit steps) on the same mesh and report them")
    ap.add_argument("--gmres", type=int, default=0,
                    help="fsolve of the implicit timing: 0 LU-SGS, 1 GMRES(10) x 2 restarts preconditioned by "
                         "LU-SGS, 2 LU-SGS then GMRES")
    ap.add_argument("--warmstart", type=int, default=0,
                    help="also run the learned warm start workload: harvest converged reconstructions of the "
                         "40x40 vortex at 8 mesh seeds (160 snapshots each, 4.1M pairs), train on the device, "
                         "report the sweeps to a 1e-10 rec increment at an unseen seed")
    ap.add_argument("--limiter-procedure", type=int, default=0, choices=[0, 1],
                    help="WBAP of the limited cases: 0 V2 (the reference solver's default), 1 V3")
    ap.add_argument("--mu-gas", type=float, default=0.0,
                    help="muGas > 0 times the laminar Navier-Stokes model on the same workload (Sutherland with "
                         "TRef 1, CSutherland 0.5, Prandtl 0.72, Rgas 1; the mixed-precision line is skipped)")
    ap.add_argument("--case", default="riemann", choices=["riemann", "vortex", "step", "plate"],
                    help="riemann: four-quadrant Riemann problem; vortex: periodic isentropic vortex at the same scale; "
                         "step: Mach-3 step on a mixed tri/quad mesh of [0,4]x[0,1] (n x n quads, --n 2048 for the "
                         "6.3M-cell case); plate: laminar flat plate at M 0.2, Re 5e6, (1.5 n) x (0.3125 n) quads "
                         "stretched to y+ = 1 (491,520 at the default n), steady: with --implicit N, N LU-SGS "
                         "iterations from the free stream at CFL 20 with local time steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    from dnds_amd import _capi as A
    from dnds_amd import cases
    from dnds_amd.host import DNDSError, EulerEvaluator, ExplicitSSPRK3LocalDt

    t0 = time.time()
    log(f"building mesh + VR setup, {args.n}x{args.n} quads ({args.case})")
    if args.case == "riemann":
        case = cases.riemann_2d(args.n, nthreads=args.threads, build_limiter_data=True)
    elif args.case == "step":
        case = cases.mach3_step(args.n, nthreads=args.threads)
    elif args.case == "plate":
        case = cases.flat_plate(args.n * 3 // 2, args.n * 5 // 16, nthreads=args.threads)
    else:
        case = cases.isentropic_vortex(args.n, nthreads=args.threads)
    nc = case.n_cell
    log(f"setup done: {nc} cells, {case.mesh.n_face} faces, {time.time() - t0:.1f}s")
    gas = case.gas or dict(muGas=args.mu_gas, prGas=0.72, Rgas=1.0, TRef=1.0, CSutherland=0.5)
    if gas["muGas"] > 0:
        args.mixed = 0  # the mixed-precision step covers the Euler model only
    ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType=case.riemann, farFieldStaticValue=case.far, **gas)
    log(f"device context: {ev.device.device_bytes() / 2**30:.2f} GiB")
    urec0 = np.zeros((nc, 36))
    lib = A.load()
    import ctypes as C

    def timed_run(mixed: bool):
        """warmup + timed SSP-RK3 steps from the initial state; per-kernel
        HIP-event times over the timed steps"""
        u = ev.BuildMean().upload(case.u0)
        r, rn = case.vfv.BuildRec().upload(urec0), case.vfv.BuildRec()
        ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=args.sweeps, CFL=args.cfl, useLimiter=case.limiter,
                                    mixedPrecision=mixed, limiterProcedure=args.limiter_procedure)
        t = 0.0
        for _ in range(args.warmup):
            t += ode.Step(u, r, rn, t)
        ev.device.sync()
        tag = "mixed" if mixed else "fp64"
        log(f"[{tag}] warmup {args.warmup} steps done")
        lib.dndsmi_profile_enable(ev.device.ctx, 1)
        ev.device.sync()
        ts = time.perf_counter()
        for _ in range(args.steps):
            t += ode.Step(u, r, rn, t)
        ev.device.sync()
        elapsed = time.perf_counter() - ts
        log(f"[{tag}] timed {args.steps} steps: {elapsed:.3f}s, t = {t:.5f}")
        prof = {}
        for k, name in enumerate(A.KERNEL_NAMES):
            cnt, tot = C.c_int64(), C.c_double()
            lib.dndsmi_profile_query(ev.device.ctx, k, C.byref(cnt), C.byref(tot))
            if cnt.value:
                b = int(lib.dndsmi_kernel_bytes(ev.device.ctx, k))
                avg_ms = tot.value / cnt.value
                prof[name] = dict(launches=cnt.value, avg_ms=avg_ms, bytes=b, gbs=b / (avg_ms * 1e-3) / 1e9)
        lib.dndsmi_profile_enable(ev.device.ctx, 0)
        for name, p in prof.items():
            log(f"[{tag}] {name}: {p['launches']} launches, {p['avg_ms']:.4f} ms avg, {p['gbs']:.0f} GB/s algorithmic")
        uf = u.download()
        if not np.isfinite(uf).all():
            raise RuntimeError("non-finite state after the timed steps")
        return elapsed, prof, uf, r

    elapsed, prof, u_fp64, r = timed_run(False)
    copy = None
    if args.copy_probe:
        rn = case.vfv.BuildRec()
        nbytes = 36 * ((nc + 63) // 64 * 64) * 8
        log(f"copy probe: {args.copy_probe} device copies of {nbytes} bytes")
        rn.copy_from(r)
        ev.device.sync()
        ts = time.perf_counter()
        for _ in range(args.copy_probe):
            rn.copy_from(r)
        ev.device.sync()
        el = time.perf_counter() - ts
        # read + write of each copy, wall clock over back-to-back copies on the library's stream
        copy = {"copies": args.copy_probe, "bytes_per_copy": 2 * nbytes,
                "gbs": 2 * nbytes * args.copy_probe / el / 1e9, "frac_of_peak": 2 * nbytes * args.copy_probe / el / 1e9
                / HBM_PEAK_GBS}
        log(f"copy probe: {copy['gbs']:.0f} GB/s read + write ({copy['frac_of_peak']:.2f} of peak)")
    mixed = None
    if args.mixed:
        m_elapsed, m_prof, u_mixed, _ = timed_run(True)
        diff = float(np.abs(u_mixed - u_fp64).max() / np.abs(u_fp64).max())
        mixed = {"value": nc * 3 * args.steps / m_elapsed, "ms_per_step": m_elapsed / args.steps * 1e3,
                 "speedup_vs_fp64": elapsed / m_elapsed, "rel_diff_u_vs_fp64": diff,
                 "dtype": "f32 sweeps/face reconstruction + f64 refinement", "kernels": m_prof}
        log(f"mixed: {mixed['value']:.3e} cell-updates/s, x{mixed['speedup_vs_fp64']:.2f} vs fp64, "
            f"rel. diff of u {diff:.1e}")

    stages = 3 * args.steps
    value = nc * stages / elapsed
    jac = prof.get("k_jacobi", {})
    roof = None
    if jac:
        roof = {"bound": "hbm", "achieved": jac["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": jac["gbs"] / HBM_PEAK_GBS, "traffic": None, "kernel": "k_jacobi",
                "bytes_per_launch": jac["bytes"], "avg_ms": jac["avg_ms"]}
        if args.case == "riemann" and args.n == 1024:
            roof["traffic"], roof["traffic_source"] = measured_traffic("k_jacobi<3, false>")

    # EvaluateRHS = face kernel (inner faces; the boundary-face launch is
    # untimed and small) + gather kernel, with the bytes those kernels need
    rhs_obj = None
    if "k_rhs_face" in prof and "k_rhs_cell" in prof:
        ks = [k for k in ("k_rhs_face", "k_rhs_bface", "k_rhs_cell") if k in prof]
        ms = sum(prof[k]["avg_ms"] for k in ks)
        nbytes = sum(prof[k]["bytes"] for k in ks)
        gbs = nbytes / (ms * 1e-3) / 1e9
        # the per-unit algorithmic figure of the hot-path scope table: 1384 B per
        # face (both sides' stored row-0 bases at 3 points, normals, weights,
        # indices) and 360 B per cell (uRec, u, volume, rhs written)
        stored_bytes = 1384 * case.mesh.n_face + 360 * nc
        stored_gbs = stored_bytes / (ms * 1e-3) / 1e9
        rhs_obj = {"ms": ms, "algorithmic_bytes": nbytes, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": gbs / HBM_PEAK_GBS, "kernels": ks,
                   "stored_basis_bytes": stored_bytes, "achieved_stored_basis": stored_gbs,
                   "frac_stored_basis": stored_gbs / HBM_PEAK_GBS}
        log(f"EvaluateRHS: {ms:.3f} ms, {nbytes / 1e9:.2f} GB, {gbs:.0f} GB/s = {gbs / HBM_PEAK_GBS:.2f} of peak; "
            f"{stored_bytes / 1e9:.2f} GB with stored bases: {stored_gbs / HBM_PEAK_GBS:.2f}")

    cpu = None
    if args.cpu_baseline:
        import oracle
        st = case.vfv.setting
        orc = oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(), riemann_solver=2 if case.riemann == "HLLC" else 1,
                            far_field=case.far, tang_weight=st.tang_weight,
                            tang_weight_direction=st.tang_weight_direction,
                            wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std,
                            limiter_procedure=args.limiter_procedure, mu_gas=gas["muGas"], pr_gas=gas["prGas"],
                            r_gas=gas["Rgas"], t_ref=gas["TRef"], c_sutherland=gas["CSutherland"])
        timed = {}
        for threads in (args.threads, 1):
            oracle.set_threads(threads)
            log(f"CPU baseline: oracle on {threads} thread(s), one SSP-RK3 step on the same mesh")
            c0 = time.perf_counter()
            orc.ssprk3_steps(case.u0, urec0, 1, args.sweeps, args.cfl, use_limiter=case.limiter)
            timed[threads] = time.perf_counter() - c0
            log(f"CPU baseline {nc * 3 / timed[threads]:.3e} cell-updates/s ({timed[threads]:.1f}s)")
        cpu = {"value": nc * 3 / timed[args.threads], "unit": "cell-updates/s", "cores": args.threads, "kind": "port",
               "sample": f"one SSP-RK3 step (3 stages x {args.sweeps} sweeps{' + WBAP' if case.limiter else ''} + RHS) "
                         f"on the same {nc}-cell mesh: {timed[args.threads]:.3g} s on {args.threads} threads, "
                         f"{timed[1]:.3g} s on 1",
               "value_1_core": nc * 3 / timed[1]}

    implicit = None
    if args.implicit:
        from dnds_amd.host import ImplicitEulerDualTimeStep
        u = ev.BuildMean().upload(case.u0)
        r, rn = case.vfv.BuildRec().upload(urec0), case.vfv.BuildRec()
        dts = ev.BuildScalar()
        dt_exp = ev.EvaluateDt(dts, u, args.cfl, 1e100, False)
        ev.SetLinearSolver(args.gmres, 10, 2)
        steady = args.case == "plate"
        # steady: pseudo-time iterations from the free stream (physical dt infinite)
        imp_cfl, imp_dt = (case.cfl, 1e100) if steady else (10.0, 5 * dt_exp)
        # steady: chunks of 100 iterations (with an infinite dt, the next chunk
        # continues the same pseudo-time iteration), so that the history up to
        # a failure is kept
        chunk = min(args.implicit, 100) if steady else args.implicit
        imp = ImplicitEulerDualTimeStep(ev, nInternalRecStep=1, CFL=imp_cfl, nTimeStepInternal=chunk,
                                        rhsThresholdInternal=0.0, UseLocaldt=True)
        if not steady:
            imp.Step(u, r, rn, dt=imp_dt)  # warm-up step
        ev.device.sync()
        lib.dndsmi_profile_enable(ev.device.ctx, 1)
        ts = time.perf_counter()
        parts, failure = [], None
        while sum(len(h) for h in parts) < args.implicit:
            try:
                parts.append(imp.Step(u, r, rn, dt=imp_dt))
            except DNDSError as e:
                if not steady:
                    raise
                failure = f"after {sum(len(h) for h in parts)} iterations: {e}"
                log(f"implicit: stopped {failure}")
                break
        hist = np.concatenate(parts)
        ev.device.sync()
        el = time.perf_counter() - ts
        iprof = {}
        for k, name in enumerate(A.KERNEL_NAMES):
            cnt, tot = C.c_int64(), C.c_double()
            lib.dndsmi_profile_query(ev.device.ctx, k, C.byref(cnt), C.byref(tot))
            if cnt.value:
                b = int(lib.dndsmi_kernel_bytes(ev.device.ctx, k))
                if name == "k_lusgs_sweep":  # bytes are per whole sweep (one launch per colour)
                    sweeps = cnt.value // max(ev.LUSGSColouring()[1], 1)
                    iprof[name] = dict(launches=cnt.value, sweeps=sweeps, ms_per_sweep=tot.value / sweeps, bytes=b,
                                       gbs=b * sweeps / (tot.value * 1e-3) / 1e9)
                else:
                    avg = tot.value / cnt.value
                    iprof[name] = dict(launches=cnt.value, avg_ms=avg, bytes=b, gbs=b / (avg * 1e-3) / 1e9 if b else 0)
        lib.dndsmi_profile_enable(ev.device.ctx, 0)
        col, ncol = ev.LUSGSColouring()
        implicit = {"linear_solver": ["LU-SGS", "GMRES(10) x 2 restarts, LU-SGS preconditioned",
                                      "LU-SGS then GMRES(10) x 2 restarts"][args.gmres],
                    "inner_iterations": len(hist), "ms_per_iteration": el / len(hist) * 1e3,
                    "cell_iterations_per_s": nc * len(hist) / el, "colours": ncol,
                    "increment_drop": float(hist[-1, 0] / hist[0, 0]), "cfl": imp_cfl,
                    "dt": "steady (infinite)" if steady else imp_dt, "kernels": iprof}
        if steady:
            implicit["increment_history"] = {k: float(hist[k, 0]) for k in
                                             sorted({0, *range(0, len(hist), max(1, len(hist) // 20)), len(hist) - 1})}
            implicit["failure"] = failure
        lus = iprof.get("k_lusgs_sweep", {})
        log(f"implicit: {implicit['ms_per_iteration']:.3f} ms per inner iteration, {ncol} colours, "
            f"LU-SGS {lus.get('ms_per_sweep', 0):.3f} ms per sweep, {lus.get('gbs', 0):.0f} GB/s algorithmic")

    warm = None
    if args.warmstart:
        from dnds_amd import learned
        w0 = time.time()
        X, T = learned.harvest(range(1, 9), n=40, snapshots=160, steps_between=1, nthreads=args.threads, log=log)
        w1 = time.time()
        log(f"warm start: {X.shape[0]} training pairs harvested in {w1 - w0:.1f}s")
        params, hist = learned.train(X, T, epochs=12, batch=4096, lr=2e-3, lr_final=1e-5, seed=0, n=40,
                                     nthreads=args.threads)
        w2 = time.time()
        log(f"warm start: trained 12 epochs in {w2 - w1:.1f}s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
        ev_ws = learned.evaluate(params, seed=9, n=40, snapshots=20, steps_between=5, tol=1e-10,
                                 nthreads=args.threads)
        log(f"warm start: {ev_ws}")
        warm = {"pairs": int(X.shape[0]), "train_s": w2 - w1, "harvest_s": w1 - w0,
                "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1]),
                "eval_seed": 9, "tolerance": 1e-10, **ev_ws}

    wbap = "V2" if args.limiter_procedure == 0 else "V3"
    out = {
        "metric": "p=3 VR+RHS cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh, analytic initial state)",
        "config": {"workload": {
                       "riemann": (f"2D Riemann, {args.n}x{args.n} quads split to {nc} triangles, jitter 0.2h, "
                                   f"p=3 VR, HLLC, WBAP {wbap} limiter, SSP-RK3 CFL {args.cfl}, {args.sweeps} Jacobi "
                                   "sweeps/stage"),
                       "vortex": (f"periodic isentropic vortex, {args.n}x{args.n} quads split to {nc} triangles, "
                                  f"p=3 VR, Roe, SSP-RK3 CFL {args.cfl}, {args.sweeps} Jacobi sweeps/stage, no limiter"),
                       "step": (f"Mach-3 step, [0,4]x[0,1], {args.n}x{args.n} quads half split: {nc} mixed cells, "
                                f"jitter 0.15h, p=3 VR, HLLC, WBAP {wbap}, SSP-RK3 CFL {args.cfl}, {args.sweeps} Jacobi "
                                "sweeps/stage"),
                       "plate": (f"laminar flat plate, M 0.2, Re 5e6, {nc} quads stretched to y+ = 1, p=3 VR, Roe + "
                                 f"viscous, SSP-RK3 CFL {args.cfl}, {args.sweeps} Jacobi sweeps/stage")}[args.case],
                   "cells": nc, "sweeps_per_stage": args.sweeps,
                   "limiter": f"WBAP {wbap}" if case.limiter else "none", "precision": "fp64",
                   "model": "laminar Navier-Stokes, muGas %g" % gas["muGas"] if gas["muGas"] > 0 else "Euler"},
        "kernels": prof,
        "roofline": roof,
        "evaluate_rhs": rhs_obj,
        "cpu_baseline": cpu,
        "copy_probe": copy,
        "mixed_precision": mixed,
        "implicit_lusgs": implicit,
        "learned_warm_start": warm,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
