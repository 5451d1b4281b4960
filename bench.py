"""Benchmark: p = 3 VR + RHS cell-updates/s of the SSP-RK3 hot path on one GPU,
plain FP64 against FP32 + FP64 refinement at equal L2 error.

Workload (default): the combined large case -- [0,4] x [0,1], 2048 x 2048
quads of which half are cut into triangles (seeded): 6,293,459 mixed cells,
interior nodes jittered by U(-0.15h, 0.15h); a Mach-3 shock at x = 0.5 with
seeded density noise of 1e-3; p = 3 VR, HLLC, WBAP V2 (the reference
solver's default limiterProcedure 0), SSP-RK3 at CFL 0.5 with a global time
step, 5 Jacobi sweeps per stage, 100 steps.  One "step" is one SSP-RK3 step:
EvaluateDt once, then three stages of (5 sweeps + WBAP + EvaluateRHS + stage
update), all on the device with the state resident in HBM.

value = cells x 3 stages x steps / wall time of the timed FP64 steps (the
plain arm).  The mixed arm (first sweep of each stage in FP64, the others as
FP32 corrections; FP32 face reconstruction and flux storage) runs the same
steps from the same initial state; the line reports its throughput, the
ratio, and the L2 difference of its state from the FP64 arm's after the
same steps.  Both arms are also compared with the oracle after one step from
the initial state (the CPU-baseline run), and the accuracy object gives both
arms' L2 error against the exact solution on the small isentropic-vortex
case.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(the Jacobi sweep: algorithmic bytes per launch / mean launch time from HIP
events recorded by the library on its stream), EvaluateRHS's fraction of the
HBM peak, a CPU baseline (the oracle compiled -O3 -march=native on this host,
one SSP-RK3 step of the same workload on the host's cores) and the
four-quadrant Riemann case at its 200 steps.
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

# the learned warm start's plain check: steps per check, steps between a
# passed check and the next, tolerance on the relative L2 difference from
# the plain run (1 % of the vortex case's L2 discretisation error, 2.5e-5)
LEARNED_CHECK = (5, 20, 2.5e-7)

# committed rocprofv3 summaries (kernel trace, request-size traffic), by workload
PROFILES = {("riemann", 1024): os.path.join(ROOT, "profiles", "riemann_2M_kernels_traffic.md"),
            ("step", 2048): os.path.join(ROOT, "profiles", "step_6M_kernels_traffic.md")}


def measured_traffic(case: str, n: int, kernel: str):
    """L2-to-fabric bytes per launch of `kernel` on a workload (an upper
    bound of its HBM bytes), from the committed rocprofv3 request-size
    summary (None if absent)."""
    path = PROFILES.get((case, n))
    try:
        with open(path) as f:
            for line in f:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) >= 7 and cols[0] == f"`{kernel}`":
                    return float(cols[4]) * 1e9, os.path.relpath(path, ROOT)
    except (OSError, TypeError):
        pass
    return None, None


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_l2(a, b, vol):
    """Per variable, sqrt(sum V (a - b)^2 / sum V b^2): the volume-weighted L2
    difference of two cell-mean states relative to b."""
    w = vol[:, None]
    return (np.sqrt((w * (a - b) ** 2).sum(0) / (w * b ** 2).sum(0))).tolist()


# the reference's RiemannSolverType codes, for the oracle
RS_CODE = {"Roe": 1, "HLLC": 2, "HLLEP": 3, "Roe_M1": 11, "Roe_M2": 12, "Roe_M3": 13, "Roe_M4": 14, "Roe_M5": 15}


def make_case(name, n, threads, rec_order=3):
    from dnds_amd import cases
    if name == "riemann":
        return cases.riemann_2d(n, nthreads=threads, build_limiter_data=True)
    if name == "step":
        return cases.mach3_step(n, nthreads=threads)
    if name == "plate":
        return cases.flat_plate(n * 3 // 2, n * 5 // 16, rec_order=rec_order, nthreads=threads)
    return cases.isentropic_vortex(n, nthreads=threads)


def kernel_profile(lib, ctx, A):
    import ctypes as C
    prof = {}
    for k, name in enumerate(A.KERNEL_NAMES):
        cnt, tot = C.c_int64(), C.c_double()
        lib.dndsmi_profile_query(ctx, k, C.byref(cnt), C.byref(tot))
        if cnt.value:
            b = int(lib.dndsmi_kernel_bytes(ctx, k))
            avg_ms = tot.value / cnt.value
            prof[name] = dict(launches=cnt.value, avg_ms=avg_ms, bytes=b, gbs=b / (avg_ms * 1e-3) / 1e9)
    return prof


class Arms:
    """Runs the plain FP64 and the mixed-precision SSP-RK3 steps of one case
    from its initial state on the device."""

    def __init__(self, case, args, gas):
        from dnds_amd import _capi as A
        from dnds_amd.host import EulerEvaluator
        self.A, self.case, self.args = A, case, args
        self.ev = EulerEvaluator(case.mesh, case.vfv, riemannSolverType=case.riemann, farFieldStaticValue=case.far,
                                 **gas)
        self.lib = A.load()
        self.urec0 = np.zeros((case.n_cell, 36))

    def fresh(self):
        c = self.case
        return self.ev.BuildMean().upload(c.u0), c.vfv.BuildRec().upload(self.urec0), c.vfv.BuildRec()

    def ode(self, mixed, ws=None):
        from dnds_amd.host import ExplicitSSPRK3LocalDt
        self.last_ode = ExplicitSSPRK3LocalDt(self.ev, nInternalRecStep=self.args.sweeps, CFL=self.args.cfl,
                                              useLimiter=self.case.limiter, mixedPrecision=mixed,
                                              limiterProcedure=self.args.limiter_procedure, warmStart=ws,
                                              checkSteps=LEARNED_CHECK[0], recheckInterval=LEARNED_CHECK[1],
                                              tolerance=LEARNED_CHECK[2])
        return self.last_ode

    def warm_start(self, params):
        """the trained network on this workload's context"""
        from dnds_amd.host import WarmStart
        ws = WarmStart(self.ev, seed=0, capacity=0)
        ws.set_params(params)
        return ws

    def steps(self, mixed, n, ws=None):
        """n steps from the initial state, untimed: (u, uRec, t)."""
        u, r, rn = self.fresh()
        ode, t = self.ode(mixed, ws), 0.0
        for _ in range(n):
            t += ode.Step(u, r, rn, t)
        return u.download(), r.download(), t

    def timed(self, mixed, steps, warmup, ws=None, until_failure=False):
        """warmup steps, then `steps` timed steps continuing from them; the
        per-kernel HIP-event times of the timed steps.  until_failure: a step
        that fails (a non-positive state) ends the run; self.completed holds
        the steps before it, and the time covers them"""
        from dnds_amd.host import DNDSError
        tag = ("mixed+learned" if ws is not None else "mixed") if mixed else "fp64"
        u, r, rn = self.fresh()
        ode, t = self.ode(mixed, ws), 0.0
        for _ in range(warmup):
            t += ode.Step(u, r, rn, t)
        self.ev.device.sync()
        self.lib.dndsmi_profile_enable(self.ev.device.ctx, 1)
        self.ev.device.sync()
        self.completed, self.failure = steps, None
        ts = time.perf_counter()
        for k in range(steps):
            try:
                t += ode.Step(u, r, rn, t)
            except DNDSError as e:
                if not until_failure:
                    raise
                # the failure surfaces when step k reads the time step back:
                # the state left by step k - 1 is not admissible
                self.completed, self.failure = k - 1, f"step {k}: {e}"
                break
        self.ev.device.sync()
        elapsed = time.perf_counter() - ts
        log(f"[{self.case.name} {tag}] timed {self.completed} steps after {warmup}: {elapsed:.3f}s, t = {t:.6f}"
            + (f" (stopped: {self.failure})" if self.failure else ""))
        prof = kernel_profile(self.lib, self.ev.device.ctx, self.A)
        self.lib.dndsmi_profile_enable(self.ev.device.ctx, 0)
        for name, p in prof.items():
            log(f"[{self.case.name} {tag}] {name}: {p['launches']} launches, {p['avg_ms']:.4f} ms avg, "
                f"{p['gbs']:.0f} GB/s algorithmic")
        uf = u.download()
        if not np.isfinite(uf).all():
            raise RuntimeError("non-finite state after the timed steps")
        return elapsed, prof, uf, r, t


def evaluate_rhs_object(prof, case):
    """EvaluateRHS = face kernel (inner faces) + boundary-face kernel + gather
    kernel, with the bytes those kernels need"""
    if "k_rhs_face" not in prof or "k_rhs_cell" not in prof:
        return None
    ks = [k for k in ("k_rhs_face", "k_rhs_bface", "k_rhs_cell") if k in prof]
    ms = sum(prof[k]["avg_ms"] for k in ks)
    nbytes = sum(prof[k]["bytes"] for k in ks)
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"EvaluateRHS: {ms:.3f} ms, {nbytes / 1e9:.2f} GB, {gbs:.0f} GB/s = {gbs / HBM_PEAK_GBS:.2f} of peak")
    return {"ms": ms, "algorithmic_bytes": nbytes, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "kernels": ks}


def vortex_accuracy(args, params=None):
    """The small isentropic-vortex case (40 x 40 quads split into 3200
    triangles, Roe, no limiter, 5 sweeps, CFL 0.5, 100 steps): density L2
    error against the exact solution for the FP64 arm, the mixed arm and
    (with trained parameters) the mixed arm with the learned warm start."""
    from dnds_amd import cases
    case = cases.isentropic_vortex(40, nthreads=args.threads)
    arms = Arms(case, argparse.Namespace(sweeps=5, cfl=0.5, limiter_procedure=0), {})
    vol = case.vfv.arrays(["volume"])["volume"]
    out = {"workload": "periodic isentropic vortex, 40x40 quads split to 3200 triangles, Roe, no limiter, "
                       "5 sweeps/stage, CFL 0.5, 100 steps"}
    runs = [("fp64", False, None), ("mixed", True, None)]
    if params is not None:
        runs.append(("mixed_learned", True, arms.warm_start(params)))
    for tag, mixed, ws in runs:
        u, _, t = arms.steps(mixed, case.steps, ws)
        ex = cases.cell_means(case.mesh, vol, lambda x, y: cases.vortex_exact(x, y, t))
        out[f"l2_density_error_{tag}"] = float(np.sqrt(np.sum((u[:, 0] - ex[:, 0]) ** 2 * vol) / vol.sum()))
        if ws is not None:
            out["warm_start_stats"] = arms.last_ode.WarmStartStats()
    out["mixed_over_fp64"] = out["l2_density_error_mixed"] / out["l2_density_error_fp64"]
    if params is not None:
        out["mixed_learned_over_fp64"] = out["l2_density_error_mixed_learned"] / out["l2_density_error_fp64"]
    log(f"vortex accuracy: {out}")
    return out


def train_warm_start(args):
    """The learned warm start's network: converged reconstructions of the
    40 x 40 vortex at mesh seeds 1..8 (`args.learned_snapshots` snapshots
    each, one step apart), 12 epochs on the device; (parameters, summary)."""
    from dnds_amd import learned
    w0 = time.time()
    X, T = learned.harvest(range(1, 9), n=40, snapshots=args.learned_snapshots, steps_between=1,
                           nthreads=args.threads, log=log)
    w1 = time.time()
    params, hist = learned.train(X, T, epochs=12, batch=4096, lr=2e-3, lr_final=1e-5, seed=0, n=40,
                                 nthreads=args.threads)
    w2 = time.time()
    log(f"warm start: {X.shape[0]} pairs harvested in {w1 - w0:.1f}s, trained 12 epochs in {w2 - w1:.1f}s, "
        f"loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    return params, {"pairs": int(X.shape[0]), "harvest_s": w1 - w0, "train_s": w2 - w1,
                    "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1]),
                    "training_data": "converged reconstructions of the 40x40 isentropic vortex, mesh seeds 1-8"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="step", choices=["step", "riemann", "vortex", "plate"],
                    help="step: the combined large case, Mach-3 shock on a mixed tri/quad mesh of [0,4]x[0,1] "
                         "(n x n quads, half split; 6.3M cells at n 2048); riemann: four-quadrant Riemann problem "
                         "on [0,1]^2 (2 n^2 triangles); vortex: periodic isentropic vortex on [0,10]^2; plate: laminar "
                         "flat plate at M 0.2, Re 1e5, (1.5 n) x (0.3125 n) quads stretched to y+ = 1, the reference's "
                         "boundary-layer VR settings, steady: with --implicit N, N LU-SGS iterations from the free "
                         "stream at CFL 20 with local time steps")
    ap.add_argument("--n", type=int, default=None, help="quads per direction (default 2048 for step, 1024 else)")
    ap.add_argument("--sweeps", type=int, default=5)
    ap.add_argument("--rec-order", type=int, default=3, choices=[1, 3],
                    help="P_ORDER of the plate case (the other cases run p = 3)")
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads for the setup and the CPU baseline")
    ap.add_argument("--cpu-baseline", type=int, default=1,
                    help="time the oracle (built -O3 -march=native here) on one step of the workload (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=10,
                    help="after the timed steps, N device copies of one rec array (a plain copy of known size "
                         "for scaling profiler byte counters), timed (0 = skip)")
    ap.add_argument("--read-probe", type=int, default=10,
                    help="after the timed steps, N launches of a kernel that streams the sweep's A^-1 B pairs as the "
                         "sweep does, without its gathers: the read ceiling of the sweep's streaming part (0 = skip)")
    ap.add_argument("--mixed", type=int, default=1,
                    help="also run the mixed-precision arm (FP32 correction sweeps + FP64 refinement) over the same "
                         "steps from the same initial state")
    ap.add_argument("--learned", type=int, default=1,
                    help="also train the learned warm start (vortex data, on the device) and run the mixed arm with "
                         "it and its plain fallback over the same steps")
    ap.add_argument("--learned-snapshots", type=int, default=40,
                    help="training snapshots per mesh seed (8 seeds of the 40x40 vortex, 3200 pairs each)")
    ap.add_argument("--accuracy", type=int, default=1,
                    help="both arms' L2 error against the exact solution on the 3200-cell isentropic vortex")
    ap.add_argument("--riemann-200", type=int, default=1,
                    help="also run the four-quadrant Riemann case (1024x1024 quads, 2.1M triangles) for its 200 "
                         "steps in both arms and report it in the riemann_2M object")
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
    ap.add_argument("--sor", type=int, default=2000,
                    help="also reconstruct the workload's initial state from zero to a 1e-10 rec increment (at most "
                         "N sweeps) with the Jacobi sweep and the SOR_Instead sweep, and report sweeps and times "
                         "(0 = skip)")
    ap.add_argument("--sor-inverse", type=int, default=0,
                    help="with --sor, also SOR_Instead with SOR_InverseScanning")
    ap.add_argument("--limiter-procedure", type=int, default=0, choices=[0, 1],
                    help="WBAP of the limited cases: 0 V2 (the reference solver's default), 1 V3")
    ap.add_argument("--mu-gas", type=float, default=0.0,
                    help="muGas > 0 times the laminar Navier-Stokes model on the same workload (Sutherland with "
                         "TRef 1, CSutherland 0.5, Prandtl 0.72, Rgas 1; the mixed-precision arm is skipped)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.n is None:
        args.n = 2048 if args.case == "step" else 1024

    from dnds_amd.host import DNDSError

    t0 = time.time()
    log(f"building mesh + VR setup, {args.n}x{args.n} quads ({args.case})")
    case = make_case(args.case, args.n, args.threads, args.rec_order)
    nc = case.n_cell
    log(f"setup done: {nc} cells, {case.mesh.n_face} faces, {time.time() - t0:.1f}s")
    gas = case.gas or dict(muGas=args.mu_gas, prGas=0.72, Rgas=1.0, TRef=1.0, CSutherland=0.5)
    if gas["muGas"] > 0:
        args.mixed = 0  # the mixed-precision step covers the Euler model only
    arms = Arms(case, args, gas)
    ev, lib = arms.ev, arms.lib
    log(f"device context: {ev.device.device_bytes() / 2**30:.2f} GiB")
    vol = case.vfv.arrays(["volume"])["volume"]

    # one step of each arm from the initial state, for the comparison with the oracle
    one = {"fp64": arms.steps(False, 1)}
    if args.mixed:
        one["mixed"] = arms.steps(True, 1)

    elapsed, prof, u_fp64, r, t_end = arms.timed(False, args.steps, args.warmup)
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
    probe = None
    if args.read_probe:
        ev.device.sync()
        lib.dndsmi_profile_enable(ev.device.ctx, 1)
        lib.dndsmi_read_probe(ev.device.ctx, args.read_probe)
        pp = kernel_profile(lib, ev.device.ctx, arms.A).get("k_read_probe")
        lib.dndsmi_profile_enable(ev.device.ctx, 0)
        if pp:
            probe = {"launches": pp["launches"], "avg_ms": pp["avg_ms"], "bytes": pp["bytes"], "gbs": pp["gbs"],
                     "frac_of_peak": pp["gbs"] / HBM_PEAK_GBS,
                     "what": "16-byte non-temporal loads of the sweep's A^-1 B pairs, as the sweep issues them, "
                             "one double written per cell"}
            log(f"read probe: {pp['gbs']:.0f} GB/s ({pp['gbs'] / HBM_PEAK_GBS:.2f} of peak)")
    mixed = None
    if args.mixed:
        m_elapsed, m_prof, u_mixed, _, m_t = arms.timed(True, args.steps, args.warmup)
        if m_t != t_end:
            log(f"note: the arms reached different times ({t_end} vs {m_t})")
        l2 = rel_l2(u_mixed, u_fp64, vol)
        mixed = {"value": nc * 3 * args.steps / m_elapsed, "ms_per_step": m_elapsed / args.steps * 1e3,
                 "speedup_vs_fp64": elapsed / m_elapsed, "rel_l2_vs_fp64_arm": l2,
                 "rel_max_vs_fp64_arm": float(np.abs(u_mixed - u_fp64).max() / np.abs(u_fp64).max()),
                 "dtype": "f32 sweeps/face reconstruction + f64 refinement", "learned_warm_start": None,
                 "kernels": m_prof, "evaluate_rhs": evaluate_rhs_object(m_prof, case)}
        log(f"mixed: {mixed['value']:.3e} cell-updates/s, x{mixed['speedup_vs_fp64']:.3f} vs fp64, "
            f"rel. L2 vs the fp64 arm after {args.warmup + args.steps} steps {l2}")
    params = None
    if args.mixed and args.learned:
        # the network is trained on the vortex and carried to this workload
        # (its inputs are scale-free stencil differences and cell moments)
        params, train_info = train_warm_start(args)
        ws = arms.warm_start(params)
        l_elapsed, l_prof, u_learned, _, _ = arms.timed(True, args.steps, args.warmup, ws)
        stats = arms.last_ode.WarmStartStats()
        l2l = rel_l2(u_learned, u_fp64, vol)
        mixed["learned_warm_start"] = {
            "value": nc * 3 * args.steps / l_elapsed, "ms_per_step": l_elapsed / args.steps * 1e3,
            "speedup_vs_fp64": elapsed / l_elapsed, "rel_l2_vs_fp64_arm": l2l,
            "rel_max_vs_fp64_arm": float(np.abs(u_learned - u_fp64).max() / np.abs(u_fp64).max()),
            "fallback": (f"the first {LEARNED_CHECK[0]} steps, and {LEARNED_CHECK[0]} steps every "
                         f"{LEARNED_CHECK[1]} after a passed check, run plain on shadow copies too; a corrected state "
                         f"more than {LEARNED_CHECK[2]:g} (relative L2) from the plain one, or a failed corrected "
                         "run, puts the plain state back and ends the correction"),
            "stats_warmup_and_timed": stats, "training": train_info}
        log(f"mixed + learned: {mixed['learned_warm_start']['value']:.3e} cell-updates/s, "
            f"x{mixed['learned_warm_start']['speedup_vs_fp64']:.3f} vs fp64, stats {stats}, rel. L2 vs fp64 {l2l}")
        del ws

    stages = 3 * args.steps
    value = nc * stages / elapsed
    jac = prof.get("k_jacobi", {})
    roof = None
    if jac:
        roof = {"bound": "hbm", "achieved": jac["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": jac["gbs"] / HBM_PEAK_GBS, "traffic": None, "kernel": "k_jacobi",
                "bytes_per_launch": jac["bytes"], "avg_ms": jac["avg_ms"]}
        kname = "k_jacobi<4, false>" if args.case == "step" else "k_jacobi<3, false>"
        roof["traffic"], roof["traffic_source"] = measured_traffic(args.case, args.n, kname)
        if probe:
            roof["algorithmic_rate_over_read_probe"] = jac["gbs"] / probe["gbs"]
            if roof["traffic"]:
                roof["traffic_rate_over_read_probe"] = roof["traffic"] / (jac["avg_ms"] * 1e-3) / 1e9 / probe["gbs"]
    rhs_obj = evaluate_rhs_object(prof, case)

    cpu = None
    oracle_cmp = None
    orc = None
    if args.cpu_baseline:
        import oracle
        log("CPU baseline: building the oracle -O3 -march=native on this host")
        oracle.build_baseline()
        st = case.vfv.setting

        def make_oracle(variant):
            return oracle.Oracle(case.mesh.arrays(), case.vfv.arrays(),
                                 riemann_solver=RS_CODE[case.riemann], far_field=case.far,
                                 tang_weight=st.tang_weight, tang_weight_direction=st.tang_weight_direction,
                                 wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale, wbap_n_std=st.wbap_n_std,
                                 limiter_procedure=args.limiter_procedure, mu_gas=gas["muGas"], pr_gas=gas["prGas"],
                                 r_gas=gas["Rgas"], t_ref=gas["TRef"], c_sutherland=gas["CSutherland"],
                                 variant=variant)
        orc = make_oracle("baseline")
        oracle.set_threads(args.threads, variant="baseline")
        log(f"CPU baseline: oracle on {args.threads} threads, one SSP-RK3 step on the same mesh")
        c0 = time.perf_counter()
        ou, _, ot = orc.ssprk3_steps(case.u0, arms.urec0, 1, args.sweeps, args.cfl, use_limiter=case.limiter)
        el = time.perf_counter() - c0
        log(f"CPU baseline {nc * 3 / el:.3e} cell-updates/s ({el:.1f}s)")
        cpu = {"value": nc * 3 / el, "unit": "cell-updates/s", "cores": args.threads, "kind": "port",
               "sample": f"one SSP-RK3 step (3 stages x {args.sweeps} sweeps{' + WBAP' if case.limiter else ''} + RHS) "
                         f"of the same {nc}-cell workload from its initial state: {el:.3g} s on {args.threads} "
                         "threads",
               "build": "g++ " + " ".join(oracle.BASELINE_FLAGS)}
        # the parity build (every product and sum rounded as written, as on
        # the device) is the checker; the speed build's contractions differ
        # within rounding, which the limiter's switches can amplify
        log("oracle (parity build): one SSP-RK3 step for the comparison")
        orc = make_oracle("parity")
        oracle.set_threads(args.threads)
        pu, _, pt = orc.ssprk3_steps(case.u0, arms.urec0, 1, args.sweeps, args.cfl, use_limiter=case.limiter)
        oracle_cmp = {"steps": 1, "oracle_build": "parity (-O2 -ffp-contract=off)", "t_oracle": pt,
                      "t_gpu": one["fp64"][2]}
        for tag, (gu, _, _) in one.items():
            oracle_cmp[f"rel_l2_{tag}"] = rel_l2(gu, pu, vol)
            oracle_cmp[f"rel_max_{tag}"] = float(np.abs(gu - pu).max() / np.abs(pu).max())
        oracle_cmp["speed_build_rel_max_vs_parity_build"] = float(np.abs(ou - pu).max() / np.abs(pu).max())
        log(f"vs oracle after one step: {oracle_cmp}")

    accuracy = vortex_accuracy(args, params) if args.accuracy else None

    implicit = None
    if args.implicit:
        import ctypes as C

        from dnds_amd import _capi as A
        from dnds_amd.host import ImplicitEulerDualTimeStep
        u = ev.BuildMean().upload(case.u0)
        r, rn = case.vfv.BuildRec().upload(arms.urec0), case.vfv.BuildRec()
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

    sor = None
    if args.sor:
        import ctypes as C
        from dnds_amd import _capi as A
        from dnds_amd.host import EulerEvaluator, VRFiniteVolume2D
        sor = {"tolerance": 1e-10, "max_sweeps": args.sor,
               "start": "the workload's initial means, uRec from zero, t = 0",
               "note": "the SOR sweeps run colour by colour in place; each sweep of this loop first copies uRec "
                       "(the loop measures the increment against the previous iterate)"}
        kj = A.KERNEL_NAMES.index("k_jacobi")
        variants = [("jacobi", (0, 0)), ("sor", (1, 0))] + ([("sor_inverse", (1, 1))] if args.sor_inverse else [])
        for name, (on, inverse) in variants:
            if on:
                st = A.VRSetting.from_buffer_copy(case.vfv.setting)
                st.sor_instead, st.sor_inverse_scanning = on, inverse
                s0 = time.time()
                vfv2 = VRFiniteVolume2D(case.mesh, st)
                vfv2.Initialization(args.threads)
                ev2 = EulerEvaluator(case.mesh, vfv2, riemannSolverType=case.riemann, farFieldStaticValue=case.far,
                                     **gas)
                log(f"sor: {name} setup {time.time() - s0:.1f}s")
            else:
                vfv2, ev2 = case.vfv, ev
            u = ev2.BuildMean().upload(case.u0)
            r, rn = vfv2.BuildRec().setConstant(0.0), vfv2.BuildRec()
            ctx = ev2.device.ctx
            lib.dndsmi_profile_enable(ctx, 1)
            ev2.device.sync()
            ts = time.perf_counter()
            n, res = vfv2.ReconstructToTolerance(u, r, rn, 1e-10, args.sor, 0.0)
            ev2.device.sync()
            el = time.perf_counter() - ts
            cnt, tot = C.c_int64(), C.c_double()
            lib.dndsmi_profile_query(ctx, kj, C.byref(cnt), C.byref(tot))
            lib.dndsmi_profile_enable(ctx, 0)
            sor[name] = {"sweeps": n, "last_increment": res, "converged": res < 1e-10, "s": el,
                         "ms_per_sweep": el / n * 1e3, "kernel_ms_per_sweep": tot.value / n,
                         "launches_per_sweep": cnt.value / n}
            log(f"sor: {name}: {sor[name]}")
            del u, r, rn
            if on:
                del ev2, vfv2

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

    riemann = None
    if args.riemann_200 and args.case != "riemann":
        orc = None
        del arms, ev  # frees the workload's device context before the next one
        log("riemann_2M: building mesh + VR setup, 1024x1024 quads")
        rcase = make_case("riemann", 1024, args.threads)
        rargs = argparse.Namespace(**vars(args))
        rarms = Arms(rcase, rargs, dict(muGas=0.0, prGas=0.72, Rgas=1.0, TRef=1.0, CSutherland=0.5))
        riemann = {"workload": f"2D Riemann, 1024x1024 quads split to {rcase.n_cell} triangles, jitter 0.2h, p=3 VR, "
                               f"HLLC, WBAP V2, SSP-RK3 CFL {args.cfl}, {args.sweeps} sweeps/stage, "
                               f"{rcase.steps} steps from the initial state", "cells": rcase.n_cell,
                   "steps": rcase.steps}
        # the step that meets a non-positive cell mean ends the run (the
        # scheme aborts there, as the reference's DNDS_assert would)
        re, rprof, ru, _, _ = rarms.timed(False, rcase.steps, 0, until_failure=True)
        n_ok = rarms.completed
        riemann.update({"steps_completed_fp64": n_ok, "failure_fp64": rarms.failure,
                        "value_fp64": rcase.n_cell * 3 * n_ok / re, "ms_per_step_fp64": re / max(n_ok, 1) * 1e3,
                        "jacobi_frac": rprof["k_jacobi"]["gbs"] / HBM_PEAK_GBS,
                        "evaluate_rhs": evaluate_rhs_object(rprof, rcase)})
        if args.mixed:
            me, mprof, mu, _, _ = rarms.timed(True, n_ok, 0, until_failure=True)
            rvol = rcase.vfv.arrays(["volume"])["volume"]
            riemann.update({"steps_mixed": rarms.completed, "value_mixed": rcase.n_cell * 3 * rarms.completed / me,
                            "ms_per_step_mixed": me / max(rarms.completed, 1) * 1e3,
                            "speedup": (re / max(n_ok, 1)) / (me / max(rarms.completed, 1)),
                            "rel_l2_mixed_vs_fp64": rel_l2(mu, ru, rvol) if rarms.completed == n_ok else None})
        log(f"riemann_2M: {riemann}")

    wbap = "V2" if args.limiter_procedure == 0 else "V3"
    out = {
        "metric": "p=3 VR+RHS cell-updates/s; plain FP64 vs FP32+learned at equal L2 error",
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
                       "plate": (f"laminar flat plate, M 0.2, Re 1e5, {nc} quads stretched to y+ = 1, "
                                 f"p={args.rec_order} VR with the reference's boundary-layer settings, Roe + viscous, "
                                 f"SSP-RK3 CFL {args.cfl}, {args.sweeps} Jacobi sweeps/stage")}[args.case],
                   "cells": nc, "sweeps_per_stage": args.sweeps,
                   "limiter": f"WBAP {wbap}" if case.limiter else "none", "precision": "fp64",
                   "model": "laminar Navier-Stokes, muGas %g" % gas["muGas"] if gas["muGas"] > 0 else "Euler"},
        "kernels": prof,
        "roofline": roof,
        "evaluate_rhs": rhs_obj,
        "cpu_baseline": cpu,
        "copy_probe": copy,
        "read_probe": probe,
        "mixed_precision": mixed,
        "oracle_comparison": oracle_cmp,
        "accuracy_vortex": accuracy,
        "riemann_2M": riemann,
        "implicit_lusgs": implicit,
        "learned_warm_start": warm,
        "sor_study": sor,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
