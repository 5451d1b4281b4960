"""The learned warm start, end to end: harvest converged reconstructions from
vortex runs, train the network on the device, and measure the Jacobi sweeps
it saves to reach a reconstruction tolerance.

Workload (the periodic isentropic vortex on [0,10]^2, n x n quads cut along
seeded diagonals, Roe, SSP-RK3 at CFL 0.5 with 5 sweeps per stage): at every
snapshot the reconstruction of the current state is converged with plain
FP64 Jacobi to `harvest_tol` and the (features, coefficients) pairs of all
cells are kept.  Evaluation is on a mesh seed not used for training: at each snapshot,
plain Jacobi runs to `tol` (the reference's rec increment,
sum |uRec - uRecNew|^2 / n_cell) from four starting points and the sweeps
are counted:
  zero;
  the network's guess from the current means;
  the previous snapshot's converged reconstruction (what a time-stepping
  solver already has);
  that reconstruction plus the network applied to the change of the means
  since then.  The converged reconstruction of the periodic case is linear
  in the means and the network's inputs are scale-free, so the same network
  predicts the change of the reconstruction.
"""
from __future__ import annotations

import time

import numpy as np

from . import cases
from .host import EulerEvaluator, ExplicitSSPRK3LocalDt, WarmStart


def _vortex(n, seed, nthreads):
    c = cases.isentropic_vortex(n, seed=seed, nthreads=nthreads)
    ev = EulerEvaluator(c.mesh, c.vfv, riemannSolverType="Roe", farFieldStaticValue=c.far)
    return c, ev


def harvest(seeds, n=40, snapshots=160, steps_between=1, harvest_tol=1e-24, max_sweeps=600, nthreads=0,
            log=None):
    """Stacked (X [N][58], T [N][36]) over seeds and snapshots."""
    Xs, Ts = [], []
    for seed in seeds:
        t0 = time.time()
        c, ev = _vortex(n, seed, nthreads)
        ws = WarmStart(ev, seed=seed, capacity=0)
        u, r, rn = ev.BuildMean().upload(c.u0), c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
        ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=5, CFL=0.5)
        t = 0.0
        for k in range(snapshots):
            c.vfv.ReconstructToTolerance(u, r, rn, harvest_tol, max_sweeps, t)
            X, _ = ws.features(u)
            Xs.append(X)
            Ts.append(ws.targets(u, r))
            for _ in range(steps_between):
                t += ode.Step(u, r, rn, t)
        if log:
            log(f"harvest seed {seed}: {snapshots} snapshots x {c.n_cell} cells, {time.time() - t0:.1f}s")
    return np.concatenate(Xs), np.concatenate(Ts)


def train(X, T, epochs=30, batch=4096, lr=2e-3, lr_final=1e-5, seed=0, n=16, nthreads=0):
    """Trains a fresh network on (X, T); returns (parameters, loss history)."""
    c, ev = _vortex(n, 1, nthreads)  # any context: training does not touch the mesh
    ws = WarmStart(ev, seed=seed, capacity=X.shape[0])
    ws.add_rows(X, T)
    hist = ws.train(epochs, batch, lr, lr_final)
    return ws.get_params(), hist


def evaluate(params, seed=9, n=40, snapshots=20, steps_between=5, tol=1e-10, max_sweeps=600, nthreads=0):
    """Sweeps to `tol` from zero, from the network's guess, from the previous
    snapshot's converged reconstruction and from that plus the network's
    correction; mean over snapshots.  guess_rel_err: the guess's error
    relative to the converged reconstruction; correction_rel_err: the
    corrected start's error relative to the change since the previous
    snapshot (1 for the previous reconstruction itself)."""
    c, ev = _vortex(n, seed, nthreads)
    ws = WarmStart(ev, seed=seed, capacity=0)
    ws.set_params(params)
    u = ev.BuildMean().upload(c.u0)
    r, rn = c.vfv.BuildRec().setConstant(0.0), c.vfv.BuildRec()
    ode = ExplicitSSPRK3LocalDt(ev, nInternalRecStep=5, CFL=0.5)
    trial, scratch, du = c.vfv.BuildRec(), c.vfv.BuildRec(), ev.BuildMean()
    prev_x = prev_u = None
    rows = []
    t = 0.0
    for k in range(snapshots):
        counts = {}
        u_now = u.download()
        trial.setConstant(0.0)
        counts["zero"], _ = c.vfv.ReconstructToTolerance(u, trial, scratch, tol, max_sweeps, t)
        converged = trial.download()
        ws.predict(u, trial)
        guess = trial.download()
        counts["learned"], _ = c.vfv.ReconstructToTolerance(u, trial, scratch, tol, max_sweeps, t)
        if k > 0:
            trial.upload(prev_x)
            counts["previous"], _ = c.vfv.ReconstructToTolerance(u, trial, scratch, tol, max_sweeps, t)
            du.upload(u_now - prev_u)
            ws.predict(du, trial)
            corrected = prev_x + trial.download()
            trial.upload(corrected)
            counts["previous_plus_learned"], _ = c.vfv.ReconstructToTolerance(u, trial, scratch, tol, max_sweeps, t)
            change = max(np.linalg.norm(converged - prev_x), 1e-300)
            counts["correction_rel_err"] = float(np.linalg.norm(corrected - converged) / change)
        err = np.linalg.norm(guess - converged) / max(np.linalg.norm(converged), 1e-300)
        counts["guess_rel_err"] = float(err)
        rows.append(counts)
        prev_x, prev_u = converged, u_now
        for _ in range(steps_between):
            t += ode.Step(u, r, rn, t)
    out = {key: float(np.mean([row[key] for row in rows if key in row])) for key in
           ("zero", "learned", "previous", "previous_plus_learned", "guess_rel_err", "correction_rel_err")}
    out["snapshots"] = snapshots
    out["sweeps_saved_vs_zero"] = out["zero"] - out["learned"]
    out["sweeps_saved_vs_previous"] = out["previous"] - out["previous_plus_learned"]
    return out
