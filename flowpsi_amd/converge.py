"""Time to a 1e-8 residual drop: plain FP64 path vs FP32 (+ learned guess).

The residual is rtrms + etrms + mtrms of src/solverResidual.loci:74-112,
relative to the first Newton step.  Paths:

  fp64_fixed   FP64 Jacobian/LU/SGS, the workload's fixed sweep count
  fp32_fixed   FP32 Jacobian/LU/SGS (flowPsi's real_fj = float), same sweeps
  fp32_tol     FP32, sweeps until the linear residual falls by `rtol`
  fp32_nn      as fp32_tol, starting from the learned guess (plain start
               whenever the guess does not lower the linear residual)

The network is trained first on (system, solution) pairs the solver
harvests from its own first Newton steps; its training time is reported
separately and not included in the solve time.
"""
from __future__ import annotations

import time

import numpy as np

from .capi import Solver


def state_err(q, ref):
    """max relative difference with T, |u| and p scales (the velocity
    components share the speed scale: they vanish at a rest state)"""
    sT = np.abs(ref[0]).max()
    su = np.sqrt((ref[1:4] ** 2).sum(0)).max() + 1e-300
    sp = np.abs(ref[4]).max()
    scale = np.array([sT, su, su, su, sp])[:, None]
    return float((np.abs(q - ref) / scale).max())


def _drop(hist):
    tot = np.array([h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist])
    return np.sqrt(tot / tot[0])


def run_fixed(g, q, prm, precision, target=1e-8, max_steps=400, chunk=10):
    s = Solver(g, dict(prm, precision=precision))
    s.set_state(q)
    s.synchronize()
    hist = []
    t0 = time.perf_counter()
    while len(hist) < max_steps:
        hist += s.newton(chunk)
        d = _drop(hist)
        if d.min() < target:
            break
    s.synchronize()
    dt = time.perf_counter() - t0
    d = _drop(hist)
    n = int(np.argmax(d < target)) if (d < target).any() else -1
    # time to the first step below target, assuming uniform step cost
    t_hit = dt * (n + 1) / len(hist) if n >= 0 else None
    return dict(steps_to_target=n, time_to_target_s=t_hit, final_drop=float(d[-1]), state=s.get_state(),
                sweeps_per_step=prm["sgs_iter"])


def run_tol(g, q, prm, precision, rtol, max_sweeps, use_guess, solver=None, target=1e-8, max_steps=400,
            chunk=10):
    s = solver or Solver(g, dict(prm, precision=precision))
    s.set_state(q)
    s.synchronize()
    hist, lin = [], []
    t0 = time.perf_counter()
    while len(hist) < max_steps:
        h, l = s.newton_adaptive(chunk, rtol, max_sweeps, use_guess)
        hist += h
        lin += l
        if _drop(hist).min() < target:
            break
    s.synchronize()
    dt = time.perf_counter() - t0
    d = _drop(hist)
    n = int(np.argmax(d < target)) if (d < target).any() else -1
    sweeps = np.array([x["sweeps"] for x in lin])
    t_hit = dt * (sweeps[:n + 1].sum() + n + 1) / (sweeps.sum() + len(lin)) if n >= 0 else None
    return dict(steps_to_target=n, time_to_target_s=t_hit, final_drop=float(d[-1]), state=s.get_state(),
                mean_sweeps=float(sweeps[:max(n, 1)].mean()),
                guess_kept=float(np.mean([x["used_guess"] for x in lin])) if use_guess else None)


def train_guess(g, q, prm, precision, harvest_steps=4, target_sweeps=30, epochs=400, lr=2e-3, seed=1):
    s = Solver(g, dict(prm, precision=precision))
    s.set_state(q)
    s.mlp_init(seed)
    t0 = time.perf_counter()
    for _ in range(harvest_steps):
        s.assemble()
        s.mlp_harvest(target_sweeps)
        s.update()
    loss0 = s.mlp_train(1, lr=0.0)
    loss = s.mlp_train(epochs, lr=lr)
    s.synchronize()
    return s, dict(train_time_s=time.perf_counter() - t0, loss_before=loss0, loss_after=loss,
                   harvest_steps=harvest_steps, epochs=epochs)


def compare(g, q, prm, rtol=0.1, max_sweeps=20, target=1e-8, max_steps=400):
    """The four paths on one workload.  rtol defaults to flowPsi's own
    relative linear tolerance, LSGSRelTol = 0.1 (src/linearSolverLSGS.loci:96-107)."""
    out = {}
    r64 = run_fixed(g, q, prm, "fp64", target, max_steps)
    out["fp64_fixed"] = {k: v for k, v in r64.items() if k != "state"}
    r32 = run_fixed(g, q, prm, "fp32", target, max_steps)
    out["fp32_fixed"] = {k: v for k, v in r32.items() if k != "state"}
    rt = run_tol(g, q, prm, "fp32", rtol, max_sweeps, False, target=target, max_steps=max_steps)
    out["fp32_tol"] = {k: v for k, v in rt.items() if k != "state"}
    s, tr = train_guess(g, q, prm, "fp32")
    rn = run_tol(g, q, prm, "fp32", rtol, max_sweeps, True, solver=s, target=target, max_steps=max_steps)
    out["fp32_nn"] = {k: v for k, v in rn.items() if k != "state"}
    out["fp32_nn"]["training"] = tr
    ref = r64["state"]
    for k, r in (("fp32_fixed", r32), ("fp32_tol", rt), ("fp32_nn", rn)):
        out[k]["final_state_err_vs_fp64"] = state_err(r["state"], ref)
    t64 = out["fp64_fixed"]["time_to_target_s"]
    tnn = out["fp32_nn"]["time_to_target_s"]
    out["speedup_fp32_nn_vs_fp64"] = (t64 / tnn) if (t64 and tnn) else None
    out["rtol"] = rtol
    return out
