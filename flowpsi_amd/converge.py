"""Time to a 1e-8 residual drop: the plain FP64 path against FP32 Jacobians
with the learned initial guess.

The residual is rtrms + etrms + mtrms of src/solverResidual.loci:74-112,
relative to the first Newton step.  Paths, all from the same initial state:

  fp64_ref   FP64 Jacobian/diagonal/sweeps, the workload's own linear solve
             (line SGS: flowPsi's update-norm test, at most LSGSMaxIter
             sweeps; block SGS: gauss_seidel_iter sweeps) -- the plain path
  fp32_ref   the same with real_fj = float
  fp32_tol   FP32, the fluid's sweeps stop once the linear residual
             ||B - A X|| falls below rtol x that of the plain start
  fp32_nn    as fp32_tol, starting from the learned guess (the plain start
             is kept whenever the guess does not lower the linear residual)
  fp64_tol   FP64 with the same tolerance (separates precision from the
             solve's stopping rule)
  fp32_max<k> FP32 with the workload's own solve capped at k sweeps
             (LSGSMaxIter = k under line SGS, gauss_seidel_iter = k under
             block SGS): flowPsi's options, no tolerance and no guess

Every path is clocked directly: one Newton step per call, the host checks
the drop after each, and the clock stops at the step whose residual first
shows the target drop.  The network is trained before its path runs, on
samples the solver harvests from its own first Newton steps on the same
grid; the training time is reported beside the path and added in
`time_to_target_incl_training_s`.

Final states are compared with an FP64 reference run continued to a drop
of `ref_drop` (well below the target): at the step each path reaches the
target, and after each path is itself continued to `ref_drop`.  Each
path's state at the target is also compared with the first path's state at
its target (the precision's own error at the same residual level).
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
    return float((np.abs(q[:5] - ref[:5]) / scale).max())


def _tot(h):
    return h["rtrms"] + h["etrms"] + h["mtrms"]


class _Run:
    """One path: a solver stepping from q, one Newton step per call."""

    def __init__(self, g, q, prm, tol=None, guess=False, solver=None):
        self.s = solver or Solver(g, prm)
        self.s.set_state(q)
        self.tol, self.guess = tol, guess
        self.hist, self.lin = [], []

    def step(self):
        if self.tol is None:
            self.hist += self.s.newton(1)
        else:
            h, l = self.s.newton_adaptive(1, self.tol[0], self.tol[1], self.guess)
            self.hist += h
            self.lin += l
        return np.sqrt(_tot(self.hist[-1]) / _tot(self.hist[0]))

    def run_to(self, target, max_steps, say=None, every_s=30.0):
        """Newton steps until the drop is below target (or max_steps in
        all); say(msg), when given, hears the progress every every_s
        seconds (the host's drop check already synchronises each step)."""
        self.s.synchronize()
        t0 = time.perf_counter()
        t_say = t0
        d = 1.0
        while len(self.hist) < max_steps:
            d = self.step()
            if d < target:
                break
            if say is not None and time.perf_counter() - t_say > every_s:
                t_say = time.perf_counter()
                say(f"  step {len(self.hist)}: drop {d:.2e} after {t_say - t0:.0f}s")
        self.s.synchronize()
        return time.perf_counter() - t0, d


def train_guess(g, q, prm, harvest_steps=8, target_sweeps=30, epochs=200, lr=2e-3, seed=1):
    """A context for the guess path: the network trained on (system, solver
    iterate after target_sweeps sweeps) pairs from the first harvest_steps
    Newton steps of the same path from the same state."""
    s = Solver(g, prm)
    s.set_state(q)
    s.mlp_init(seed)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(harvest_steps):
        s.assemble()
        s.mlp_harvest(target_sweeps)
        s.update()
    loss0 = s.mlp_train(1, lr=0.0)
    loss = s.mlp_train(epochs, lr=lr)
    s.synchronize()
    return s, dict(train_time_s=time.perf_counter() - t0, loss_before=loss0, loss_after=loss,
                   harvest_steps=harvest_steps, target_sweeps=target_sweeps, epochs=epochs)


def compare(g, q, prm, rtol=0.1, max_sweeps=20, target=1e-8, max_steps=3000, ref_drop=1e-11,
            paths=("fp64_ref", "fp32_ref", "fp32_tol", "fp32_nn", "fp64_tol"), log=None,
            continue_paths=True, train_kw=None, sweep_fp32=None):
    """The paths on one workload, one solver alive at a time; rtol is the
    relative linear-residual tolerance of the _tol and _nn paths.  The
    first path is the reference: it runs on to ref_drop before the others
    start.  continue_paths=False skips continuing the others to ref_drop
    (their converged-state error is then not reported).  train_kw goes to
    train_guess (harvest_steps, target_sweeps, epochs, lr).  sweep_fp32,
    when given, is the sweep mode of the FP32 paths (each precision runs
    its faster dataflow form; every form gives the same bits)."""
    say = log or (lambda m: None)
    out = {"target": target, "rtol": rtol, "max_sweeps": max_sweeps, "ref_drop": ref_drop}
    qref = q_base = None
    for name in paths:
        prec, kind = name.split("_", 1)
        p = dict(prm, precision=prec)
        if prec == "fp32" and sweep_fp32:
            p["sweep"] = sweep_fp32
        if kind.startswith("max"):
            k = int(kind[3:])
            p.update(lsgs_maxiter=k) if prm.get("solver") == "lsgs" else p.update(sgs_iter=k)
        tol = (rtol, max_sweeps) if kind in ("tol", "nn") else None
        tr = None
        solver = None
        if name.endswith("_nn"):
            solver, tr = train_guess(g, q, p, **(train_kw or {}))
            say(f"{name}: trained in {tr['train_time_s']:.1f}s, loss {tr['loss_before']:.3e} -> {tr['loss_after']:.3e}")
        r = _Run(g, q, p, tol, name.endswith("_nn"), solver)
        t, d = r.run_to(target, max_steps, say)
        n = len(r.hist) - 1 if d < target else None
        res = dict(steps_to_target=n, time_to_target_s=t if n is not None else None, final_drop=float(d),
                   ms_per_step=1e3 * t / max(len(r.hist), 1), sweep=p.get("sweep"))
        if r.lin:
            sw = np.array([x["sweeps"] for x in r.lin])
            res["mean_sweeps"] = float(sw.mean())
            if name.endswith("_nn"):
                res["guess_kept"] = float(np.mean([x["used_guess"] for x in r.lin]))
                res["mean_rguess_over_r0"] = float(np.mean([x["rguess"] / x["r0"] for x in r.lin]))
        else:
            st = r.s.lsgs_stats() if prm.get("solver") == "lsgs" else None
            res["last_lsgs"] = st
        if tr is not None:
            res["training"] = tr
            if res["time_to_target_s"] is not None:
                res["time_to_target_incl_training_s"] = res["time_to_target_s"] + tr["train_time_s"]
        q_target = r.s.get_state()
        say(f"{name}: {n} steps, {res['time_to_target_s']} s, drop {d:.2e}; {res}")
        out[name] = res
        if q_base is not None:
            res["state_err_at_target_vs_first_path_at_target"] = state_err(q_target, q_base)
        if qref is None:  # the reference converged state: this path continued
            q_base = q_target.copy()
            r.run_to(ref_drop, max_steps, say)
            qref = r.s.get_state()
            out["reference"] = dict(path=name, steps=len(r.hist),
                                    drop=float(np.sqrt(_tot(r.hist[-1]) / _tot(r.hist[0]))))
            say(f"reference: {name} continued to drop {out['reference']['drop']:.2e} after {len(r.hist)} steps")
            res["state_err_converged_vs_converged_fp64"] = 0.0
        elif continue_paths:
            r.run_to(ref_drop, max_steps, say)
            res["state_err_converged_vs_converged_fp64"] = state_err(r.s.get_state(), qref)
        res["state_err_at_target_vs_converged_fp64"] = state_err(q_target, qref)
        r.s.close()
        del r, q_target
    t64 = out.get("fp64_ref", {}).get("time_to_target_s")
    for name in paths:
        tn = out[name].get("time_to_target_s")
        if name != "fp64_ref" and t64 and tn:
            out[name]["speedup_vs_fp64_ref"] = t64 / tn
    tnn = out.get("fp32_nn", {}).get("time_to_target_s")
    out["speedup_fp32_nn_vs_fp64"] = (t64 / tnn) if (t64 and tnn) else None
    return out


def guess_gain(g, q, prm, transient=5, harvest_steps=4, target_sweeps=30, epochs=300, lr=5e-3, eval_steps=4,
               rtol=1e-3, max_sweeps=60, seed=1, log=None):
    """What the trained network gains on Newton steps it was not trained on.

    After `transient` plain Newton steps, `harvest_steps` steps harvest
    (system, `target_sweeps`-sweep solution) pairs, one per cell each, and
    the network trains `epochs` epochs on them.  From the state after the
    harvest, `eval_steps` Newton steps then run with the fluid's sweeps
    stopped at a linear residual of `rtol` x the plain start's, first from
    the learned guess and then again from the same state without it (the
    guess reads the previous step's solution, so the learned run goes
    first).  Reported per path: the sweeps of each step, their total, the
    device time per step (host clock around synchronised calls), and for
    the guess the linear residual of the guess over the plain start's."""
    say = log or (lambda m: None)
    s = Solver(g, prm)
    try:
        s.set_state(q)
        s.mlp_init(seed)
        s.newton(transient)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(harvest_steps):
            s.assemble()
            s.mlp_harvest(target_sweeps)
            s.update()
        loss0 = s.mlp_train(1, lr=0.0)
        loss = s.mlp_train(epochs, lr=lr)
        s.synchronize()
        out = dict(cells=g.nc, harvested_pairs=g.nc * harvest_steps, rtol=rtol,
                   training=dict(train_time_s=time.perf_counter() - t0, loss_before=loss0, loss_after=loss,
                                 harvest_steps=harvest_steps, target_sweeps=target_sweeps, epochs=epochs))
        say(f"trained on {g.nc * harvest_steps} pairs in {out['training']['train_time_s']:.1f}s, "
            f"loss {loss0:.3e} -> {loss:.3e}")
        qh = s.get_state()
        for use in (True, False):
            s.set_state(qh)
            s.synchronize()
            c0 = time.perf_counter()
            _, lin = s.newton_adaptive(eval_steps, rtol, max_sweeps, use_guess=use)
            s.synchronize()
            dt = time.perf_counter() - c0
            r = dict(sweeps=[x["sweeps"] for x in lin], total_sweeps=int(sum(x["sweeps"] for x in lin)),
                     ms_per_step=1e3 * dt / eval_steps,
                     reached=[bool(x["rfinal"] <= rtol * x["r0"]) for x in lin])
            if use:
                r["used_guess"] = [int(x["used_guess"]) for x in lin]
                r["rguess_over_r0"] = [float(x["rguess"] / x["r0"]) for x in lin]
            out["learned" if use else "plain"] = r
            say(f"{'learned' if use else 'plain'}: {r}")
        out["sweeps_saved"] = out["plain"]["total_sweeps"] - out["learned"]["total_sweeps"]
    finally:
        s.close()
    return out


if __name__ == "__main__":
    import json
    import sys

    from . import workloads

    # the learned guess on the full 8M-cell prism-layer + hex grid (SA, block SGS)
    t0 = time.time()
    g, q, prm = workloads.c4_hybrid(nx=200, nz=200, n_prism=30, n_hex=140)
    lg = lambda m: print(f"[guess c4] {m}", file=sys.stderr, flush=True)
    lg(f"grid {g.nc} cells built in {time.time() - t0:.1f}s")
    for prec in sys.argv[1:] or ["fp32"]:
        r = guess_gain(g, q, dict(prm, precision=prec, sweep="persistent"), log=lg)
        r.update(workload="c4", precision=prec)
        print(json.dumps(r), flush=True)
