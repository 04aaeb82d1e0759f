"""FP32 Jacobians against the FP64 path, per configuration, at full size.

The mixed-precision path keeps the residual, the state update, primitive
recovery and the norms in FP64 and stores the flux Jacobians, the diagonal
blocks, their factors and the right-hand side in FP32 (flowPsi's `real_fj`
= float build, include/flowPsi.lh:109).  For one workload this module runs
both precisions side by side from the same initial state and reports

* the first linear update: the system assembled at the initial state and
  solved with the workload's own sweeps (`gauss_seidel_iter` block SGS
  sweeps, or line SGS with flowPsi's update-norm test and at most
  `LSGSMaxIter` sweeps), FP32 against FP64, relative to the FP64 update's
  largest magnitude per component, and on the state's T, |u|, p scales;
* the state after every Newton step, FP32 against FP64, with T, |u| and p
  scales (`converge.state_err`), and both residual drops;
* optionally the states once both paths have converged to `conv_drop`.

The FP64 path is the reference here: at the sizes the CPU oracle finishes
in seconds it equals the oracle to 1e-12 (residuals) and 1e-10 (states).

Run as `python -m flowpsi_amd.precision [c1 c2 c3 c4 c5]`; one JSON line per
workload.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

from .capi import Solver
from .converge import state_err

CONFIGS = ("c1", "c2", "c3", "c4", "c5")


def workload(name, small=False):
    """(grid, state, options, description) of a configuration at its full
    size, or at a small size of the same kind (`small`)."""
    from . import workloads as w
    if name == "c1":
        n = 8 if small else 32
        g, q, prm = w.box(n)
        return g, q, prm, f"Euler box {n}^3, Roe, SGS x{prm['sgs_iter']}"
    if name == "c2":
        g, q, prm = w.channel(nx=12, ny=16, nz=8) if small else w.channel(128)
        return g, q, prm, f"laminar channel ({g.nc} cells), HLLC+viscous, SGS x{prm['sgs_iter']}"
    if name == "c3":
        n = 4 if small else 88
        g, q, prm = w.c3_tets(n)
        return g, q, prm, f"SA on {g.nc} jittered tetrahedra, SGS x{prm['sgs_iter']}"
    if name == "c4":
        g, q, prm = (w.c4_hybrid(nx=6, nz=6, n_prism=4, n_hex=4) if small else
                     w.c4_hybrid(nx=200, nz=200, n_prism=30, n_hex=140))
        return g, q, prm, f"SA on a {g.nc}-cell prism-layer + hex grid, SGS x{prm['sgs_iter']}"
    if name == "c5":
        n = 6 if small else 577
        g, q, prm = w.c5_bl(nx=n, ny=60, nz=n)
        return g, q, prm, f"SST on a {g.nc}-cell boundary-layer grid ({n}x60x{n}), line SGS"
    raise ValueError(name)


def _sweeps(prm):
    return int(prm.get("lsgs_maxiter", 10)) if prm.get("solver") == "lsgs" else int(prm.get("sgs_iter", 5))


def _drop(h, h0):
    t = h["rtrms"] + h["etrms"] + h["mtrms"]
    t0 = h0["rtrms"] + h0["etrms"] + h0["mtrms"]
    return float(np.sqrt(t / t0)) if t0 > 0 else 0.0


def compare(g, q, prm, nsteps, conv_drop=None, conv_max=0, log=None):
    """FP32 against FP64 on one workload (see the module docstring); both
    contexts are alive together, so the device holds both systems."""
    say = log or (lambda m: None)
    prm = dict(prm)
    prm.pop("precision", None)
    sv = {p: Solver(g, dict(prm, precision=p)) for p in ("fp64", "fp32")}
    out = {"cells": g.nc, "steps": nsteps}
    try:
        dq = {}
        for p, s in sv.items():
            s.set_state(q)
            s.residual(want_faces=False)
            s.assemble()
            dq[p] = s.linear_solve(_sweeps(prm))
            s.set_state(q)
        scale = np.abs(dq["fp64"]).max(axis=0) + 1e-300
        out["linear_update_err"] = float((np.abs(dq["fp32"] - dq["fp64"]) / scale).max())
        # the same difference on the state's T, |u|, p scales
        out["linear_update_err_state_scale"] = state_err(q[:5] + dq["fp32"].T, q[:5] + dq["fp64"].T)
        say(f"first linear update: fp32 vs fp64 {out['linear_update_err']:.3e} of the update, "
            f"{out['linear_update_err_state_scale']:.3e} of the state")
        del dq
        per = []
        h0 = {}
        t0 = time.perf_counter()
        for k in range(nsteps):
            st = {}
            for p, s in sv.items():
                h = s.newton(1)[0]
                h0.setdefault(p, h)
                st[p] = (s.get_state(), _drop(h, h0[p]))
            e = state_err(st["fp32"][0], st["fp64"][0])
            per.append({"step": k + 1, "state_err": e, "drop_fp64": st["fp64"][1], "drop_fp32": st["fp32"][1]})
            say(f"step {k + 1}: state fp32 vs fp64 {e:.3e}, drops {st['fp64'][1]:.3e} / {st['fp32'][1]:.3e} "
                f"({time.perf_counter() - t0:.1f}s)")
        out["per_step"] = per
        out["final_state_err"] = per[-1]["state_err"] if per else None
        if conv_drop:
            done = {}
            for p, s in sv.items():
                n = 0
                d = per[-1]["drop_" + p] if per else 1.0
                while d > conv_drop and n < conv_max:
                    d = _drop(s.newton(1)[0], h0[p])
                    n += 1
                done[p] = (s.get_state(), d, nsteps + n)
                say(f"{p} continued to a drop of {d:.3e} at step {nsteps + n}")
            out["converged"] = {"target_drop": conv_drop,
                                "steps_fp64": done["fp64"][2], "drop_fp64": done["fp64"][1],
                                "steps_fp32": done["fp32"][2], "drop_fp32": done["fp32"][1],
                                "state_err": state_err(done["fp32"][0], done["fp64"][0])}
            say(f"converged states: fp32 vs fp64 {out['converged']['state_err']:.3e}")
    finally:
        for s in sv.values():
            s.close()
    return out


def main(argv):
    names = [a for a in argv if a in CONFIGS] or list(CONFIGS)
    small = "--small" in argv
    steps = {"c1": 50, "c2": 10, "c3": 10, "c4": 10, "c5": 10}
    for name in names:
        t0 = time.time()
        g, q, prm, what = workload(name, small)
        log = lambda m, n=name: print(f"[precision {n}] {m}", file=sys.stderr, flush=True)
        log(f"{what}: {g.nc} cells, grid built in {time.time() - t0:.1f}s")
        kw = dict(conv_drop=1e-11, conv_max=400) if name == "c1" else {}
        r = compare(g, q, prm, 3 if small else steps[name], log=log, **kw)
        r.update(config=name, workload=what)
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
