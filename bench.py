"""Newton-step throughput of the HIP path on one MI355X.

Default workload: laminar Navier-Stokes on the 128^3 stretched channel
(2,097,152 cells), HLLC + viscous fluxes, FP64 everywhere (the plain path),
10 symmetric block-SGS sweeps per steady Newton step.  A step is one full
Newton step: residual, Jacobian assembly, block LU, SGS, update.  The state
is resident in HBM before the timed region.

Prints one JSON line with the metric, the roofline of the dominant kernel
(HIP events on the library's stream, algorithmic bytes from the library's
byte model) and a CPU baseline (the oracle on this host, OpenMP threads and
one thread, on a bounded sample of the same workload).  The Gauss-Seidel
half-sweeps run as one cooperative dataflow launch each (--sweep
persistent, the fastest mode measured); per_level and tiled give the same
bits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calib_copy(nbytes):
    """A plain device-to-device copy of known size through the HIP runtime,
    used to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def cpu_baseline(g, q, prm, steps):
    """The oracle (C restatement of the reference path) timed on the host:
    real_fj = float as the reference's default build, face and cell loops
    on OMP_NUM_THREADS threads (16 at most), sweeps sequential as in the
    reference; the same sample again on one thread."""
    import ctypes
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    threads = max(1, min(16, threads))
    o = oracle.Oracle(g, prm, "f32", omp=True)
    gomp = ctypes.CDLL("libgomp.so.1")
    res = {}
    for n in (threads, 1):
        gomp.omp_set_num_threads(n)
        log(f"cpu baseline: oracle real_fj=float, {steps} Newton step(s) on the same grid, {n} thread(s)")
        c0 = time.perf_counter()
        o.newton(q, steps)
        res[n] = time.perf_counter() - c0
    gomp.omp_set_num_threads(threads)
    cpu = {"value": g.nc * steps / res[threads], "unit": "cells*Newton-steps/s", "cores": threads,
           "kind": "port", "value_1core": g.nc * steps / res[1],
           "sample": f"{steps} Newton step(s) of the same workload ({g.nc} cells) from the same initial "
                     f"state, oracle with real_fj=float (the reference default), OpenMP face/cell loops, "
                     f"sequential SGS: {res[threads]:.1f}s on {threads} threads, {res[1]:.1f}s on 1"}
    log(f"cpu baseline {cpu['value']:.3e} cells*steps/s ({threads} threads), {cpu['value_1core']:.3e} (1 thread)")
    return cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="channel128",
                    choices=["channel128", "channel128_sa", "channel128_sst", "box32", "c3", "c4", "c5"],
                    help="channel128 is the metric's workload; the turbulent channels (SA, SST) and the "
                         "full-size tetrahedral (c3, 4.1M cells), hybrid (c4, 8M) and boundary-layer "
                         "(c5, 20M, line SGS) grids are extra lines")
    ap.add_argument("--n", type=int, default=None, help="grid cells per axis (default 128 / 32)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--sweep", default=None, choices=["per_level", "tiled", "persistent"],
                    help="default persistent (the dataflow); per_level for c3, whose six dependency levels "
                         "of ~690k cells each fill the GPU in one launch per level")
    ap.add_argument("--time-scheme", default="steady", choices=["steady", "bdf2", "kec"],
                    help="bdf2/kec: time-accurate dual time stepping; a step is one time step of "
                         "--newton-iter Newton iterations and the value counts Newton iterations")
    ap.add_argument("--newton-iter", type=int, default=3)
    ap.add_argument("--dt", type=float, default=2e-6, help="physical time step of bdf2/kec (s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-steps", type=int, default=1)
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, copy 1 GiB device-to-device (counter calibration)")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the time-to-1e-8 comparison on the 32^3 box")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        sys.exit(2)

    import numpy as np
    from flowpsi_amd import workloads
    from flowpsi_amd.capi import Solver

    t0 = time.time()
    if a.workload == "channel128":
        n = a.n or 128
        g, q, prm = workloads.channel(n)
        wname = f"laminar NS channel {n}^3 ({g.nc} cells), HLLC+viscous, SGS x{prm['sgs_iter']}"
    elif a.workload in ("channel128_sa", "channel128_sst"):
        n = a.n or 128
        turb = a.workload.split("_")[1]
        g, q, prm = (workloads.channel_sa if turb == "sa" else workloads.channel_sst)(n)
        wname = (f"{turb.upper()} NS channel {n}^3 ({g.nc} cells), HLLC+viscous, segregated "
                 f"{'nuTilde' if turb == 'sa' else 'k and omega'}, SGS x{prm['sgs_iter']}")
    elif a.workload == "c3":
        n = a.n or 88
        g, q, prm = workloads.c3_tets(n, precision=a.precision)
        wname = (f"SA on {g.nc} jittered tetrahedra ({n}^3 hexes x 6), M 0.5, Re 1e6/m, HLLC+viscous, "
                 f"SGS x{prm['sgs_iter']}")
    elif a.workload == "c4":
        n = a.n or 200
        g, q, prm = workloads.c4_hybrid(nx=n, nz=n, n_prism=30, n_hex=140)
        wname = (f"SA on a {g.nc}-cell prism-layer + jittered-hex grid ({n}x{n} columns, 30 prism and 140 "
                 f"hex layers), M 0.3, HLLC+viscous, SGS x{prm['sgs_iter']}")
    elif a.workload == "c5":
        n = a.n or 577
        g, q, prm = workloads.c5_bl(nx=n, ny=60, nz=n)
        wname = (f"SST on a {g.nc}-cell boundary-layer grid ({n}x60x{n}, first-cell aspect 1e3), M 0.6, "
                 f"Re 5e6/m, line SGS (max {prm.get('lsgs_maxiter', 10)} sweeps, turbulence scalars fused)")
    else:
        n = a.n or 32
        g, q, prm = workloads.box(n)
        wname = f"Euler box {n}^3 ({g.nc} cells), Roe, SGS x{prm['sgs_iter']}"
    if a.sweep is None:
        a.sweep = "per_level" if a.workload == "c3" else "persistent"
    prm = dict(prm, precision=a.precision, sweep=a.sweep)
    unsteady = a.time_scheme != "steady"
    iters = a.newton_iter if unsteady else 1
    if unsteady:
        prm = dict(prm, unsteady=True, time_scheme=a.time_scheme, newton_iter=a.newton_iter, dtmax=a.dt)
        wname += f", {a.time_scheme.upper()} dt={a.dt:g}s x{a.newton_iter} Newton iterations"
    log(f"grid {g.name}: nc={g.nc} nf={g.nf} built in {time.time() - t0:.1f}s")
    s = Solver(g, prm)
    info = s.info()
    log(f"context: {info}")
    s.set_state(q)
    step = (lambda n: [h for st in s.unsteady(n) for h in st]) if unsteady else s.newton
    t1 = time.time()
    step(a.warmup)
    s.synchronize()
    log(f"warmup {a.warmup} steps {time.time() - t1:.2f}s")

    s.profile(True)
    s.synchronize()
    t0 = time.perf_counter()
    hist = step(a.steps)
    s.synchronize()
    t1 = time.perf_counter()
    prof = s.profile_read()
    s.profile(False)
    dt = t1 - t0
    ms_per_step = 1e3 * dt / a.steps
    value = g.nc * a.steps * iters / dt
    log(f"timed {a.steps} steps: {ms_per_step:.2f} ms/step; groups " +
        ", ".join(f"{k}={v['ms'] / max(v['launches'], 1):.3f}ms x{v['launches']}" for k, v in prof.items()))

    # dominant kernel group by time; a half-sweep of the per-level path is
    # nlevels launches of k_sgs_level, reported per launch
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    avg_s = d["ms"] / max(d["launches"], 1) / 1e3
    bytes_launch = d["bytes"]
    kname = dom
    tiled = a.sweep == "tiled"
    line = prm.get("solver", "sgs") == "lsgs"
    if dom == "sgs_sweep" and line:
        if a.sweep == "persistent" and info.get("lsgs_dataflow"):
            kname = "k_lsgs_sweep"         # one cooperative launch per half-sweep
        else:
            avg_s /= info["lsgs_nlevels"]
            bytes_launch /= info["lsgs_nlevels"]
            kname = "k_lsgs_level"
    elif dom == "sgs_sweep":
        if tiled:
            kname = "k_sgs_tiles"          # one cooperative launch per half-sweep
        elif a.sweep == "persistent":
            # one cooperative launch per half-sweep: flag-free for the fluid
            # alone, with flags when the turbulence scalars ride along
            kname = "k_sgs_flow" if prm.get("turb", "none") == "none" else "k_sgs_sweep"
        else:
            avg_s /= info["nlevels"]       # nlevels launches per half-sweep
            bytes_launch /= info["nlevels"]
            kname = "k_sgs_level"
    achieved = bytes_launch / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"{kname[2:]}_channel128_traffic.json")
    if a.workload == "channel128" and (a.n or 128) == 128 and os.path.exists(tpath):
        t = json.load(open(tpath))
        key = f"fpsi::{kname}<double>" if a.precision == "fp64" else f"fpsi::{kname}<float>"
        if t.get("precision") == a.precision and key in t["kernels"]:
            traffic = t["kernels"][key]["bytes"]
    # whole-step algorithmic bytes (per Newton step) for the step-level roofline
    # half-sweeps per Newton iteration as run (line SGS stops by its own test)
    sweeps = prof["sgs_sweep"]["launches"] / (a.steps * iters)
    step_bytes = (prof["residual"]["bytes"] + prof["face_jac"]["bytes"] + prof["assemble"]["bytes"] +
                  sweeps * prof["sgs_sweep"]["bytes"] + prof["update"]["bytes"])
    step_gbs = step_bytes * a.steps * iters / dt / 1e9

    if a.calib_copy:
        calib_copy(1 << 30)
        log("calibration copy: 1 GiB read + 1 GiB written (hipMemcpy device to device)")

    tot = [h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist]
    cpu = None
    if not a.no_cpu_baseline and not unsteady and a.workload.startswith("channel"):
        cpu = cpu_baseline(g, q, prm, a.cpu_sample_steps)

    conv = None
    if not a.no_converge and not unsteady and a.workload == "channel128":
        from flowpsi_amd import converge
        gb, qb, pb = workloads.box(32, sweep="persistent")
        log("time to 1e-8 drop on the 32^3 Euler box (dataflow sweeps): fp64 fixed / fp32 fixed / fp32 tol / fp32+NN")
        conv = converge.compare(gb, qb, pb)
        conv["workload"] = "Euler box 32^3 (32768 cells), Roe, CFL 10, SGS x5 (fixed paths)"
        log(f"converge: " + ", ".join(f"{k}={v.get('time_to_target_s')}" for k, v in conv.items()
                                      if isinstance(v, dict) and "time_to_target_s" in v))

    out = {
        "metric": "cells*Newton-steps/s",
        "value": value,
        "unit": "cells*Newton-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64 residual/update, f32 Jacobian/LU/SGS matrices",
        "data": "synthetic (seeded grid and initial state)",
        "config": {"workload": wname, "cells": g.nc, "faces": g.nf, "levels": info["nlevels"],
                   "precision": a.precision, "sweep": "tiled" if tiled else a.sweep,
                   "tiles": info["sweep_tiles"], "sgs_sweeps": prm["sgs_iter"],
                   "parallelism": "single GPU", "time_scheme": a.time_scheme,
                   "newton_iterations_per_step": iters},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": avg_s * 1e3, "bytes_per_launch": bytes_launch},
        "step_roofline": {"achieved": step_gbs, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_step": step_bytes},
        "residual_drop": float(np.sqrt(tot[-1] / tot[0])) if tot[0] > 0 else None,
        "cpu_baseline": cpu,
        "time_to_1e-8": conv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
