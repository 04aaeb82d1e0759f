"""Newton-step throughput of the HIP path on one MI355X.

Default workload: the 20M-cell RANS boundary-layer grid (577 x 60 x 577 =
19,975,740 hexes, first-cell aspect ratio 1e3), Menter SST with k and omega
solved segregated as flowPsi does, Mach 0.6, Re 5e6 per metre, HLLC +
viscous fluxes, line SGS (flowPsi's default linear solver: at most 10
relaxed symmetric sweeps, stopped by its update-norm test) with the
turbulence scalars swept in the same launches, FP64 everywhere (the plain
path).  A step is one full steady Newton step: residual, Jacobian assembly,
diagonal factors, line SGS, update.  The state is resident in HBM before the
timed region.

Prints one JSON line with the metric, the roofline of the dominant kernel
(HIP events on the library's stream, algorithmic bytes from the library's
byte model, HBM bytes from the rocprofv3 counter summary under profiles/),
a measured device-copy bandwidth beside the vendor peak, a CPU baseline
(the oracle on this host, all the threads this process may use and one
thread, on a bounded sample of the same workload) and the time to a 1e-8
residual drop of the plain FP64 path against FP32 Jacobians with the
learned initial guess on a smaller grid of the same kind.

`--pmc-summary FETCH_DIR WRITE_DIR` turns two `rocprofv3 --pmc FETCH_SIZE`
/ `--pmc WRITE_SIZE` runs of this script (with --calib-copy) into the
traffic summary that the bench line reads.
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _hip():
    import ctypes as C
    return C, C.CDLL("libamdhip64.so")


def calib_copy(nbytes):
    """A plain device-to-device copy of known size through the HIP runtime,
    used to scale the FETCH_SIZE / WRITE_SIZE counters."""
    C, hip = _hip()
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


def stream_copy_gbs(nbytes=1 << 30, reps=20):
    """Measured copy bandwidth: device-to-device hipMemcpyAsync of `nbytes`
    (the runtime's blit kernel) `reps` times on the null stream, timed with
    HIP events; 2 x nbytes move per copy."""
    C, hip = _hip()
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)  # warm
    hip.hipDeviceSynchronize()
    hip.hipEventRecord(e0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(e1, None)
    hip.hipEventSynchronize(e1)
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), e0, e1)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def host_threads():
    """Threads the CPU baseline may use: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when it is set (the GPU boxes set it to the
    share of the host that belongs to one GPU)."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(avail, env) if env > 0 else avail), avail


def cpu_baseline(g, q, prm, steps, what):
    """The oracle (C restatement of the reference path) timed on the host:
    real_fj = float as the reference's default build, face and cell loops
    on OpenMP threads, sweeps sequential as in the reference; the same sample
    again on one thread."""
    import ctypes
    import oracle
    threads, avail = host_threads()
    o = oracle.Oracle(g, prm, "f32", omp=True)
    gomp = ctypes.CDLL("libgomp.so.1")
    res = {}
    for n in (threads, 1):
        gomp.omp_set_num_threads(n)
        log(f"cpu baseline: oracle real_fj=float, {steps} Newton step(s) of {what}, {n} thread(s)")
        c0 = time.perf_counter()
        o.newton(q, steps)
        res[n] = time.perf_counter() - c0
    gomp.omp_set_num_threads(threads)
    cpu = {"value": g.nc * steps / res[threads], "unit": "cells*Newton-steps/s", "cores": threads,
           "kind": "port", "value_1core": g.nc * steps / res[1], "host_nproc": os.cpu_count(),
           "host_cpus_available": avail,
           "sample": f"{steps} Newton step(s) of {what} ({g.nc} cells) from the workload's initial state, "
                     f"oracle with real_fj=float (the reference default), OpenMP face/cell loops, sequential "
                     f"sweeps: {res[threads]:.1f}s on {threads} threads ({avail} CPUs available to the process, "
                     f"nproc {os.cpu_count()}), {res[1]:.1f}s on 1"}
    log(f"cpu baseline {cpu['value']:.3e} cells*steps/s ({threads} threads), {cpu['value_1core']:.3e} (1 thread)")
    return cpu


def build_workload(a):
    from flowpsi_amd import workloads
    if a.workload == "c5":
        n = a.n or 577
        g, q, prm = workloads.c5_bl(nx=n, ny=60, nz=n)
        wname = (f"SST on a {g.nc}-cell boundary-layer grid ({n}x60x{n}, first-cell aspect 1e3), M 0.6, "
                 f"Re 5e6/m, HLLC+viscous, line SGS (max {prm.get('lsgs_maxiter', 10)} sweeps, k and omega "
                 f"swept in the same launches)")
        tag = "c5_20M" if n == 577 else None
    elif a.workload == "channel128":
        n = a.n or 128
        g, q, prm = workloads.channel(n)
        wname = f"laminar NS channel {n}^3 ({g.nc} cells), HLLC+viscous, SGS x{prm['sgs_iter']}"
        tag = "channel128" if n == 128 else None
    elif a.workload in ("channel128_sa", "channel128_sst"):
        n = a.n or 128
        turb = a.workload.split("_")[1]
        g, q, prm = (workloads.channel_sa if turb == "sa" else workloads.channel_sst)(n)
        wname = (f"{turb.upper()} NS channel {n}^3 ({g.nc} cells), HLLC+viscous, segregated "
                 f"{'nuTilde' if turb == 'sa' else 'k and omega'}, SGS x{prm['sgs_iter']}")
        tag = None
    elif a.workload == "c3":
        n = a.n or 88
        g, q, prm = workloads.c3_tets(n, precision=a.precision)
        wname = (f"SA on {g.nc} jittered tetrahedra ({n}^3 hexes x 6), M 0.5, Re 1e6/m, HLLC+viscous, "
                 f"SGS x{prm['sgs_iter']}")
        tag = None
    elif a.workload == "c4":
        n = a.n or 200
        g, q, prm = workloads.c4_hybrid(nx=n, nz=n, n_prism=30, n_hex=140)
        wname = (f"SA on a {g.nc}-cell prism-layer + jittered-hex grid ({n}x{n} columns, 30 prism and 140 "
                 f"hex layers), M 0.3, HLLC+viscous, SGS x{prm['sgs_iter']}")
        tag = None
    else:
        n = a.n or 32
        g, q, prm = workloads.box(n)
        wname = f"Euler box {n}^3 ({g.nc} cells), Roe, SGS x{prm['sgs_iter']}"
        tag = None
    return g, q, prm, wname, tag


def kernel_name(prof, prm, sweep, info, precision):
    """The dominant kernel group and the kernel that runs it, with its average
    launch time (s) and algorithmic bytes per launch."""
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    avg_s = d["ms"] / max(d["launches"], 1) / 1e3
    nbytes = d["bytes"]
    kname = dom
    solver = prm.get("solver", "sgs")
    if dom == "sgs_sweep" and solver in ("lsgs", "fsgs"):
        if sweep in ("persistent", "persistent_flags") and info.get("lsgs_dataflow"):
            # line SGS: flag-free (persistent) or chunk flags; FSGS always flags
            kname = "k_lsgs_flow" if (sweep == "persistent" and solver == "lsgs") else "k_lsgs_sweep"
        else:
            avg_s /= info["lsgs_nlevels"]
            nbytes /= info["lsgs_nlevels"]
            kname = "k_lsgs_level"
    elif dom == "sgs_sweep":
        if sweep == "tiled":
            kname = "k_sgs_tiles"
        elif sweep == "persistent":
            kname = "k_sgs_flow" if prm.get("turb", "none") == "none" else "k_sgs_sweep"
        elif sweep == "persistent_flags":
            kname = "k_sgs_sweep"
        else:
            avg_s /= info["nlevels"]
            nbytes /= info["nlevels"]
            kname = "k_sgs_level"
    return dom, kname, avg_s, nbytes


def traffic_lookup(kname, tag, precision):
    """HBM bytes per launch of `kname` from the committed counter summary of
    this workload, averaged over its template instances (both sweep
    directions), or None."""
    if tag is None:
        return None, None
    path = os.path.join(ROOT, "profiles", f"{kname[2:]}_{tag}_traffic.json")
    if not os.path.exists(path):
        return None, None
    t = json.load(open(path))
    if t.get("precision") != precision:
        return None, None
    ks = [v for k, v in t["kernels"].items() if k.split("<")[0] == f"fpsi::{kname}"]
    if not ks:
        return None, None
    tot = sum(v["bytes"] * v["dispatches"] for v in ks)
    n = sum(v["dispatches"] for v in ks)
    return tot / n, os.path.relpath(path, ROOT)


def pmc_summary(fetch_dir, write_dir, tag, precision, calib_bytes=1 << 30):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 counter runs
    (FETCH_SIZE, WRITE_SIZE in KiB per dispatch), each scaled by what the
    same counter reads for the calibration copy of `calib_bytes`."""
    def read(d, counter):
        rows = {}
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        for f in files:
            for r in csv.DictReader(open(f)):
                if r["Counter_Name"] != counter:
                    continue
                # "void fpsi::k_x<double, true>(fpsi::DevGrid, ...)" -> "fpsi::k_x<double, true>"
                name = r["Kernel_Name"]
                if name.startswith("void "):
                    name = name[5:]
                depth, cut = 0, len(name)
                for i, ch in enumerate(name):
                    depth += ch == "<"
                    depth -= ch == ">"
                    if ch == "(" and depth == 0:
                        cut = i
                        break
                rows.setdefault(name[:cut], []).append(float(r["Counter_Value"]))
        return rows
    fr, wr = read(fetch_dir, "FETCH_SIZE"), read(write_dir, "WRITE_SIZE")
    cal = [k for k in fr if "copyBuffer" in k]
    if not cal:
        raise SystemExit("no calibration copy (run bench.py with --calib-copy)")
    kib = calib_bytes / 1024.0
    fscale = kib / max(fr[cal[0]])     # the copy reads calib_bytes
    wscale = kib / max(wr[cal[0]])     # and writes calib_bytes
    out = {"workload": tag, "precision": precision, "fetch_scale": fscale, "write_scale": wscale, "kernels": {}}
    for k in fr:
        if not k.startswith("fpsi::"):
            continue
        f = sum(fr[k]) / len(fr[k])
        w = sum(wr.get(k, [0.0])) / max(len(wr.get(k, [0.0])), 1)
        out["kernels"][k] = {"fetch_kib": f, "write_kib": w, "bytes": 1024.0 * (f * fscale + w * wscale),
                             "dispatches": len(fr[k])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5",
                    choices=["c5", "channel128", "channel128_sa", "channel128_sst", "box32", "c3", "c4"],
                    help="c5 (20M cells, SST, line SGS) is the metric's workload; the laminar and turbulent "
                         "128^3 channels, the 32^3 box, the tetrahedral (c3, 4.1M cells) and hybrid (c4, 8M) "
                         "grids are extra lines")
    ap.add_argument("--n", type=int, default=None, help="cells per axis (c5: x and z; default 577)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--sweep", default=None, choices=["per_level", "tiled", "persistent", "persistent_flags"],
                    help="default persistent (the flag-free dataflow); per_level for c3, whose six dependency "
                         "levels of ~690k cells each fill the GPU in one launch per level")
    ap.add_argument("--sweep-fp32", default=None, choices=["per_level", "persistent", "persistent_flags"],
                    help="c5: the sweep of the FP32 paths of the time-to-1e-8 comparison (default: --sweep)")
    ap.add_argument("--time-scheme", default="steady", choices=["steady", "bdf2", "kec"],
                    help="bdf2/kec: time-accurate dual time stepping; a step is one time step of "
                         "--newton-iter Newton iterations and the value counts Newton iterations")
    ap.add_argument("--newton-iter", type=int, default=3)
    ap.add_argument("--dt", type=float, default=2e-6, help="physical time step of bdf2/kec (s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-steps", type=int, default=1)
    ap.add_argument("--cpu-sample-n", type=int, default=128,
                    help="c5: the CPU baseline runs the same kind of grid at n x 60 x n cells")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, copy 1 GiB device-to-device (counter calibration)")
    ap.add_argument("--no-converge", action="store_true", help="skip the time-to-1e-8 comparison")
    ap.add_argument("--converge-n", type=int, default=48,
                    help="c5: the time-to-1e-8 comparison runs on the n x 60 x n grid of the same kind")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/<kernel>_<workload>_traffic.json from two counter runs and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(sharding the cells over GPUs is not implemented)"}))
        sys.exit(2)

    import numpy as np
    from flowpsi_amd.capi import Solver

    if a.pmc_summary:
        tag = {"c5": "c5_20M" if (a.n or 577) == 577 else None,
               "channel128": "channel128" if (a.n or 128) == 128 else None}.get(a.workload)
        t = pmc_summary(*a.pmc_summary, tag or a.workload, a.precision)
        print(json.dumps(t, indent=1))
        return
    t0 = time.time()
    g, q, prm, wname, tag = build_workload(a)
    if a.sweep is None:
        a.sweep = {"c3": "per_level"}.get(a.workload, "persistent")
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
    lin = s.lsgs_stats() if prm.get("solver") == "lsgs" else None
    dt = t1 - t0
    ms_per_step = 1e3 * dt / a.steps
    value = g.nc * a.steps * iters / dt
    log(f"timed {a.steps} steps: {ms_per_step:.2f} ms/step; groups " +
        ", ".join(f"{k}={v['ms'] / max(v['launches'], 1):.3f}ms x{v['launches']}" for k, v in prof.items()))

    dom, kname, avg_s, bytes_launch = kernel_name(prof, prm, a.sweep, info, a.precision)
    achieved = bytes_launch / avg_s / 1e9
    traffic, tsrc = traffic_lookup(kname, tag, a.precision)
    # whole-step algorithmic bytes: half-sweeps per Newton iteration as run
    # (line SGS stops by its own test)
    sweeps = prof["sgs_sweep"]["launches"] / (a.steps * iters)
    step_bytes = (prof["residual"]["bytes"] + prof["face_jac"]["bytes"] + prof["assemble"]["bytes"] +
                  sweeps * prof["sgs_sweep"]["bytes"] + prof["update"]["bytes"])
    step_gbs = step_bytes * a.steps * iters / dt / 1e9
    tot = [h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist]
    s.close()
    del s

    if a.calib_copy:
        calib_copy(1 << 30)
        log("calibration copy: 1 GiB read + 1 GiB written (hipMemcpy device to device)")
    copy_gbs = stream_copy_gbs()
    log(f"measured device copy {copy_gbs:.0f} GB/s (read + write)")

    cpu = None
    if not a.no_cpu_baseline and not unsteady:
        if a.workload == "c5":
            from flowpsi_amd import workloads
            ncs = a.cpu_sample_n
            gc, qc, pc = workloads.c5_bl(nx=ncs, ny=60, nz=ncs)
            cpu = cpu_baseline(gc, qc, dict(pc, precision="fp32"), a.cpu_sample_steps,
                               f"the same boundary-layer SST workload at {ncs}x60x{ncs}")
        elif a.workload.startswith("channel"):
            cpu = cpu_baseline(g, q, prm, a.cpu_sample_steps, "the same workload")

    conv = None
    if not a.no_converge and not unsteady and a.workload in ("c5", "channel128"):
        from flowpsi_amd import converge, workloads
        if a.workload == "c5":
            nn = a.converge_n
            gb, qb, pb = workloads.c5_bl(nx=nn, ny=60, nz=nn)
            cname = (f"SST boundary-layer grid {nn}x60x{nn} ({gb.nc} cells), line SGS, CFL "
                     f"{pb['cflmax']:g}, urelax {pb['urelax']:g}")
        else:
            gb, qb, pb = workloads.box(32, sweep="persistent")
            cname = "Euler box 32^3 (32768 cells), Roe, CFL 10, SGS x5 (fixed paths)"
        log(f"time to 1e-8 drop on the {cname}")
        if a.workload == "c5":
            # line SGS: the tolerance paths stop once the linear residual has
            # halved; fp32_max3 caps flowPsi's own loop at LSGSMaxIter = 3
            conv = converge.compare(gb, qb, dict(pb, sweep=a.sweep), rtol=0.5, log=log,
                                    paths=("fp64_ref", "fp32_ref", "fp32_max3", "fp32_tol", "fp32_nn"),
                                    sweep_fp32=a.sweep_fp32)
        else:
            conv = converge.compare(gb, qb, dict(pb, sweep="persistent"), log=log)
        conv["workload"] = cname
        log("converge: " + ", ".join(f"{k}={v.get('time_to_target_s')}" for k, v in conv.items()
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
                   "precision": a.precision, "sweep": a.sweep, "linear_solver": prm.get("solver", "sgs"),
                   "half_sweeps_per_step": sweeps, "parallelism": "single GPU", "time_scheme": a.time_scheme,
                   "newton_iterations_per_step": iters, "lsgs_last": lin},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": tsrc, "avg_launch_ms": avg_s * 1e3, "bytes_per_launch": bytes_launch,
                     "measured_copy_gbs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs},
        "step_roofline": {"achieved": step_gbs, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_step": step_bytes},
        "residual_drop": float(np.sqrt(tot[-1] / tot[0])) if tot[0] > 0 else None,
        "cpu_baseline": cpu,
        "time_to_1e-8": conv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
    import gc
    gc.collect()
