"""Line SGS and FSGS as one cooperative dataflow launch per half-sweep
(sweep = persistent, flag-free; persistent_flags, chunk flags) against one
launch per level: the same bits for the
solution, the line-SGS iteration count, relaxation and update norm, and
the state after Newton steps, in FP64 and FP32 blocks, on grids whose
breadth-first levels are the SGS levels (where the dataflow runs; with a
turbulence model the scalars' line-SGS loops run in the same launches) and
on a tetrahedral grid where they are not (where the persistent mode falls
back to the level launches)."""
import numpy as np
import pytest

from flowpsi_amd import workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu


def _cases():
    return {
        "channel": lambda: workloads.channel(nx=24, ny=20, nz=12),
        "bl_sst": lambda: workloads.c5_bl(nx=12, ny=30, nz=10),
        "box40": lambda: workloads.box(40),
        "tets": lambda: workloads.c3_tets(5, precision="fp64"),
        "channel_sa": lambda: workloads.channel_sa(nx=10, ny=16, nz=6),
        # prism-to-hex transition cells have rows wider than the padded six
        "hybrid": lambda: workloads.c4_hybrid(nx=6, nz=5, n_prism=4, n_hex=5),
    }


@pytest.mark.parametrize("solver", ["lsgs", "fsgs"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(_cases()))
def test_dataflow_line_sweeps_bitwise(name, prec, solver):
    g, q, prm = _cases()[name]()
    prm = dict(prm, solver=solver, precision=prec, lsgs_maxiter=6, sgs_iter=4, lsgs_reltol=1e-3)
    print(f"\n[{name} {solver} {prec}] nc={g.nc}", flush=True)
    out = {}
    for sweep in ("per_level", "persistent", "persistent_flags"):
        s = Solver(g, dict(prm, sweep=sweep))
        info = s.info()
        s.set_state(q)
        s.assemble()
        dq = s.linear_solve(prm["lsgs_maxiter"] if solver == "lsgs" else prm["sgs_iter"])
        st = s.lsgs_stats() if solver == "lsgs" else None
        hist = s.newton(2)
        # turbulence scalars: their line-SGS loops of the last Newton step
        tst = [s.turb_lsgs_stats(j) for j in range(s.nvar - 5)] if solver == "lsgs" else None
        out[sweep] = (dq, st, s.get_state(), hist, info, tst)
        s.close()
    print(f"lsgs_dataflow {out['persistent'][4]['lsgs_dataflow']}", flush=True)
    if name not in ("tets", "hybrid"):
        assert out["persistent"][4]["lsgs_dataflow"] == 1
    dq0, st0, q0, h0, _, t0 = out["per_level"]
    for mode in ("persistent", "persistent_flags"):
        dq1, st1, q1, h1, _, t1 = out[mode]
        assert t0 == t1, mode
        assert np.array_equal(dq0, dq1), mode
        assert st0 == st1, mode
        assert np.array_equal(q0, q1), mode
        assert h0 == h1, mode
