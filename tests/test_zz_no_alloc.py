"""fpsi_newton_step_noalloc against fpsi_newton_step through the C-ABI: the
same Newton steps from the same state give the same norm history and the
same state, bit for bit, on two calls in a row on one context, the second
with fewer steps (the library's fixed per-GPU history array then still
holds the first call's longer history), and again on a second, smaller
grid.  Step counts beyond the array's maximum, and contexts with profiling
on, are refused before anything runs."""
import numpy as np
import pytest

from flowpsi_amd import workloads
from flowpsi_amd.capi import NOALLOC_MAX_STEPS, FpsiError, Solver

pytestmark = pytest.mark.gpu


def _solver(g, q, prm):
    s = Solver(g, prm)
    s.set_state(q)
    return s


CASES = [
    # line SGS with SST on the boundary-layer grid (dataflow sweeps), then
    # the 8^3 Euler box under block SGS: the second input is smaller
    ("bl_sst_lsgs", lambda: workloads.c5_bl(nx=10, ny=24, nz=8, sweep="persistent"), (5, 2)),
    ("box8_sgs", lambda: workloads.box(8), (4, 1)),
]


@pytest.mark.parametrize("name,make,steps", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_newton_step(name, make, steps):
    g, q, prm = make()
    ref, na = _solver(g, q, prm), _solver(g, q, prm)
    try:
        for n in steps:
            h_ref = ref.newton(n)
            h_na = na.newton_noalloc(n)
            print(f"\n{name}: {n} steps, rtrms {[h['rtrms'] for h in h_na]}")
            assert len(h_na) == n
            assert h_na == h_ref
            assert np.array_equal(na.get_state(), ref.get_state())
    finally:
        ref.close()
        na.close()


def test_noalloc_refuses_oversize_and_profiling():
    g, q, prm = workloads.box(6)
    s = _solver(g, q, prm)
    try:
        q0 = s.get_state()
        with pytest.raises(FpsiError, match="FPSI_ERR_ARG"):
            s.newton_noalloc(NOALLOC_MAX_STEPS + 1)
        assert np.array_equal(s.get_state(), q0)
        s.profile(True)
        with pytest.raises(FpsiError, match="FPSI_ERR_STATE"):
            s.newton_noalloc(1)
        s.profile(False)
        assert len(s.newton_noalloc(1)) == 1
    finally:
        s.close()
