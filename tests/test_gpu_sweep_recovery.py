"""A dataflow sweep whose dependency wait runs past its spin limit ends
early and the call returns FPSI_ERR_HIP; the context then recovers: the
error word is cleared, the system must be assembled again, and the next
solve gives the same bits as a context that never failed.  The failure is
forced from the host with a spin limit of one poll (every wait in the
kernels is bounded by it, so every wave exits normally)."""
import numpy as np
import pytest

from flowpsi_amd import workloads
from flowpsi_amd.capi import FpsiError, Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["sgs", "lsgs", "lsgs_sst"])
def test_spin_limit_error_then_recovery(name):
    if name == "lsgs_sst":
        g, q, prm = workloads.c5_bl(nx=16, ny=30, nz=12)
    else:
        g, q, prm = workloads.box(24, solver=name)
    prm = dict(prm, sweep="persistent", lsgs_maxiter=4, sgs_iter=3)
    print(f"\n[{name}] nc={g.nc}", flush=True)
    ref = Solver(g, prm)
    ref.set_state(q)
    ref.assemble()
    dq0 = ref.linear_solve(3)
    h0 = ref.newton(2)
    q0 = ref.get_state()
    ref.close()

    s = Solver(g, dict(prm, sweep_spin_limit=1))
    s.set_state(q)
    s.assemble()
    with pytest.raises(FpsiError, match="dependency wait"):
        s.linear_solve(3)
    # without a new assembly the solve is refused
    with pytest.raises(FpsiError):
        s.linear_solve(3)
    s.set_params(prm)
    s.assemble()
    dq = s.linear_solve(3)
    assert np.array_equal(dq, dq0)
    h = s.newton(2)
    assert h == h0
    assert np.array_equal(s.get_state(), q0)
    s.close()
