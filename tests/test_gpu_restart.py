"""Restart continuity on the GPU, as quickTest/restart checks flowPsi
(tests.sh there: run, restart from an earlier restart file, compare): 10
steps in one run equal 5 steps, restart files written and read into a fresh
context, and 5 more steps, bit for bit, for a steady SST channel and for
BDF2 and KEC time steps (SA channel, laminar box)."""
import numpy as np
import pytest

from flowpsi_amd import h5lib, restart, workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu

try:
    h5lib.lib()
    HAVE_LIB = True
except OSError:
    HAVE_LIB = False


def _advance(s, prm, n):
    return s.unsteady(n) if prm.get("unsteady") else s.newton(n)


@pytest.mark.skipif(not HAVE_LIB, reason="libhdf5 not installed")
@pytest.mark.parametrize("case", ["steady_sst", "bdf2_sa", "kec_sa", "bdf2_box"])
def test_restart_continuity(tmp_path, case):
    if case == "steady_sst":
        g, q, prm = workloads.channel_sst(nx=5, ny=8, nz=4, sweep="persistent")
    elif case == "bdf2_box":
        g, q, prm = workloads.box(6, unsteady=True, time_scheme="bdf2", dtmax=2e-5, newton_iter=3)
    else:
        g, q, prm = workloads.channel_sa(nx=5, ny=8, nz=4, unsteady=True, time_scheme=case[:-3], dtmax=5e-5,
                                         newton_iter=2)
    print(f"\n[restart {case}] nc={g.nc}", flush=True)
    a = Solver(g, prm)
    a.set_state(q)
    _advance(a, prm, 10)
    q_full = a.get_state()
    b = Solver(g, prm)
    b.set_state(q)
    _advance(b, prm, 5)
    restart.write(str(tmp_path), "case", b, ncycle=5, stime=5 * prm.get("dtmax", 0.0))
    b.close()
    c = Solver(g, prm)
    info = restart.read(str(tmp_path), "case", c)
    assert info["ncycle"] == 5
    _advance(c, prm, 5)
    assert np.array_equal(c.get_state(), q_full)


@pytest.mark.skipif(not HAVE_LIB, reason="libhdf5 not installed")
@pytest.mark.parametrize("case", ["steady_sst", "bdf2_sa", "bdf2_box"])
def test_restart_from_oracle_state(tmp_path, case):
    """The oracle runs k steps and its state goes into restart files
    (flowVars_<case>, flowVarsPrev_<case> and the scalar files,
    src/solverRestart.loci:117-164); a fresh GPU context reads them
    (src/solverIC.loci:572-660) and continues; the oracle continues from its
    own state; the two agree to 1e-10 relative."""
    import oracle
    if case == "steady_sst":
        g, q, prm = workloads.channel_sst(nx=5, ny=8, nz=4, sweep="persistent")
    elif case == "bdf2_box":
        g, q, prm = workloads.box(6, unsteady=True, time_scheme="bdf2", dtmax=2e-5, newton_iter=3)
    else:
        g, q, prm = workloads.channel_sa(nx=5, ny=8, nz=4, unsteady=True, time_scheme="bdf2", dtmax=5e-5,
                                         newton_iter=2)
    print(f"\n[oracle restart {case}] nc={g.nc}", flush=True)
    o = oracle.Oracle(g, prm)
    unsteady = prm.get("unsteady", False)
    if unsteady:
        qa, qm1, _ = o.unsteady(q, q, 3)
    else:
        qa, _ = o.newton(q, 3)
        qm1 = None
    restart.write_state(str(tmp_path), "case", g, prm, qa, qm1, ncycle=3, stime=0.0)
    s = Solver(g, prm)
    info = restart.read(str(tmp_path), "case", s)
    assert info["ncycle"] == 3
    assert np.array_equal(s.get_state(), qa)
    _advance(s, prm, 3)
    q_gpu = s.get_state()
    if unsteady:
        q_ref, _, _ = o.unsteady(qa, qm1, 3)
    else:
        q_ref, _ = o.newton(qa, 3)
    scale = np.abs(q_ref).max(axis=1, keepdims=True)
    err = float((np.abs(q_gpu - q_ref) / scale).max())
    print(f"state err after restart + 3 steps {err:.2e}")
    assert err < 1e-10
