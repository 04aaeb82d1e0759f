"""The HIP Newton step on the reference's own grids (quickTest VOG files read
by flowpsi_amd/vog.py) against the oracle: face fluxes and residuals to
1e-12, the state after 3 Newton steps to 1e-10, under block SGS with the
per-level and dataflow sweeps and under line SGS; the shock tube (200 hexes,
the reference case's diaphragm state) advanced by BDF2 and KEC time steps."""
import numpy as np
import pytest

import oracle
from flowpsi_amd import workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / (np.abs(b).max() + 1e-300))


@pytest.mark.parametrize("name", ["testGrid", "nozzle", "periodic", "shock_tube"])
@pytest.mark.parametrize("solver,sweep", [("sgs", "per_level"), ("sgs", "persistent"), ("lsgs", "per_level")])
def test_newton_on_reference_grids(name, solver, sweep):
    g, q, prm = workloads.vog_case(name, solver=solver, sweep=sweep)
    print(f"\n[vog {name} {solver} {sweep}] nc={g.nc} nf={g.nf}", flush=True)
    o = oracle.Oracle(g, prm)
    s = Solver(g, prm)
    s.set_state(q)
    src, ifl, vfl, _ = s.residual()
    src_r, if_r, vf_r, _ = o.residual(q)
    assert rel(ifl, if_r) < 1e-12 and rel(src, src_r) < 1e-11
    q_r, h_r = o.newton(q, 3)
    hist = s.newton(3)
    err = rel(s.get_state(), q_r)
    print(f"state err after 3 steps {err:.2e}")
    assert err < 1e-10
    tot = np.array([h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist])
    assert np.allclose(tot, h_r[:, 0] + h_r[:, 1] + h_r[:, 2], rtol=1e-9)


@pytest.mark.parametrize("scheme", ["bdf2", "kec"])
def test_shock_tube_time_steps(scheme):
    g, q, prm = workloads.vog_case("shock_tube", unsteady=True, time_scheme=scheme, dtmax=1e-5, newton_iter=4,
                                   urelax=0.9)
    o = oracle.Oracle(g, prm)
    q_r, _, _ = o.unsteady(q, q, 5)
    s = Solver(g, prm)
    s.set_state(q)
    s.unsteady(5)
    q_g = s.get_state()
    err = rel(q_g, q_r)
    print(f"{scheme} shock tube state err after 5 steps {err:.2e}; max |u| {np.abs(q_g[1]).max():.1f} m/s")
    assert err < 1e-10
    assert np.abs(q_g[1]).max() > 1.0  # the diaphragm has started to move gas
