"""The boundary conditions of the quickTest cases on the GPU, through the C
ABI: residual and state against the oracle on the reference's grids
(reflecting, inflow, outflow, supersonic inflow and outflow, extrapolate,
isentropic inflow, fixed mass), and the quickTest regression itself: the
cases run on the GPU and their extracted nodal values are compared with the
reference's own output files at the cases' tolerances (the cases, the
extract and the comparison: flowpsi_amd/quicktest.py; what each pins:
tests/test_quicktest.py)."""
import numpy as np
import pytest

from flowpsi_amd import grids
from flowpsi_amd import quicktest as qt
from flowpsi_amd.capi import Solver
import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / (np.abs(b).max() + 1e-300))


def _extrapolate_case():
    g, q, prm, v = qt.setup("superSonic")
    codes = {s: t for s, t in zip(["BC_1", "BC_2", "BC_3", "BC_4"], [grids.SYMMETRY] * 4)}
    codes["BC_5"] = grids.bc_code(grids.SUPERSONIC_INFLOW, 0)
    codes["BC_6"] = grids.EXTRAPOLATE
    from flowpsi_amd import vog
    g = vog.to_grid(v, codes, None, "testGrid")
    return g, q, dict(prm, limiter="venkatakrishnan"), v


CASES = {
    "inflow_outflow": lambda: qt.setup("inflow"),
    "isentropic": lambda: qt.setup("isentropic"),
    "supersonic": lambda: (lambda g, q, p, v: (g, q, dict(p, limiter="zero"), v))(*qt.setup("superSonic")),
    "extrapolate": _extrapolate_case,
    "fixed_mass_reflecting_laminar": lambda: qt.setup("fixedMass"),
    "reflecting_bdf2": lambda: (lambda g, q, p, v: (g, q, dict(p, newton_iter=3), v))(*qt.setup("shock_tube")),
}


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(CASES))
def test_boundary_conditions_match_oracle(name, prec):
    g, q, prm, v = CASES[name]()
    prm = dict(prm, precision=prec)
    print(f"\n[{name} {prec}] nc={g.nc}", flush=True)
    o = oracle.Oracle(g, prm, "f32" if prec == "fp32" else "f64")
    s = Solver(g, prm)
    try:
        s.set_state(q)
        src, ifl, vfl, _ = s.residual()
        src_r, if_r, vf_r, _ = o.residual(q)
        assert rel(ifl, if_r) < 1e-12 and rel(src, src_r) < 1e-11
        if prm.get("unsteady"):
            q_r = o.unsteady(q, q, 3)[0]
            s.unsteady(3)
        else:
            q_r, _ = o.newton(q, 5)
            s.newton(5)
        err = rel(s.get_state(), q_r)
        print(f"state err {err:.2e}")
        assert err < 1e-10
    finally:
        s.close()


@pytest.mark.parametrize("name", ["inflow", "isentropic", "fixedMass"])
def test_quicktest_regression(name):
    cols, qs = qt.run_gpu(name)
    assert np.isfinite(qs).all()
    ok, r = qt.ndiff(qt.golden(name), cols, *qt.CASES[name]["tol"])
    print(f"\n{name}: relative rms error per column (P, t, u) {r}")
    assert ok


def test_quicktest_supersonic_boundaries_with_a_limiter():
    cols, qs = qt.run_gpu("superSonic", limiter="zero")
    ok, r = qt.ndiff(qt.golden("superSonic"), cols, *qt.CASES["superSonic"]["tol"])
    print(f"\nsuperSonic (first order): {r}")
    assert ok


def test_shock_tube_full_run_matches_oracle():
    """600 BDF2 steps of 12 Newton iterations: the GPU's nodal values equal
    the oracle's (the comparison with the reference's output is the
    documented expected failure of tests/test_quicktest.py)"""
    cols, qs = qt.run_gpu("shock_tube")
    g, q, prm, v = qt.setup("shock_tube")
    q_r = oracle.Oracle(g, prm).unsteady(q, q, qt.CASES["shock_tube"]["steps"])[0]
    err = rel(qs, q_r)
    print(f"\nshock tube GPU vs oracle after 600 steps: {err:.2e}")
    assert err < 1e-9
