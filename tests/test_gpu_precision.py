"""FP32 Jacobians against the FP64 path on small grids of each configuration
(flowpsi_amd/precision.py): the first linear update and the state after
every Newton step.  The FP64 path itself equals the oracle to 1e-10 in the
parity tests; here the FP32 path must stay within what single-precision
blocks allow (a linear update good to a few significant digits, 1e-6 on
the state's scales, states whose error does not grow past 1e-5 over the
steps) while the residual drops the
same way in both."""
import numpy as np
import pytest

from flowpsi_amd import precision

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", precision.CONFIGS)
def test_fp32_against_fp64_small(name):
    g, q, prm, what = precision.workload(name, small=True)
    r = precision.compare(g, q, prm, 4)
    print(name, what, r["linear_update_err"], [p["state_err"] for p in r["per_step"]])
    # FP32 blocks leave a few significant digits of the update, and a
    # difference far below the state's own scale
    assert 0.0 < r["linear_update_err"] < 1e-3
    assert r["linear_update_err_state_scale"] < 1e-6
    errs = np.array([p["state_err"] for p in r["per_step"]])
    assert np.all(errs < 1e-5)
    for p in r["per_step"]:
        assert p["drop_fp32"] == pytest.approx(p["drop_fp64"], rel=1e-2)


def test_converged_states_agree():
    """The Euler box converged by both precisions: the Newton iteration acts
    as iterative refinement, so the FP32 path's converged state is the FP64
    path's within the 1e-8 relative bar."""
    g, q, prm, _ = precision.workload("c1", small=True)
    r = precision.compare(g, q, prm, 2, conv_drop=1e-11, conv_max=300)
    c = r["converged"]
    print(c)
    assert c["drop_fp64"] <= 1e-11 and c["drop_fp32"] <= 1e-11
    assert c["state_err"] < 1e-8
