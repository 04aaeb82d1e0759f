"""Parity of the line-SGS solver (linear_solver = FPSI_SOLVER_LSGS, through
the C ABI) against the oracle's restatement of src/linearSolverLSGS.loci.

The unfactored diagonal, B and the line-Jacobi start are compared bitwise.
The sweeps follow the oracle's arithmetic order; the update norm is a tree
sum on the GPU and a sequential sum in the oracle, so the relaxation factor
can differ in its last bits once it is shrunk, and the solution is held to
1e-11 relative (FP64) / 1e-5 (FP32 blocks) instead.  A renumbered grid
makes the breadth-first levels differ from the cell-index levels, so the
indexed level path of the kernel runs too.
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads
from flowpsi_amd.capi import FpsiError, Solver

pytestmark = pytest.mark.gpu


def _shuffled_box():
    g, q, prm = workloads.box(6, flux="hllc")
    perm = np.random.default_rng(11).permutation(g.nc)
    g2, q2 = grids.renumber_cells(g, perm, q)
    return g2, q2, prm


def _cases():
    return {
        "box_roe": lambda: workloads.box(8),
        "channel_visc": lambda: workloads.channel(nx=6, ny=10, nz=4),
        "channel_visc_nolim": lambda: workloads.channel(nx=5, ny=8, nz=6, limiter="none"),
        "box_renumbered": _shuffled_box,
    }


@pytest.fixture(scope="module", params=list(_cases()))
def case(request):
    g, q, prm = _cases()[request.param]()
    prm = dict(prm, solver="lsgs")
    print(f"\n[lsgs case {request.param}] nc={g.nc} nf={g.nf}", flush=True)
    return request.param, g, q, prm


def test_levels(case):
    name, g, q, prm = case
    s = Solver(g, prm)
    info = s.info()
    print(f"lsgs levels {info['lsgs_nlevels']} contiguous {info['lsgs_contiguous']}; sgs levels {info['nlevels']}")
    assert 1 <= info["lsgs_nlevels"] <= g.nc
    if name == "box_renumbered":
        assert not info["lsgs_contiguous"]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_system(case, prec):
    name, g, q, prm = case
    prm = dict(prm, precision=prec)
    o = oracle.Oracle(g, prm, "f64" if prec == "fp64" else "f32")
    ref = o.assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    s.assemble()
    sy = s.get_system()
    assert np.array_equal(sy["D"], ref["Draw"])
    assert np.array_equal(sy["B"], ref["B"])
    assert np.array_equal(sy["piv"], np.tile(np.arange(5, dtype=np.int32), (g.nc, 1)))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_solve(case, prec):
    name, g, q, prm = case
    prm = dict(prm, precision=prec)
    o = oracle.Oracle(g, prm, "f64" if prec == "fp64" else "f32")
    ref = o.assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    for maxiter, reltol in ((0, 0.1), (1, 0.1), (3, 0.1), (10, 0.1), (25, 1e-6)):
        s.set_params(dict(prm, lsgs_maxiter=maxiter, lsgs_reltol=reltol))
        s.assemble()
        dq = s.linear_solve(maxiter)
        st = s.lsgs_stats()
        dq_r, info = o.lsgs(ref, maxiter=maxiter, reltol=reltol)
        scale = np.abs(dq_r).max(axis=0, keepdims=True)
        err = float((np.abs(dq - dq_r) / scale).max())
        print(f"{prec} maxiter={maxiter} reltol={reltol}: iters {st['iters']}/{info['iters']} "
              f"w {st['w']:.6g}/{info['w']:.6g} resid {st['resid']:.3e}/{info['resid']:.3e} dq err {err:.2e}")
        if maxiter == 0:
            assert np.array_equal(dq, dq_r)
        assert st["iters"] == info["iters"]
        assert st["w"] == pytest.approx(info["w"], rel=1e-13)
        assert st["resid"] == pytest.approx(info["resid"], rel=1e-10, abs=1e-300)
        assert err < (1e-11 if prec == "fp64" else 1e-5)


def test_newton_steps(case):
    name, g, q, prm = case
    nsteps = 5
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-6)):
        p = dict(prm, precision=prec)
        o = oracle.Oracle(g, p, "f64" if prec == "fp64" else "f32")
        q_r, h_r = o.newton(q, nsteps)
        s = Solver(g, p)
        s.set_state(q)
        hist = s.newton(nsteps)
        q_g = s.get_state()
        scale = np.abs(q_r).max(axis=1, keepdims=True)
        err = float((np.abs(q_g - q_r) / scale).max())
        tot_r = h_r[:, 0] + h_r[:, 1] + h_r[:, 2]
        tot = np.array([h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist])
        print(f"{prec} lsgs state err after {nsteps} steps {err:.2e}; drop {np.sqrt(tot[-1] / tot[0]):.3e}")
        assert err < tol
        assert np.allclose(tot, tot_r, rtol=1e-8)


def test_tolerance_calls_under_line_solvers():
    """The linear residual is defined under every solver; a solve to a
    tolerance needs block SGS or line SGS run as the dataflow sweep (FSGS
    and line SGS in level launches refuse it)."""
    g, q, prm = workloads.box(4)
    s = Solver(g, dict(prm, solver="lsgs"))  # level launches
    s.set_state(q)
    s.assemble()
    assert s.linear_residual() > 0
    with pytest.raises(FpsiError):
        s.linear_solve_tol(0.1, 5)
    s = Solver(g, dict(prm, solver="fsgs", sweep="persistent"))
    s.set_state(q)
    s.assemble()
    with pytest.raises(FpsiError):
        s.linear_solve_tol(0.1, 5)


def test_sgs_on_renumbered_grid():
    """Block SGS on a grid whose cell order is far from its level order."""
    g, q, prm = _shuffled_box()
    nsteps = 4
    o = oracle.Oracle(g, prm)
    q_r, _ = o.newton(q, nsteps)
    s = Solver(g, prm)
    s.set_state(q)
    s.newton(nsteps)
    q_g = s.get_state()
    scale = np.abs(q_r).max(axis=1, keepdims=True)
    err = float((np.abs(q_g - q_r) / scale).max())
    print(f"sgs renumbered state err {err:.2e}")
    assert err < 1e-10
