"""Pins of the oracle's FSGS (src/linearSolverFSGS.loci) by what the
mathematics fixes (no GPU).

* a two-cell chain iterated by hand from X = 0 (exact binary fractions):
  no Jacobi start, no relaxation, lower-then-upper couplings, block and
  scalar alike;
* the sweeps reach the dense block / scalar solve;
* the diagonal is factored in the Jacobian precision (FP32 factors move the
  result by float rounding only).
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads
from test_oracle import dense_system
from test_oracle_lsgs import _chain2
from test_oracle_sa import _scalar_chain


def test_block_chain_by_hand():
    """Rows X0 + 2 X1 = 1, 2 X0 + X1 = 0 (blocks multiples of I), from zero:
    sweep 1 forward (1, -2), backward (5, -2); sweep 2 forward (5, -10),
    backward (21, -10)."""
    g, s = _chain2()
    o = oracle.Oracle(g, dict(workloads.BASE, solver="fsgs"))
    assert np.array_equal(o.fsgs(s, 0), np.zeros((2, 5)))
    assert np.array_equal(o.fsgs(s, 1), [[5.0] * 5, [-2.0] * 5])
    assert np.array_equal(o.fsgs(s, 2), [[21.0] * 5, [-10.0] * 5])


def test_scalar_chain_by_hand():
    g, s = _scalar_chain()
    o = oracle.Oracle(g, workloads.BASE)
    assert np.array_equal(o.sa_fsgs(s, 0), [0.0, 0.0])
    assert np.array_equal(o.sa_fsgs(s, 1), [5.0, -2.0])
    assert np.array_equal(o.sa_fsgs(s, 2), [21.0, -10.0])


@pytest.mark.parametrize("mk", ["box", "channel"])
def test_block_converges_to_dense_solve(mk):
    if mk == "box":
        g, q, prm = workloads.box(4, cflmax=2.0)
    else:
        g, q, prm = workloads.channel(nx=4, ny=6, nz=4, cflmax=2.0)
    o = oracle.Oracle(g, dict(prm, solver="fsgs"))
    s = o.assemble(q)
    A, b = dense_system(g, s)
    x = np.linalg.solve(A, b)
    X = o.fsgs(s, 300).ravel()
    assert np.allclose(X, x, rtol=1e-9, atol=1e-9 * np.abs(x).max())


def test_scalar_converges_to_dense_solve():
    g, q, prm = workloads.channel_sa(nx=4, ny=6, nz=4, cflmax=2.0)
    o = oracle.Oracle(g, prm)
    s = o.sa_assemble(q)
    A = np.diag(s["D"])
    for f in range(g.nf):
        if g.btype[f] == grids.INTERIOR:
            A[g.cl[f], g.cr[f]] += s["fjm"][f]
            A[g.cr[f], g.cl[f]] += -s["fjp"][f]
        elif g.btype[f] == grids.PERIODIC:
            A[g.cl[f], g.cl[g.pface[f]]] += s["fjm"][f]
    x = np.linalg.solve(A, s["B"])
    assert np.allclose(o.sa_fsgs(s, 400), x, rtol=1e-9, atol=1e-9 * np.abs(x).max())


def test_fp32_factors_differ_from_fp64():
    """The diagonal is factored in the Jacobian precision: with real_fj =
    float the result departs from the FP64 build by float rounding, not
    more (both converge to the same solve)."""
    g, q, prm = workloads.box(4, cflmax=2.0)
    p = dict(prm, solver="fsgs")
    s64 = oracle.Oracle(g, p, "f64").assemble(q)
    s32 = oracle.Oracle(g, p, "f32").assemble(q)
    x64 = oracle.Oracle(g, p, "f64").fsgs(s64, 5)
    x32 = oracle.Oracle(g, p, "f32").fsgs(s32, 5)
    d = np.abs(x32 - x64).max() / np.abs(x64).max()
    assert 1e-9 < d < 1e-4
