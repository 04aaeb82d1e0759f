"""Pins of the oracle's line SGS (src/linearSolverLSGS.loci) by what the
mathematics fixes (no GPU).

* the equation order is the breadth-first order of the interior-face graph
  from cell 0, each distance class ascending (checked against scipy's
  unweighted shortest paths);
* the relaxed iteration converges to the dense block solve;
* on a box (no periodic faces) one sweep with w = 1 is one symmetric block
  Gauss-Seidel sweep of the separately pinned SGS oracle;
* the relaxation is affine in w;
* a two-cell chain whose Gauss-Seidel diverges, iterated by hand: the
  line-Jacobi start, the forward/backward sweeps, the relaxation, the update
  norm and the shrinking of w are all exact binary fractions there.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import shortest_path

import oracle
from flowpsi_amd import grids, workloads
from test_oracle import dense_system


def _case(mk):
    if mk == "box":
        return workloads.box(4, cflmax=2.0)
    return workloads.channel(nx=4, ny=6, nz=4, cflmax=2.0)


@pytest.mark.parametrize("mk", ["box", "channel"])
def test_order_is_breadth_first(mk):
    g, q, prm = _case(mk)
    order = oracle.Oracle(g, prm).lsgs_order()
    assert np.array_equal(np.sort(order), np.arange(g.nc))
    it = g.btype == grids.INTERIOR
    A = coo_matrix((np.ones(it.sum()), (g.cl[it], g.cr[it])), shape=(g.nc, g.nc))
    d = shortest_path(A, directed=False, unweighted=True, indices=0)
    key = d[order]
    assert np.all(np.diff(key) >= 0)
    for lev in np.unique(key):
        cells = order[key == lev]
        assert np.all(np.diff(cells) > 0)


@pytest.mark.parametrize("fj", ["f64", "f32"])
@pytest.mark.parametrize("mk", ["box", "channel"])
def test_raw_diagonal_is_the_factored_one(mk, fj):
    """Draw is fluid_D itself: P^T L U of the SGS factors rebuilds it."""
    g, q, prm = _case(mk)
    o = oracle.Oracle(g, prm, fj)
    s = o.assemble(q)
    A, _ = dense_system(g, s)
    tol = 1e-12 if fj == "f64" else 1e-5
    for c in range(g.nc):
        Dc = s["Draw"][c].astype(np.float64).reshape(5, 5).T
        ref = A[5 * c:5 * c + 5, 5 * c:5 * c + 5]
        assert np.allclose(Dc, ref, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize("fj", ["f64", "f32"])
@pytest.mark.parametrize("mk", ["box", "channel"])
def test_converges_to_dense_solve(mk, fj):
    g, q, prm = _case(mk)
    o = oracle.Oracle(g, prm, fj)
    s = o.assemble(q)
    # the dense system of the stored blocks, diagonal taken unfactored
    A, b = dense_system(g, s)
    for c in range(g.nc):
        A[5 * c:5 * c + 5, 5 * c:5 * c + 5] = s["Draw"][c].astype(np.float64).reshape(5, 5).T
    x = np.linalg.solve(A, b)
    X, info = o.lsgs(s, maxiter=400, reltol=1e-14, abstol=0.0)
    assert info["iters"] < 400
    assert np.allclose(X.ravel(), x, rtol=1e-9, atol=1e-9 * np.abs(x).max())
    # zero sweeps: the line-Jacobi start D^-1 B
    X0, info0 = o.lsgs(s, maxiter=0)
    assert info0["iters"] == 0
    for c in range(g.nc):
        Dc = A[5 * c:5 * c + 5, 5 * c:5 * c + 5]
        bc = b[5 * c:5 * c + 5]
        assert np.allclose(Dc @ X0[c], bc, rtol=1e-12, atol=1e-12 * np.abs(bc).max())


def test_one_unrelaxed_sweep_is_block_sgs():
    """On a box every face joins consecutive breadth-first layers and cl < cr,
    so the breadth-first and the index order orient every face alike and one
    sweep with w = 1 is one SGS sweep (the factorisations differ: FP64
    unpivoted here, pivoted there, so only rounding separates them)."""
    g, q, prm = workloads.box(4, cflmax=2.0)
    o = oracle.Oracle(g, prm)
    s = o.assemble(q)
    X, info = o.lsgs(s, maxiter=1, w=1.0)
    ref = o.sgs(s, 1)
    assert info["iters"] == 1
    assert np.allclose(X, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    # and it is not the Jacobi start
    assert not np.allclose(X, o.sgs(s, 0), rtol=1e-6)


def test_relaxation_is_affine_in_w():
    """After one sweep X(w) = w X_sweep + (1 - w) X_forward."""
    g, q, prm = _case("channel")
    o = oracle.Oracle(g, prm)
    s = o.assemble(q)
    X1, _ = o.lsgs(s, maxiter=1, w=1.0)
    Xa, _ = o.lsgs(s, maxiter=1, w=0.25)
    Xb, _ = o.lsgs(s, maxiter=1, w=0.5)
    scale = np.abs(X1).max()
    assert np.allclose(Xa - X1, 1.5 * (Xb - X1), rtol=0, atol=1e-12 * scale)
    assert np.abs(Xb - X1).max() > 1e-6 * scale


def _chain2(a=2.0, c=2.0):
    """Two cells joined by one face: row 0 is X0 + a X1 = 1, row 1 is c X0 + X1 = 0
    (every block a multiple of the identity)."""
    x = np.linspace(0.0, 1.0, 3)
    y = np.linspace(0.0, 0.5, 2)
    X = np.stack(np.meshgrid(x, y, y, indexing="ij"))
    g = grids.structured_hex(X, {s: grids.FARFIELD for s in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")})
    assert g.nc == 2
    fi = np.flatnonzero(g.btype == grids.INTERIOR)
    assert len(fi) == 1 and g.cl[fi[0]] == 0 and g.cr[fi[0]] == 1
    I = np.eye(5).T.ravel()
    fjp = np.zeros((g.nf, 25)); fjm = np.zeros((g.nf, 25))
    fjm[fi[0]] = a * I           # U in row 0
    fjp[fi[0]] = -c * I          # L = -fjp in row 1
    sys_ = dict(fjp=fjp, fjm=fjm, Draw=np.tile(I, (2, 1)),
                B=np.array([[1.0] * 5, [0.0] * 5]))
    return g, sys_


def test_divergent_chain_by_hand():
    """Hand iteration (w0 = 1/4): start (1, 0); sweep 1 forward (1, -2),
    backward (5, -2), relaxed (2, -2), norm sqrt(5), w kept; sweep 2 forward
    (5, -10), backward (21, -10), relaxed (9, -10), norm 4 sqrt(5) = 4 x the
    first, w = 1/16; sweep 3 forward (21, -42), backward (85, -42), relaxed
    with 1/16 to (25, -42), norm 4 sqrt(5), w = 1/64."""
    g, s = _chain2()
    o = oracle.Oracle(g, workloads.BASE)
    X, info = o.lsgs(s, maxiter=0)
    assert np.array_equal(X, [[1.0] * 5, [0.0] * 5])
    X, info = o.lsgs(s, maxiter=1)
    assert np.array_equal(X, [[2.0] * 5, [-2.0] * 5])
    assert info["w"] == 0.25 and info["resid"] == pytest.approx(np.sqrt(5), rel=1e-15)
    X, info = o.lsgs(s, maxiter=2)
    assert np.array_equal(X, [[9.0] * 5, [-10.0] * 5])
    assert info["w"] == 0.0625 and info["resid"] == pytest.approx(4 * np.sqrt(5), rel=1e-15)
    X, info = o.lsgs(s, maxiter=3)
    assert np.array_equal(X, [[25.0] * 5, [-42.0] * 5])
    assert info["w"] == 0.015625 and info["iters"] == 3


def test_stopping_rules():
    """Uncoupled cells: the first sweep changes nothing, the norm is 0 and the
    solve stops after one sweep; an abstol above the first norm also stops it."""
    g, s = _chain2(a=0.0, c=0.0)
    o = oracle.Oracle(g, workloads.BASE)
    X, info = o.lsgs(s, maxiter=10)
    assert info["iters"] == 1 and info["resid"] == 0.0
    assert np.array_equal(X, [[1.0] * 5, [0.0] * 5])
    g, s = _chain2(a=0.1, c=0.1)
    _, info = o.lsgs(s, maxiter=10, abstol=1.0, reltol=0.0)
    assert info["iters"] == 1
    _, info = o.lsgs(s, maxiter=10, abstol=0.0, reltol=1e-3)
    # contraction 0.01 per sweep: the relative norm is 1, 1e-2 (x relax), ...
    assert 2 <= info["iters"] < 10
