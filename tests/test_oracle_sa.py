"""Pins of the oracle's Spalart-Allmaras path (turbulence/SA/SA_turb.loci,
src/solverScalarTransport.loci) by what the mathematics fixes (no GPU).

* the cell terms against the published baseline model (Spalart & Allmaras,
  AIAA 92-0439, without ft2), written here from the paper's formulas:
  chi = nuTilde/nu, fv1 = chi^3/(chi^3+cv1^3), fv2 = 1 - chi/(1+chi fv1),
  S~ = max(Omega + nuTilde fv2/(k d)^2, 0.3 Omega), r = min(nuTilde/(S~ (k d)^2), 10),
  g = r + cw2 (r^6 - r), fw = g ((1+cw3^6)/(g^6+cw3^6))^(1/6),
  source = rho V (cb1 S~ nuTilde - cw1 fw (nuTilde/d)^2), tmu = rho nuTilde fv1;
* the destruction Jacobian against a central difference of the destruction
  term in nuTilde (away from the S~ and r clips, where it is exact);
* scalar SGS converges to the dense solve of the assembled nuTilde system;
* nuTilde = 0 gives tmu = 0 and the fluid residual of the laminar path bit
  for bit (the eddy-viscosity coupling adds exactly nothing);
* the projected wall distance of the channel equals h - |y|;
* scalar line SGS (src/linearSolverLSGS.loci:1112-1404): it reaches the
  dense solve; on a box one unrelaxed sweep is one scalar SGS sweep; and a
  two-cell chain iterated by hand, where the relaxation blends with the
  value before the forward sweep (the block solver blends with the value
  before the backward sweep, and the chain tells the two apart).
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads

CB1, K, CV1, CB2, CW2, CW3, SIG = 0.1355, 0.41, 7.1, 0.622, 0.3, 2.0, 2.0 / 3.0
CW1 = CB1 / K ** 2 + (1 + CB2) / SIG
R = 287.0


def mu_suth(T):
    return 1.458e-6 * T ** 1.5 / (T + 110.4)


def sa_textbook(T, pg, nut, Om, d, V):
    rho = pg / (R * T)
    chi = nut * rho / mu_suth(T)
    fv1 = chi ** 3 / (chi ** 3 + CV1 ** 3)
    fv2 = 1 - chi / (1 + chi * fv1)
    S = max(Om + nut * fv2 / (K * d) ** 2, 0.3 * Om)
    r = min(nut / (S * (K * d) ** 2), 10.0)
    g = r + CW2 * (r ** 6 - r)
    fw = g * ((1 + CW3 ** 6) / (g ** 6 + CW3 ** 6)) ** (1 / 6)
    prod = rho * V * CB1 * S * nut
    dest = rho * V * CW1 * fw * (nut / d) ** 2
    return rho * nut * fv1, prod - dest, dest


def states(rng, n):
    for _ in range(n):
        T = rng.uniform(250, 350)
        pg = rng.uniform(0.8e5, 1.2e5)
        nu = mu_suth(T) / (pg / (R * T))
        nut = nu * 10 ** rng.uniform(-1, 2)
        Om = 10 ** rng.uniform(1, 4)
        d = 10 ** rng.uniform(-4, -2)
        yield T, pg, nut, Om, d, rng.uniform(1e-9, 1e-6)


def test_cell_terms_match_the_published_model():
    rng = np.random.default_rng(31)
    for T, pg, nut, Om, d, V in states(rng, 200):
        tmu, src, _ = oracle.sa_point(workloads.BASE, T, pg, nut, Om, d, V)
        rt, rs, _ = sa_textbook(T, pg, nut, Om, d, V)
        assert tmu == pytest.approx(rt, rel=1e-12)
        assert src == pytest.approx(rs, rel=1e-9, abs=1e-12 * abs(rs))


def test_destruction_jacobian_is_its_derivative():
    rng = np.random.default_rng(32)
    checked = 0
    for T, pg, nut, Om, d, V in states(rng, 400):
        rho = pg / (R * T)
        nu = mu_suth(T) / rho
        chi = nut / nu
        fv2 = 1 - chi / (1 + chi * chi ** 3 / (chi ** 3 + CV1 ** 3))
        S = Om + nut * fv2 / (K * d) ** 2
        if S < 0.4 * Om or nut / (S * (K * d) ** 2) > 5:
            continue  # keep away from the clips of S~ and r
        _, _, jac = oracle.sa_point(workloads.BASE, T, pg, nut, Om, d, V)
        h = 1e-6 * nut
        dp = sa_textbook(T, pg, nut + h, Om, d, V)[2]
        dm = sa_textbook(T, pg, nut - h, Om, d, V)[2]
        fd = -(dp - dm) / (2 * h)
        assert jac == pytest.approx(fd, rel=1e-5)
        checked += 1
    assert checked > 50


def test_wall_distance_of_the_channel():
    g, q, prm = workloads.channel_sa(nx=4, ny=12, nz=4)
    h = g.meta["h"]
    assert np.allclose(g.meta["dist"], h - np.abs(g.cc[1]), rtol=0, atol=1e-12 * h)


def test_scalar_sgs_converges_to_dense_solve():
    g, q, prm = workloads.channel_sa(nx=4, ny=6, nz=4, cflmax=2.0)
    o = oracle.Oracle(g, prm)
    s = o.sa_assemble(q)
    A = np.diag(s["D"])
    for f in range(g.nf):
        if g.btype[f] == grids.INTERIOR:
            l, r = g.cl[f], g.cr[f]
            A[l, r] += s["fjm"][f]
            A[r, l] += -s["fjp"][f]
        elif g.btype[f] == grids.PERIODIC:
            A[g.cl[f], g.cl[g.pface[f]]] += s["fjm"][f]
    x = np.linalg.solve(A, s["B"])
    X = o.sa_sgs(s, 300)
    assert np.allclose(X, x, rtol=1e-9, atol=1e-9 * np.abs(x).max())
    assert np.allclose(o.sa_sgs(s, 0), s["B"] / s["D"], rtol=1e-15)
    # the assembled system is diagonally dominant with positive diagonal
    assert np.all(s["D"] > 0)


def test_zero_nutilde_is_laminar():
    g, q, prm = workloads.channel_sa(nx=5, ny=8, nz=4)
    q0 = q.copy()
    q0[5] = 0.0
    o = oracle.Oracle(g, prm)
    ol = oracle.Oracle(g, dict(prm, turb="none"))
    assert np.all(o.sa_assemble(q0)["tmu"] == 0.0)
    r_sa = o.residual(q0)
    r_lam = ol.residual(q0[:5])
    assert np.array_equal(r_sa[0], r_lam[0])
    a_sa = o.assemble(q0)
    a_lam = ol.assemble(q0[:5])
    for k in ("fjp", "fjm", "D", "B", "dtau"):
        assert np.array_equal(a_sa[k], a_lam[k])


def _scalar_chain(a=2.0, c=2.0):
    """Two cells joined by one face: row 0 is X0 + a X1 = 1, row 1 is c X0 + X1 = 0."""
    x = np.linspace(0.0, 1.0, 3)
    y = np.linspace(0.0, 0.5, 2)
    Xn = np.stack(np.meshgrid(x, y, y, indexing="ij"))
    g = grids.structured_hex(Xn, {s: grids.FARFIELD for s in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")})
    fi = np.flatnonzero(g.btype == grids.INTERIOR)
    assert len(fi) == 1 and g.cl[fi[0]] == 0 and g.cr[fi[0]] == 1
    fjp = np.zeros(g.nf); fjm = np.zeros(g.nf)
    fjm[fi[0]] = a        # U in row 0
    fjp[fi[0]] = -c       # L = -fjp in row 1
    return g, dict(fjp=fjp, fjm=fjm, D=np.ones(2), B=np.array([1.0, 0.0]))


def test_scalar_lsgs_chain_by_hand():
    """w0 = 1/4.  Start (1, 0).  Sweep 1: forward (1, -2), backward (5, -2),
    blended with the start (1, 0) to (2, -1/2), norm sqrt(5)/2.  Sweep 2:
    forward (2, -4), backward (9, -4), blended with (2, -1/2) to
    (15/4, -11/8), norm 7 sqrt(5)/8 = 1.75 x the first, so w = 1/7."""
    g, s = _scalar_chain()
    o = oracle.Oracle(g, workloads.BASE)
    X, info = o.sa_lsgs(s, maxiter=0)
    assert np.array_equal(X, [1.0, 0.0])
    X, info = o.sa_lsgs(s, maxiter=1)
    assert np.array_equal(X, [2.0, -0.5])
    assert info["w"] == 0.25 and info["resid"] == pytest.approx(np.sqrt(5) / 2, rel=1e-15)
    X, info = o.sa_lsgs(s, maxiter=2)
    assert np.array_equal(X, [3.75, -1.375])
    assert info["w"] == pytest.approx(1 / 7, rel=1e-14) and info["iters"] == 2
    # uncoupled cells: the first sweep changes nothing and the solve stops
    g, s = _scalar_chain(0.0, 0.0)
    X, info = o.sa_lsgs(s, maxiter=10)
    assert info["iters"] == 1 and info["resid"] == 0.0 and np.array_equal(X, [1.0, 0.0])


def test_scalar_lsgs_converges_to_dense_solve():
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
    X, info = o.sa_lsgs(s, maxiter=400, abstol=0.0, reltol=0.0, w=1.0)
    assert np.allclose(X, x, rtol=1e-9, atol=1e-9 * np.abs(x).max())


def test_scalar_lsgs_unrelaxed_sweep_is_scalar_sgs_on_box():
    """On a box the breadth-first order is a topological order of cl -> cr,
    so one sweep with w = 1 equals one scalar SGS sweep (random
    diagonally dominant system on the box's faces)."""
    g, q, prm = workloads.box(5)
    rng = np.random.default_rng(33)
    it = g.btype == grids.INTERIOR
    fjp = np.where(it, -rng.uniform(0.1, 1.0, g.nf), 0.0)
    fjm = np.where(it, rng.uniform(-1.0, -0.1, g.nf), 0.0)
    s = dict(fjp=fjp, fjm=fjm, D=rng.uniform(7.0, 9.0, g.nc), B=rng.standard_normal(g.nc))
    o = oracle.Oracle(g, prm)
    X, info = o.sa_lsgs(s, maxiter=1, w=1.0)
    ref = o.sa_sgs(s, 1)
    assert info["iters"] == 1
    assert np.allclose(X, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert not np.allclose(X, o.sa_sgs(s, 0), rtol=1e-6)
