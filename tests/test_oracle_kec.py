"""Pins of the oracle's KEC time scheme (timeStepScheme KEC: a Crank-Nicolson
variant whose Newton iterate is the mid-step state q*,
src/solverTimestep.loci:340-396, src/solverLinearSystem.loci:226-378,
src/solverNewton.loci:93-157, src/solverScalarTransport.loci:367-395,
514-539, 592-600, 691-700) by what the mathematics fixes (no GPU):

* the mid-step state is the harmonic mean of T, the (1 + eps, 1 - eps)
  weighted mean of p and the Roe (sqrt(rho)-weighted) mean of u of the levels
  n and n+1: the right-hand side at q* equals src(q*) - V/dt (Q^(n+1) - Q^n)
  with q^(n+1) built forward from those means (eps fixed by
  timePressureBiasControl);
* at q* = q^n the time block is V/dtau dQ/dq . (dq*/dq^(n+1))^-1, the
  inverse of the finite-difference Jacobian of the forward means;
* the scalar level n+1 is the midpoint extrapolation 2 X* - X^n when X* >=
  X^n and (3 X* - X^n)^2 / (4 X*) below (positive, C1 at X* = X^n); the
  scalar rows are the midpoint rule in rho X with the stabilisation of X^(n+1)
  and the diagonal gains rhodot and (2 rho^n - rho) V/dtau;
* converged steps satisfy the discrete midpoint rule V (Q^(n+1) - Q^n) / dt =
  src(q*) cell by cell, and converge at second order in dt.
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import workloads
from test_oracle_bdf2 import blocks, cons, dQdq, levels

R, G = 287.0, 1.4


def star(qn, qp, eps):
    """mid-step state of levels qn and qp (rows T, u, v, w, p)."""
    Ts = 2.0 * qn[0] * qp[0] / (qn[0] + qp[0])
    ps = ((1 + eps) * qp[4] + (1 - eps) * qn[4]) / 2.0
    rn, rp = qn[4] / (R * qn[0]), qp[4] / (R * qp[0])
    sn, sp = np.sqrt(rn), np.sqrt(rp)
    us = (sp * qp[1:4] + sn * qn[1:4]) / (sp + sn)
    return np.vstack([Ts, us, ps])


def test_rhs_is_the_midpoint_rule():
    g, q, prm = workloads.box(6)
    dt = 1e-5
    rng = np.random.default_rng(7)
    qn = q * (1.0 + 1e-3 * rng.standard_normal(q.shape))
    qp = q * (1.0 + 1e-3 * rng.standard_normal(q.shape))
    for eps in (0.0, 0.3):
        o = oracle.Oracle(g, dict(prm, unsteady=True, time_scheme="kec", dtmax=dt, pbias_control=eps))
        qs = star(qn, qp, eps)
        a = o.assemble(qs, qn)
        assert np.all(a["pbias"] == max(eps, 1e-40 ** 0.7))
        src = o.residual(qs)[0]
        ref = src - g.vol / dt * (cons(qp) - cons(qn))
        scale = np.abs(src).max(axis=1, keepdims=True) + np.abs(g.vol / dt * cons(qp)).max(axis=1, keepdims=True)
        err = np.abs(a["B"].T - ref) / scale
        assert err.max() < 1e-11, (eps, err.max())


def test_time_block_at_the_first_iterate():
    g, q, prm = workloads.box(5)
    base = dict(prm, cflmax=7.0, urelax=1e30)
    o = oracle.Oracle(g, dict(base, dtmax=1e-5, unsteady=True, time_scheme="kec", pbias_control=0.2))
    D0 = blocks(oracle.Oracle(g, dict(base, cflmax=0.0, dtmax=1e30)).assemble(q)["Draw"])
    a = o.assemble(q, q)
    T_kec = blocks(a["Draw"]) - D0
    # d q* / d q^(n+1) at q^(n+1) = q^n by central differences of the means
    J = np.zeros((g.nc, 5, 5))
    for j in range(5):
        h = 1e-6 * (np.abs(q[1:4]).max() if 1 <= j <= 3 else np.abs(q[j]).max())
        qp, qm = q.copy(), q.copy()
        qp[j] += h
        qm[j] -= h
        J[:, :, j] = ((star(q, qp, 0.2) - star(q, qm, 0.2)) / (2 * h)).T
    ref = (g.vol / a["dtau"])[:, None, None] * dQdq(q) @ np.linalg.inv(J)
    scale = np.abs(ref).max(axis=(1, 2), keepdims=True)
    err = (np.abs(T_kec - ref) / scale).max()
    assert err < 1e-7, err


@pytest.mark.parametrize("turb", ["sa", "sst"])
def test_scalar_level_and_rows(turb):
    xs = np.array([0.3, 0.9, 1.0, 1.2, 5.0])
    xn = 1.0
    fn = workloads.channel_sa if turb == "sa" else workloads.channel_sst
    g, q, prm = fn(nx=4, ny=8, nz=4)
    dt = 2e-5
    qn, _ = levels(q, 9, amp=5e-2)
    ou = oracle.Oracle(g, dict(prm, unsteady=True, time_scheme="kec", dtmax=dt))
    os_ = oracle.Oracle(g, dict(prm, dtmax=dt))
    asm = "sa_assemble" if turb == "sa" else "sst_assemble"
    su, ss = getattr(ou, asm)(q, qn), getattr(os_, asm)(q)
    dtau = os_.assemble(q[:5] if turb == "sa" else q)["dtau"]
    rhodot = os_.residual(q)[0][0]
    rho, rn = (x[4] / (R * x[0]) for x in (q, qn))
    for j in range(1 if turb == "sa" else 2):
        pick = (lambda a: a) if turb == "sa" else (lambda a: a[j])
        X, Xn = q[5 + j], qn[5 + j]
        assert (X < Xn).any() and (X > Xn).any()
        Xp = np.where(X >= Xn, 2 * X - Xn, (3 * X - Xn) ** 2 / (4 * X))
        src = pick(ss["src"]) - (Xp - X) * rhodot
        assert np.allclose(pick(su["src"]), src, rtol=1e-12, atol=1e-12 * np.abs(src).max())
        tt = g.vol / dt * (Xp - Xn) * rn
        B = -(tt - src)
        assert np.abs(pick(su["B"]) - B).max() <= 1e-11 * (np.abs(tt) + np.abs(src)).max()
        D = pick(ss["D"]) + rhodot + g.vol / dtau * (2 * rn - rho)
        sc = (np.abs(pick(ss["D"])) + np.abs(rhodot) + g.vol / dtau * rho).max()
        assert np.abs(pick(su["D"]) - D).max() <= 1e-12 * sc
    # the two branches of the level map meet with equal value and slope
    f = lambda x: np.where(x >= xn, 2 * x - xn, (3 * x - xn) ** 2 / (4 * x))
    h = 1e-7
    assert f(xn) == xn
    assert abs((f(xn + h) - f(xn - h)) / (2 * h) - 2.0) < 1e-6
    assert np.all(f(xs) > 0)


def test_converged_steps_are_the_midpoint_rule_and_second_order():
    g, q, prm = workloads.box(6)
    prm = dict(prm, limiter="none", unsteady=True, time_scheme="kec", newton_iter=14, sgs_iter=20,
               cflmax=100.0)
    ends = {}
    for k in (4, 8, 16):
        dt = 1e-4 / k
        o = oracle.Oracle(g, dict(prm, dtmax=dt))
        if k == 4:
            q1, _, h = o.unsteady(q, q, 1)
            assert np.sqrt(h[0, -1, 0]) < 1e-10 * np.sqrt(h[0, 0, 0])
            qs = star(q, q1, 1e-40 ** 0.7)
            lhs = g.vol * (cons(q1) - cons(q)) / dt
            src = o.residual(qs)[0]
            err = np.abs(lhs - src) / np.abs(lhs).max(axis=1, keepdims=True)
            assert err.max() < 1e-8, err.max()
        ends[k], _, _ = o.unsteady(q, q, 4 * k)
    d1, d2 = ends[4] - ends[8], ends[8] - ends[16]
    for i in range(5):
        ratio = np.abs(d1[i]).max() / np.abs(d2[i]).max()
        assert 3.5 < ratio < 4.7, (i, ratio)
