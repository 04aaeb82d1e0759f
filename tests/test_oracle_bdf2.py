"""Pins of the oracle's time-accurate BDF2 path (timeStepMode unsteady,
src/solverTimestep.loci:400-428 rhs, src/solverLinearSystem.loci:384-434
fluid_D, src/solverNewton.loci:160-198 update,
src/solverScalarTransport.loci:534-539, 602-616, 702-716 scalars) by what
the mathematics fixes (no GPU):

* the fluid right-hand side is the steady residual minus V/(2 dt) (3 Q -
  4 Q^n + Q^(n-1)) in the conservative variables rho, rho u, rho e0 of the
  three time levels (the textbook BDF2 derivative);
* the time block of the diagonal is 1.5 V/dtau x the textbook Jacobian
  dQ/dq of the conservative variables for an unpreconditioned blend
  (pseudo_cflmax -> inf), 1.5 x the steady preconditioned block for
  pseudo_cflmax = 0, and affine in the blended 1/eta in between;
* the scalar rows are the BDF2 derivative of rho X with the new density
  extrapolated by the continuity equation, rho* = (4 rho^n - rho^(n-1))/3,
  and the diagonal gains 1.5 rho* V/dtau and 0.5 x the mass imbalance
  (the stabilisation -1.5 rhodot instead of the steady -rhodot);
* the residual norms are the volume-weighted RMS of the BDF2 right-hand side;
* converged dual time stepping is second order in dt from a start pair of
  third-order accuracy (Richardson ratio 4), and loses an order from the
  constant start q^(-1) = q^0 that the reference uses.
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import workloads

R, G = 287.0, 1.4


def cons(q):
    rho = q[4] / (R * q[0])
    return np.vstack([rho, rho * q[1:4], rho * (R * q[0] / (G - 1) + 0.5 * (q[1:4] ** 2).sum(0))])


def prim(Q):
    rho = Q[0]
    u = Q[1:4] / rho
    T = (Q[4] / rho - 0.5 * (u ** 2).sum(0)) * (G - 1) / R
    return np.vstack([T, u, rho * R * T])


def dQdq(q):
    """[nc, 5, 5] Jacobian of (rho, rho u, rho e0) with respect to
    (T, u, v, w, p) for rho = p / (R T), e0 = R T / (G - 1) + |u|^2 / 2."""
    T, u, p = q[0], q[1:4], q[4]
    rho = p / (R * T)
    k = 0.5 * (u ** 2).sum(0)
    J = np.zeros((q.shape[1], 5, 5))
    J[:, 0, 0] = -rho / T
    J[:, 0, 4] = 1.0 / (R * T)
    for i in range(3):
        J[:, 1 + i, 0] = -rho * u[i] / T
        J[:, 1 + i, 1 + i] = rho
        J[:, 1 + i, 4] = u[i] / (R * T)
        J[:, 4, 1 + i] = rho * u[i]
    J[:, 4, 0] = -rho * k / T
    J[:, 4, 4] = 1.0 / (G - 1) + k / (R * T)
    return J


def blocks(D):
    """[nc, 25] column-major blocks as [nc, 5, 5] (row i, column j)."""
    return np.asarray(D, float).reshape(-1, 5, 5).transpose(0, 2, 1)


def levels(q, seed, amp=2e-3):
    rng = np.random.default_rng(seed)
    qn = q * (1.0 + amp * rng.standard_normal(q.shape))
    qm1 = q * (1.0 + amp * rng.standard_normal(q.shape))
    return qn, qm1


def test_fluid_rhs_is_the_bdf2_derivative():
    g, q, prm = workloads.box(6)
    dt = 3e-5
    o = oracle.Oracle(g, dict(prm, unsteady=True, dtmax=dt))
    qn, qm1 = levels(q, 1)
    a = o.assemble(q, qn, qm1)
    src = o.residual(q)[0]
    ref = src - g.vol / (2 * dt) * (3 * cons(q) - 4 * cons(qn) + cons(qm1))
    scale = np.abs(src).max(axis=1, keepdims=True) + np.abs(g.vol / dt * cons(q)).max(axis=1, keepdims=True)
    assert np.all(np.abs(a["B"].T - ref) <= 1e-12 * scale)
    assert np.array_equal(a["src"], src)
    # the steady parts of the system do not see the time levels
    s = oracle.Oracle(g, dict(prm, dtmax=dt)).assemble(q)
    for key in ("fjp", "fjm", "dtau"):
        assert np.array_equal(a[key], s[key])


def test_residual_norms_are_those_of_the_bdf2_rhs():
    g, q, prm = workloads.box(5)
    o = oracle.Oracle(g, dict(prm, unsteady=True, dtmax=1e-4, newton_iter=1))
    _, qm1 = levels(q, 2)
    rhs = o.assemble(q, q, qm1)["B"].T
    w = g.vol / g.vol.sum()
    r = rhs / g.vol
    _, _, hist = o.unsteady(q, qm1, 1)
    assert hist[0, 0, 0] == pytest.approx((w * r[0] ** 2).sum(), rel=1e-12)
    assert hist[0, 0, 1] == pytest.approx((w * r[4] ** 2).sum(), rel=1e-12)
    assert hist[0, 0, 2] == pytest.approx((w / 3 * (r[1:4] ** 2).sum(0)).sum(), rel=1e-12)
    assert abs(hist[0, 0, 3]) == pytest.approx(np.abs(r[0]).max(), rel=1e-14)


def test_time_block_of_the_diagonal():
    g, q, prm = workloads.box(5)
    cfl = 7.0
    qn, qm1 = levels(q, 3)
    base = dict(prm, cflmax=cfl, urelax=1e30, Minf=0.05)  # low-Mach preconditioning on
    # srcJ: the steady diagonal with the time term switched off (dtau = 1e30)
    D0 = blocks(oracle.Oracle(g, dict(base, cflmax=0.0, dtmax=1e30)).assemble(q)["Draw"])
    st = oracle.Oracle(g, dict(base, dtmax=1.0)).assemble(q)
    Ds, dtau = blocks(st["Draw"]), st["dtau"]
    unst = lambda pc: oracle.Oracle(g, dict(base, dtmax=1.0, unsteady=True, pseudo_cflmax=pc))
    binf = unst(1e250).assemble(q, qn, qm1)
    assert np.array_equal(binf["dtau"], dtau)
    T_inf = blocks(binf["Draw"]) - D0
    scale = np.abs(1.5 * (g.vol / dtau)[:, None, None] * dQdq(q)).max()
    assert np.abs(T_inf - 1.5 * (g.vol / dtau)[:, None, None] * dQdq(q)).max() <= 1e-11 * scale
    T_0 = blocks(unst(0.0).assemble(q, qn, qm1)["Draw"]) - D0
    assert np.abs(T_0 - 1.5 * (Ds - D0)).max() <= 1e-11 * scale
    # pseudo_cflmax = cflmax: w = 1/2 for CFL-limited dtau, 1/eta blended
    eta = oracle.Oracle(g, base).eta(q)
    T_h = blocks(unst(cfl).assemble(q, qn, qm1)["Draw"]) - D0
    r = 1.0 / (1.0 + 0.5 * (eta - 1.0))
    m = np.abs(1.0 / eta - 1.0) > 1e-3
    assert m.sum() > 0.5 * g.nc
    lam = ((r - 1.0) / (1.0 / eta - 1.0))[m][:, None, None]
    ref = T_inf[m] + lam * (T_0[m] - T_inf[m])
    assert np.abs(T_h[m] - ref).max() <= 1e-10 * scale


@pytest.mark.parametrize("turb", ["sa", "sst"])
def test_scalar_rows(turb):
    fn = workloads.channel_sa if turb == "sa" else workloads.channel_sst
    g, q, prm = fn(nx=4, ny=8, nz=4)
    dt = 2e-5
    qn, qm1 = levels(q, 4)
    ou = oracle.Oracle(g, dict(prm, unsteady=True, dtmax=dt))
    os_ = oracle.Oracle(g, dict(prm, dtmax=dt))
    asm = "sa_assemble" if turb == "sa" else "sst_assemble"
    su, ss = getattr(ou, asm)(q, qn, qm1), getattr(os_, asm)(q)
    dtau = os_.assemble(q[:5])["dtau"] if turb == "sa" else os_.assemble(q)["dtau"]
    rhodot = os_.residual(q)[0][0]
    rho, rn, rm = (x[4] / (R * x[0]) for x in (q, qn, qm1))
    rstar = (4 * rn - rm) / 3
    ns = 1 if turb == "sa" else 2
    for j in range(ns):
        pick = (lambda a: a) if turb == "sa" else (lambda a: a[j])
        X, Xn, Xm = q[5 + j], qn[5 + j], qm1[5 + j]
        tt = g.vol / dt * (1.5 * rstar * X - 2 * rn * Xn + 0.5 * rm * Xm)
        B = -(tt - pick(ss["src"]))
        sc = np.abs(tt) + np.abs(pick(ss["src"]))
        assert np.array_equal(pick(su["src"]), pick(ss["src"]))
        assert np.all(np.abs(pick(su["B"]) - B) <= 1e-12 * sc.max())
        D = pick(ss["D"]) + 0.5 * rhodot + g.vol / dtau * (1.5 * rstar - rho)
        sd = np.abs(pick(ss["D"])) + np.abs(rhodot) + np.abs(g.vol / dtau * rho)
        assert np.all(np.abs(pick(su["D"]) - D) <= 1e-12 * sd.max())
        assert np.array_equal(pick(su["fjp"]), pick(ss["fjp"]))


def _back(o0, g, q0, dt):
    """q at t = -dt from the Taylor series of the semi-discrete ODE
    dQ/dt = src / V (error O(dt^3))."""
    Rf = lambda Q: o0.residual(prim(Q))[0] / g.vol
    Q0 = cons(q0)
    R0 = Rf(Q0)
    eps = 1e-6 * np.abs(Q0).max() / np.abs(R0).max()
    dR = (Rf(Q0 + eps * R0) - Rf(Q0 - eps * R0)) / (2 * eps)
    return prim(Q0 - dt * R0 + 0.5 * dt * dt * dR)


def test_second_order_in_time():
    g, q, prm = workloads.box(6)
    prm = dict(prm, limiter="none")
    o0 = oracle.Oracle(g, prm)
    ends, ends_c = {}, {}
    for k in (4, 8, 16):
        dt = 1e-4 / k
        o = oracle.Oracle(g, dict(prm, unsteady=True, dtmax=dt, newton_iter=14, sgs_iter=20, cflmax=100.0))
        ends[k], _, h = o.unsteady(q, _back(o0, g, q, dt), 4 * k)
        # the inner iterations converge (10+ orders), so dt alone sets the error
        assert np.sqrt(h[:, -1, 0]).max() < 1e-10 * np.sqrt(h[:, 0, 0]).max()
        if k < 16:
            ends_c[k], _, _ = o.unsteady(q, q, 4 * k)
    d1, d2 = ends[4] - ends[8], ends[8] - ends[16]
    for i in range(5):
        ratio = np.abs(d1[i]).max() / np.abs(d2[i]).max()
        assert 3.6 < ratio < 4.7, (i, ratio)
    # the constant start q(-dt) = q(0) converges at first order only
    e4, e8 = ends_c[4] - ends[16], ends_c[8] - ends[16]
    ratio = np.abs(e4[4]).max() / np.abs(e8[4]).max()
    assert 1.5 < ratio < 2.8, ratio
