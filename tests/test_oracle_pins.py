"""Pins of oracle functions that no closed form elsewhere fixes (no GPU).

Each test is chosen so that a plausible slip in the oracle fails it:

* Venkatakrishnan limiter (Loci FVM, Venkatakrishnan AIAA 93-0880, the
  option described at guide/guide.tex:495-516): with the threshold off
  (Kl = 0) every cell's phi is min(1, min over its faces of
  (y^2 + 2y)/(y^2 + y + 2)), y = Delta+/Delta-, computed here from the
  unlimited gradients and the neighbour extremes (a swapped Delta+/Delta-
  or a missing stencil member fails); the threshold eps^2 = (Kl h)^3 with
  h^3 = vol is pinned by a scaling law: phi(a dq, Kl) = phi(dq, Kl a^(-2/3))
  (an exponent other than 3 fails); Kl -> inf gives phi -> 1.
* Newton damping (src/solverNewton.loci:62-78, BDF2 :160-180, KEC
  :93-125): the update is q + alpha dq with one alpha per cell; where
  alpha < 1 the binding limit lands exactly on its bound: T or p moves by
  90 % (steady) or 75 % (BDF2) of its value, the KEC reconstruction of the
  new level keeps exactly 20 % of its absolute pressure.
* KEC localPBias (src/solverTimestep.loci:346-351): its closed form in the
  cell CFL number cflpdt * dtmax, read back from the oracle's own dtau.
* Thin-layer viscous Jacobian (src/fluxViscous.loci:315-403): at a uniform
  state on a uniform hex grid the viscous parts of fjp / fjm equal the
  central finite differences of -vflux with respect to the two cells'
  states (the 1/3 normal-stress factor, mu A/dist and k A/dist, the u_f
  weights).  The 0.5 vflux term of the energy row vanishes there; it is
  restated with the reference's sign, which is the opposite of the exact
  derivative's (the Jacobian is an approximation; the solution does not
  depend on it).
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads


def _stencil_extremes(g, q, qf):
    """per cell and variable: min and max over the cell and its stencil
    (interior neighbour, periodic partner, boundary face state)"""
    lo, hi = q.copy(), q.copy()
    for f in range(g.nf):
        l, r, t = g.cl[f], g.cr[f], g.btype[f]
        if t == grids.INTERIOR:
            for a, b in ((l, r), (r, l)):
                lo[:, a] = np.minimum(lo[:, a], q[:, b]); hi[:, a] = np.maximum(hi[:, a], q[:, b])
        elif t == grids.PERIODIC:
            o = g.cl[g.pface[f]]
            lo[:, l] = np.minimum(lo[:, l], q[:, o]); hi[:, l] = np.maximum(hi[:, l], q[:, o])
        else:
            lo[:, l] = np.minimum(lo[:, l], qf[:, f]); hi[:, l] = np.maximum(hi[:, l], qf[:, f])
    return lo, hi


def _face_lists(g):
    out = [[] for _ in range(g.nc)]
    for f in range(g.nf):
        out[g.cl[f]].append(f)
        if g.btype[f] == grids.INTERIOR:
            out[g.cr[f]].append(f)
    return out


def test_venkatakrishnan_unthresholded_closed_form():
    g, q, prm = workloads.box(5, seed=11)
    rng = np.random.default_rng(4)
    q = q * (1.0 + 0.02 * rng.standard_normal(q.shape))
    o0 = oracle.Oracle(g, dict(prm, limiter="none"))
    qf = o0.face_states(q)
    grad, phi_none = o0.gradients(q, qf)
    assert np.all(phi_none == 1.0)
    _, phi = oracle.Oracle(g, dict(prm, limiter="venkatakrishnan", Kl=0.0)).gradients(q, qf)
    lo, hi = _stencil_extremes(g, q, qf)
    lists = _face_lists(g)
    ref = np.ones_like(phi)
    for c in range(g.nc):
        for f in lists[c]:
            dm = grad[:, c].reshape(5, 3) @ (g.fc[:, f] - g.cc[:, c])
            dp = np.where(dm >= 0, hi[:, c] - q[:, c], lo[:, c] - q[:, c])
            with np.errstate(divide="ignore", invalid="ignore"):
                y = dp / dm
                v = np.where(dm == 0.0, 1.0, (y * y + 2 * y) / (y * y + y + 2))
            ref[:, c] = np.minimum(ref[:, c], v)
    ref = np.clip(ref, 0.0, 1.0)
    assert np.abs(phi - ref).max() <= 1e-12
    # the limiter acts: some phi well below 1, and a swapped Delta+/Delta- would differ
    assert (phi < 0.9).sum() > 10


def _symmetry_box(n):
    x = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"))
    return grids.structured_hex(X, {s: grids.SYMMETRY for s in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")},
                                name=f"sym{n}")


def test_venkatakrishnan_threshold_scaling():
    """phi depends on Delta+/Delta- and eps^2/Delta-^2 with eps^2 = (Kl h)^3:
    velocities scaled by a = 8 with Kl = 1 limit exactly like the unscaled
    velocities with Kl = 8^(-2/3) = 1/4 (symmetry walls: the mirrored face
    states scale with the velocity)."""
    g = _symmetry_box(5)
    rng = np.random.default_rng(7)
    q = np.zeros((5, g.nc))
    q[0] = 300.0
    q[4] = 1e5
    q[1:4] = rng.standard_normal((3, g.nc))
    prm = dict(workloads.BASE, flux="roe", viscous=0, u0=(0.0, 0.0, 0.0), cflmax=10.0, sgs_iter=5,
               limiter="venkatakrishnan")
    qa = q.copy()
    qa[1:4] *= 8.0
    oa = oracle.Oracle(g, dict(prm, Kl=1.0))
    o1 = oracle.Oracle(g, dict(prm, Kl=0.25))
    _, pa = oa.gradients(qa, oa.face_states(qa))
    _, p1 = o1.gradients(q, o1.face_states(q))
    assert np.abs(pa[1:4] - p1[1:4]).max() <= 1e-13
    assert (p1[1:4] < 0.95).sum() > 10
    # the threshold matters: the same scaled field with Kl = 1 differs
    _, pk = o1.gradients(qa, o1.face_states(qa))
    assert np.abs(pk[1:4] - pa[1:4]).max() > 1e-3
    # Kl -> infinity switches the limiter off
    ob = oracle.Oracle(g, dict(prm, Kl=1e8))
    _, pb = ob.gradients(q, ob.face_states(q))
    assert pb[1:4].min() > 1.0 - 1e-9


def _alpha_cells(q0, q1, dq):
    """per cell: the single alpha with q1 - q0 = alpha dq (all five
    components), checked collinear"""
    d = q1 - q0
    k = np.argmax(np.abs(dq), axis=0)
    a = d[k, np.arange(q0.shape[1])] / dq[k, np.arange(q0.shape[1])]
    assert np.allclose(d, a * dq, rtol=0, atol=1e-12 * (np.abs(q0) + np.abs(dq)).max(axis=1, keepdims=True))
    return a


def test_newton_alpha_clip_lands_on_ninety_percent():
    g, q, prm = workloads.box(4, seed=3, cflmax=1e8)
    rng = np.random.default_rng(9)
    q = q.copy()
    q[4] *= 1.0 + 0.3 * rng.standard_normal(g.nc).clip(-2, 2)
    o = oracle.Oracle(g, prm)
    q1, _ = o.newton(q, 1)
    dq = o.sgs(o.assemble(q), prm["sgs_iter"]).T
    a = _alpha_cells(q, q1, dq)
    assert np.all(a <= 1.0 + 1e-12) and np.all(a > 0)
    clip = a < 1.0 - 1e-12  # (q + dq) - q reproduces dq only to rounding
    assert clip.sum() >= 3
    # where clipped, T or p changed by exactly 90 % of its value
    rT = np.abs(q1[0] - q[0]) / q[0]
    rp = np.abs(q1[4] - q[4]) / q[4]  # Pambient = 0
    land = np.minimum(np.abs(rT - 0.9), np.abs(rp - 0.9))
    assert land[clip].max() <= 1e-13
    assert np.all(np.maximum(rT, rp)[~clip] < 0.9)


def test_bdf2_alpha_clip_lands_on_seventy_five_percent():
    g, q, prm = workloads.box(4, seed=3)
    rng = np.random.default_rng(12)
    q = q.copy()
    q[4] *= 1.0 + 0.3 * rng.standard_normal(g.nc).clip(-2, 2)
    p = dict(prm, unsteady=True, time_scheme="bdf2", dtmax=1.0, newton_iter=1, cflmax=1e8)
    o = oracle.Oracle(g, p)
    q1, _, _ = o.unsteady(q, q, 1)
    dq = o.sgs(o.assemble(q, q, q), prm["sgs_iter"]).T
    a = _alpha_cells(q, q1, dq)
    clip = a < 1.0 - 1e-12
    assert clip.sum() >= 3 and np.all(a <= 1.0 + 1e-12)
    rT = np.abs(q1[0] - q[0]) / q[0]
    rp = np.abs(q1[4] - q[4]) / q[4]
    land = np.minimum(np.abs(rT - 0.75), np.abs(rp - 0.75))
    assert land[clip].max() <= 1e-13
    assert np.all(np.maximum(rT, rp)[~clip] < 0.75)


def test_kec_local_pressure_bias_closed_form():
    """localPBias = max(control, min(1, max(1e-40, 0.1 (cfl - 1)))^0.7) with
    cfl = cflpdt dtmax; with a small cflmax the CFL limit sets dtau, so
    cflpdt = cflmax / dtau."""
    g, q, prm = workloads.box(5)
    cflmax = 0.01
    seen = set()
    for dt, ctl in ((4e-5, 0.0), (3e-4, 0.0), (2e-3, 0.0), (3e-4, 0.9)):
        o = oracle.Oracle(g, dict(prm, unsteady=True, time_scheme="kec", dtmax=dt, cflmax=cflmax, urelax=1e30,
                                  pbias_control=ctl))
        a = o.assemble(q, q)
        assert np.all(a["dtau"] < dt)  # the CFL limit binds everywhere
        cfl = cflmax / a["dtau"] * dt
        fac = np.minimum(1.0, np.maximum(1e-40, 0.1 * (cfl - 1.0)))
        ref = np.maximum(ctl, fac ** 0.7)
        assert np.allclose(a["pbias"], ref, rtol=1e-12, atol=0)
        seen.update(("low" if c <= 1 else "mid" if c < 11 else "sat") for c in cfl)
        # special values: (1e-40)^0.7 = 1e-28 at cfl <= 1, exactly 1 at cfl >= 11
        assert np.all(a["pbias"][cfl >= 11.0] == max(ctl, 1.0))
        assert np.allclose(a["pbias"][cfl <= 1.0], max(ctl, 1e-28), rtol=1e-13, atol=0)
    assert seen == {"low", "mid", "sat"}


def test_kec_alpha_limits_land_on_their_bounds():
    """One KEC iteration from q* = q^n with localPBias = 1 (control 1): the
    limits allow T* to move by 40 % of T (0.8 (2 T^n - T*) / 2) and p* by
    80 % of p ((1 + eps) 0.8 p^(n+1) / 2 with p^(n+1) = p* at eps = 1);
    where alpha < 1 one of them is met exactly.  The level-(n+1) state that
    comes back is the KEC reconstruction: p^(n+1) = p*, T^(n+1) the harmonic
    form T^n T* / (2 T^n - T*)."""
    g, q, prm = workloads.box(4, seed=3)
    rng = np.random.default_rng(13)
    q = q.copy()
    q[4] *= 1.0 + 0.3 * rng.standard_normal(g.nc).clip(-2, 2)
    p = dict(prm, unsteady=True, time_scheme="kec", dtmax=1.0, newton_iter=1, cflmax=1e8, pbias_control=1.0)
    o = oracle.Oracle(g, p)
    sys_ = o.assemble(q, q)
    assert np.all(sys_["pbias"] == 1.0)
    dq = o.sgs(sys_, prm["sgs_iter"]).T
    q1, _, _ = o.unsteady(q, q, 1)
    alpha = (q1[4] - q[4]) / dq[4]
    Ts = q[0] + alpha * dq[0]
    assert np.allclose(q1[0], q[0] * Ts / (2 * q[0] - Ts), rtol=1e-12)
    assert np.all(alpha <= 1.0 + 1e-9) and np.all(alpha > 0)
    clip = alpha < 1.0 - 1e-9
    assert clip.sum() >= 2
    mT = np.abs(Ts - q[0]) / q[0]
    mp = np.abs(q1[4] - q[4]) / q[4]
    land = np.minimum(np.abs(mT - 0.4), np.abs(mp - 0.8))
    assert land[clip].max() <= 1e-12
    assert np.all((mT < 0.4 + 1e-12) & (mp < 0.8 + 1e-12))


def test_thin_layer_viscous_jacobian_is_the_normal_derivative():
    g, q, prm = workloads.box(6)
    q = np.tile(q[:, :1], (1, g.nc))  # uniform state
    q[1:4, :] = np.array([[40.0], [-25.0], [15.0]])
    pv = dict(prm, viscous=1, flux="hllc", limiter="none", Twall=300.0)
    sv = oracle.Oracle(g, pv).assemble(q)
    si = oracle.Oracle(g, dict(pv, viscous=0)).assemble(q)
    col = lambda b: np.asarray(b, float).reshape(5, 5).T  # column-major -> [row, col]
    # faces whose two cells sit away from the boundary (symmetric LSQ stencils)
    touch = np.zeros(g.nc, bool)
    touch[g.cl[g.btype != grids.INTERIOR]] = True
    faces = [f for f in range(g.nf) if g.btype[f] == grids.INTERIOR and not touch[g.cl[f]] and not touch[g.cr[f]]]
    assert len(faces) > 20
    ov = oracle.Oracle(g, pv)
    h = dict(zip(range(5), (1e-3, 1e-3, 1e-3, 1e-3, 1.0)))
    checked = 0
    for f in faces[::7]:
        Jp = col(sv["fjp"][f]) - col(si["fjp"][f])
        Jm = col(sv["fjm"][f]) - col(si["fjm"][f])
        for side, J in ((g.cl[f], Jp), (g.cr[f], Jm)):
            fd = np.zeros((5, 5))
            for k in range(5):
                qp, qm = q.copy(), q.copy()
                qp[k, side] += h[k]
                qm[k, side] -= h[k]
                vp, vm = ov.fluxes(qp)["vflux"][:, f], ov.fluxes(qm)["vflux"][:, f]
                fd[1:, k] = -(vp - vm) / (2 * h[k])  # the face flux carries -vflux
            scale = np.abs(J).max()
            assert scale > 0
            assert np.abs(J - fd).max() <= 1e-7 * scale, (f, side, J, fd)
            checked += 1
    assert checked >= 6


def _dQdq(T, u, p, R=287.0, G=1.4):
    """textbook Jacobian of (rho, rho u, rho e0) with respect to (T, u, v, w, p)"""
    rho = p / (R * T)
    k = 0.5 * float(u @ u)
    J = np.zeros((5, 5))
    J[0, 0], J[0, 4] = -rho / T, 1.0 / (R * T)
    for i in range(3):
        J[1 + i, 0] = -rho * u[i] / T
        J[1 + i, 1 + i] = rho
        J[1 + i, 4] = u[i] / (R * T)
        J[4, 1 + i] = rho * u[i]
    J[4, 0] = -rho * k / T
    J[4, 4] = 1.0 / (G - 1) + k / (R * T)
    return J


def test_temporal_damping_jacobian_is_scaled_conservative_jacobian():
    """temporalDamping (src/fluxInviscid.loci:794-866,1224-1251): fjp gains
    +beta (|u_L.n| + a_L) A/2 dQ/dq(q_L), fjm gains -beta (|u_R.n| + a_R) A/2
    dQ/dq(q_R), on interior faces only (first-order reconstruction, so the
    face states are the cell states)."""
    g, q, prm = workloads.box(4, seed=5)
    base = dict(prm, limiter="zero")
    beta = 0.3
    s0 = oracle.Oracle(g, base).assemble(q)
    s1 = oracle.Oracle(g, dict(base, temporal_damping=beta)).assemble(q)
    col = lambda b: np.asarray(b, float).reshape(5, 5).T
    nint = 0
    for f in range(g.nf):
        dP = col(s1["fjp"][f]) - col(s0["fjp"][f])
        dM = col(s1["fjm"][f]) - col(s0["fjm"][f])
        if g.btype[f] != grids.INTERIOR:
            assert not dP.any() and not dM.any()
            continue
        n = g.n[:, f]
        for c, d, sgn in ((g.cl[f], dP, 1.0), (g.cr[f], dM, -1.0)):
            T, u, p = q[0, c], q[1:4, c], q[4, c]
            ref = sgn * beta * (abs(u @ n) + np.sqrt(1.4 * 287.0 * T)) * g.area[f] / 2 * _dQdq(T, u, p)
            assert np.abs(d - ref).max() <= 1e-12 * np.abs(ref).max()
        nint += 1
    assert nint > 50
    # the damped system stays the same on boundary blocks and the rhs
    assert np.array_equal(s0["B"], s1["B"]) and np.array_equal(s0["dtau"], s1["dtau"])
