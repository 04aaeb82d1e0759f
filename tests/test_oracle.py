"""Pins of the CPU oracle by what the mathematics fixes (no GPU).

The flowPsi reference cannot be built here (it is Loci DSL and Loci is
absent), so every oracle function is checked against an independent
formulation: the textbook Euler flux and Roe matrix (numpy eigen-
decomposition), finite differences, closed forms, conservation, and a dense
block solve.  Each check is chosen so that a dropped term, a wrong sign, a
transposed block or a wrong neighbour fails it.
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads

G, R = 1.4, 287.0


def euler_flux_cons(U, n, gamma=G):
    """Textbook Euler flux of conserved U = (rho, rho u, rho E) through unit normal n."""
    rho = U[0]
    u = U[1:4] / rho
    E = U[4] / rho
    p = (gamma - 1) * rho * (E - 0.5 * u @ u)
    un = u @ n
    return np.array([rho * un, *(rho * u * un + p * n), (rho * E + p) * un])


def prim_to_cons(q, gamma=G, Rt=R):
    T, u, pg = q[0], np.asarray(q[1:4]), q[4]
    rho = pg / (Rt * T)
    E = Rt * T / (gamma - 1) + 0.5 * u @ u
    return np.array([rho, *(rho * u), rho * E])


def rand_state(rng, mach=0.3, jitter=0.05):
    T = 300 * (1 + jitter * rng.standard_normal())
    a = np.sqrt(G * R * T)
    u = mach * a * rng.standard_normal(3)
    p = 101325 * (1 + jitter * rng.standard_normal())
    return np.array([T, *u, p])


def unit(rng):
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("flux", [oracle.hllc_flux, oracle.roe_flux])
def test_consistency(flux):
    """F(q, q, n) equals area times the physical Euler flux."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        q = rand_state(rng, mach=rng.uniform(0.05, 2.0))
        n = unit(rng)
        area = rng.uniform(0.1, 3.0)
        f = flux(q, q, n, area=area)
        ref = area * euler_flux_cons(prim_to_cons(q), n)
        assert np.allclose(f, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("flux", [oracle.hllc_flux, oracle.roe_flux])
def test_antisymmetry(flux):
    """F(qL, qR, n) = -F(qR, qL, -n)."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        qL, qR, n = rand_state(rng), rand_state(rng), unit(rng)
        a = flux(qL, qR, n)
        b = flux(qR, qL, -n)
        assert np.allclose(a, -b, rtol=1e-11, atol=1e-10 * np.abs(a).max())


def test_hllc_supersonic_upwind():
    rng = np.random.default_rng(2)
    n = np.array([1.0, 0.0, 0.0])
    for _ in range(20):
        qL = rand_state(rng, jitter=0.01); qR = rand_state(rng, jitter=0.01)
        qL[1:4] = 0; qR[1:4] = 0
        qL[1] = qR[1] = 3.0 * np.sqrt(G * R * 300)
        assert np.allclose(oracle.hllc_flux(qL, qR, n), euler_flux_cons(prim_to_cons(qL), n), rtol=1e-12)
        qL[1] = qR[1] = -3.0 * np.sqrt(G * R * 300)
        assert np.allclose(oracle.hllc_flux(qL, qR, n), euler_flux_cons(prim_to_cons(qR), n), rtol=1e-12)


def roe_average_cons(qL, qR):
    UL, UR = prim_to_cons(qL), prim_to_cons(qR)
    rl, rr = UL[0], UR[0]
    uL, uR = UL[1:4] / rl, UR[1:4] / rr
    HL = (UL[4] + qL[4]) / rl
    HR = (UR[4] + qR[4]) / rr
    sl, sr = np.sqrt(rl), np.sqrt(rr)
    u = (sl * uL + sr * uR) / (sl + sr)
    H = (sl * HL + sr * HR) / (sl + sr)
    E = (H + (G - 1) * 0.5 * u @ u) / G
    return np.array([1.0, *u, E])  # any density: A depends on u, H only


def jac_fd(fun, x, h=1e-7):
    x = np.asarray(x, float)
    J = np.zeros((len(fun(x)), len(x)))
    for j in range(len(x)):
        d = h * max(abs(x[j]), 1.0)
        xp = x.copy(); xp[j] += d
        xm = x.copy(); xm[j] -= d
        J[:, j] = (fun(xp) - fun(xm)) / (2 * d)
    return J


def test_roe_flux_matches_textbook_roe():
    """At eta = 1, roeCT_flux = 1/2(F_L + F_R) - 1/2 |A_roe| (U_R - U_L) with
    |A_roe| = R |Lambda| R^-1 of the conserved Euler Jacobian at the Roe state."""
    rng = np.random.default_rng(3)
    for _ in range(30):
        qL = rand_state(rng, mach=0.4, jitter=0.02)
        qR = qL + np.array([2.0, 3.0, -2.0, 1.0, 800.0]) * rng.standard_normal(5)
        n = unit(rng)
        Ur = roe_average_cons(qL, qR)
        A = jac_fd(lambda U: euler_flux_cons(U, n), Ur)
        lam, Rv = np.linalg.eig(A)
        absA = (Rv @ np.diag(np.abs(lam)) @ np.linalg.inv(Rv)).real
        UL, UR = prim_to_cons(qL), prim_to_cons(qR)
        ref = 0.5 * (euler_flux_cons(UL, n) + euler_flux_cons(UR, n)) - 0.5 * absA @ (UR - UL)
        f = oracle.roe_flux(qL, qR, n)
        assert np.allclose(f, ref, rtol=1e-7, atol=1e-7 * np.abs(ref).max())


def test_hllc_matches_toro_hllc():
    """At eta = 1 the HLLC flux equals Toro's HLLC with the same wave-speed
    estimates (Einfeldt/Roe bounds), in the subsonic star regions."""
    rng = np.random.default_rng(4)
    for _ in range(40):
        qL = rand_state(rng, mach=0.5, jitter=0.05)
        qR = rand_state(rng, mach=0.5, jitter=0.05)
        n = unit(rng)
        UL, UR = prim_to_cons(qL), prim_to_cons(qR)
        FL, FR = euler_flux_cons(UL, n), euler_flux_cons(UR, n)
        rl, rr = UL[0], UR[0]
        ul, ur = qL[1:4] @ n, qR[1:4] @ n
        pl, pr = qL[4], qR[4]
        al, ar = np.sqrt(G * R * qL[0]), np.sqrt(G * R * qR[0])
        Ua = roe_average_cons(qL, qR)
        uroe = Ua[1:4] @ n
        Hroe = G * Ua[4] - (G - 1) * 0.5 * Ua[1:4] @ Ua[1:4]
        aroe = np.sqrt((G - 1) * (Hroe - 0.5 * Ua[1:4] @ Ua[1:4]))
        SL = min(uroe - aroe, ul - al)
        SR = max(uroe + aroe, ur + ar)
        SM = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))

        def ustar(U, q, S, un):
            rho = U[0]
            fac = rho * (S - un) / (S - SM)
            u = q[1:4]
            E = U[4] / rho
            us = u + (SM - un) * n
            Es = E + (SM - un) * (SM + q[4] / (rho * (S - un)))
            return fac * np.array([1.0, *us, Es])

        if SL >= 0:
            ref = FL
        elif SR <= 0:
            ref = FR
        elif SM >= 0:
            ref = FL + SL * (ustar(UL, qL, SL, ul) - UL)
        else:
            ref = FR + SR * (ustar(UR, qR, SR, ur) - UR)
        f = oracle.hllc_flux(qL, qR, n)
        assert np.allclose(f, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def dQdq(q):
    return jac_fd(prim_to_cons, q, h=1e-7)


def test_roe_jacobian_sum_is_flux_jacobian():
    """RoeJacobian is linear in the eigenvalues and fjp + fjm uses eigenvalues
    summing to the unsmoothed ones; at qL = qR and eta = 1 it is d(area F.n)/dq."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rand_state(rng, mach=rng.uniform(0.1, 1.5))
        n = unit(rng)
        area = rng.uniform(0.5, 2.0)
        J = oracle.roe_jacobian(q, q, n, 1, area=area) + oracle.roe_jacobian(q, q, n, -1, area=area)
        ref = jac_fd(lambda x: oracle.inviscid_flux(x, n, area=area), q, h=1e-7)
        assert np.allclose(J, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def test_roe_jacobian_difference_is_smoothed_abs_a():
    """fjp - fjm = R |Lambda|_eps R^-1 dQ/dq with |lambda|_eps = sqrt(lambda^2 + eps^2),
    eps = 0.02 (lambda_max - lambda_min)."""
    rng = np.random.default_rng(6)
    for _ in range(20):
        q = rand_state(rng, mach=rng.uniform(0.1, 1.5))
        n = unit(rng)
        J = oracle.roe_jacobian(q, q, n, 1) - oracle.roe_jacobian(q, q, n, -1)
        U = prim_to_cons(q)
        A = jac_fd(lambda x: euler_flux_cons(x, n), U)
        lam, Rv = np.linalg.eig(A)
        lam = lam.real
        eps = 0.02 * (lam.max() - lam.min())
        absA = (Rv @ np.diag(np.sqrt(lam ** 2 + eps ** 2)) @ np.linalg.inv(Rv)).real
        ref = absA @ dQdq(q)
        assert np.allclose(J, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def test_characteristic_farfield_states():
    rng = np.random.default_rng(7)
    a0 = np.sqrt(G * R * 300.0)
    q0 = np.array([300.0, 0.3 * a0, 0.0, 0.0, 101325.0])
    for kind in ("inflow", "outflow"):
        n = unit(rng)
        # interior equal to the reference state: the face takes it exactly
        assert np.allclose(oracle.cv_state(kind, q0, q0, n), q0, rtol=1e-14)
        # beta = 1: textbook linearised characteristic pressure and velocity
        qi = q0 + np.array([1.0, 2.0, -1.0, 0.5, 300.0])
        qb = oracle.cv_state(kind, qi, q0, n)
        r0 = q0[4] / (R * q0[0])
        uin, u0n = qi[1:4] @ n, q0[1:4] @ n
        pb = 0.5 * (qi[4] + q0[4] + r0 * a0 * (uin - u0n))
        assert np.isclose(qb[4], pb, rtol=1e-13)
        base = q0[1:4] if kind == "inflow" else qi[1:4]
        pbase = q0[4] if kind == "inflow" else qi[4]
        ub = base + n * (pb - pbase) / (r0 * a0) * (1 if kind == "inflow" else -1)
        assert np.allclose(qb[1:4], ub, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("fj", ["f64", "f32"])
def test_block_lu(fj):
    """Packed pivoted LU: P A = L U, same pivots as LAPACK getrf, solve = numpy."""
    import scipy.linalg as sl
    rng = np.random.default_rng(8)
    for _ in range(30):
        A = rng.standard_normal((5, 5)) + 0.5 * np.eye(5)
        if fj == "f32":
            A = A.astype(np.float32).astype(np.float64)
        a, piv, rc = oracle.lu_decompose(A, fj)
        assert rc == 0
        LU = a.reshape(5, 5).T.astype(np.float64)
        Lm = np.tril(LU, -1) + np.eye(5)
        Um = np.triu(LU)
        tol = 1e-12 if fj == "f64" else 1e-5
        assert np.allclose(A[piv], Lm @ Um, atol=tol * np.abs(A).max())
        lu, ip = sl.lu_factor(A)
        perm = np.arange(5)
        for i, p in enumerate(ip):
            perm[[i, p]] = perm[[p, i]]
        assert np.array_equal(perm, piv)
        b = rng.standard_normal(5)
        x = oracle.lu_solve(a, piv, b, fj)
        assert np.allclose(A @ x, b, atol=(1e-12 if fj == "f64" else 1e-5) * np.abs(b).max() * 10)


def dense_system(g, sys_):
    """Dense block matrix from the assembled blocks: rows = equations of a cell,
    D on the diagonal, U = fjm at (cl, cr), L = -fjp at (cr, cl), periodic
    U = fjm at (cl, partner cell)."""
    nc = g.nc
    A = np.zeros((5 * nc, 5 * nc))
    for c in range(nc):
        a = sys_["D"][c].astype(np.float64)
        LU = a.reshape(5, 5).T
        Lm = np.tril(LU, -1) + np.eye(5)
        Um = np.triu(LU)
        PD = Lm @ Um
        D = np.zeros((5, 5))
        D[sys_["piv"][c]] = PD
        A[5 * c:5 * c + 5, 5 * c:5 * c + 5] = D
    for f in range(g.nf):
        fjp = sys_["fjp"][f].astype(np.float64).reshape(5, 5).T
        fjm = sys_["fjm"][f].astype(np.float64).reshape(5, 5).T
        if g.btype[f] == grids.INTERIOR:
            l, r = g.cl[f], g.cr[f]
            A[5 * l:5 * l + 5, 5 * r:5 * r + 5] += fjm
            A[5 * r:5 * r + 5, 5 * l:5 * l + 5] += -fjp
        elif g.btype[f] == grids.PERIODIC:
            l, r = g.cl[f], g.cl[g.pface[f]]
            A[5 * l:5 * l + 5, 5 * r:5 * r + 5] += fjm
    b = sys_["B"].astype(np.float64).ravel()
    return A, b


@pytest.mark.parametrize("mk", ["box", "channel"])
def test_sgs_converges_to_dense_solve(mk):
    """Many symmetric sweeps reach the solution of the dense block system."""
    if mk == "box":
        g, q, prm = workloads.box(4, cflmax=2.0)
    else:
        g, q, prm = workloads.channel(nx=4, ny=6, nz=4, cflmax=2.0)
    o = oracle.Oracle(g, prm)
    s = o.assemble(q)
    A, b = dense_system(g, s)
    x = np.linalg.solve(A, b)
    X = o.sgs(s, 300).ravel()
    assert np.allclose(X, x, rtol=1e-9, atol=1e-9 * np.abs(x).max())
    # Jacobi start is D^-1 B
    X0 = o.sgs(s, 0).ravel()
    for c in range(g.nc):
        Dc = A[5 * c:5 * c + 5, 5 * c:5 * c + 5]
        assert np.allclose(Dc @ X0[5 * c:5 * c + 5], b[5 * c:5 * c + 5], rtol=1e-10,
                           atol=1e-10 * np.abs(b[5 * c:5 * c + 5]).max())


def test_time_term_is_dQdq():
    """D(dtau1) - D(dtau2) = V (1/dtau1 - 1/dtau2) dQ/dq at eta = 1."""
    g, q, prm = workloads.box(4)
    o1 = oracle.Oracle(g, dict(prm, cflmax=5.0))
    o2 = oracle.Oracle(g, dict(prm, cflmax=50.0))
    s1, s2 = o1.assemble(q), o2.assemble(q)
    A1, _ = dense_system(g, s1)
    A2, _ = dense_system(g, s2)
    for c in (0, 7, g.nc - 1):
        dD = A1[5 * c:5 * c + 5, 5 * c:5 * c + 5] - A2[5 * c:5 * c + 5, 5 * c:5 * c + 5]
        M = dQdq(q[:, c])
        coef = g.vol[c] * (1 / s1["dtau"][c] - 1 / s2["dtau"][c])
        assert np.allclose(dD, coef * M, rtol=1e-6, atol=1e-9 * np.abs(coef * M).max())


def test_conservation_and_freestream():
    g, q, prm = workloads.box(6, flux="hllc")
    o = oracle.Oracle(g, prm)
    src, ifl, vfl, _ = o.residual(q)
    bnd = g.btype != grids.INTERIOR
    # sum over cells of src = - sum of boundary fluxes (interior fluxes telescope)
    assert np.allclose(src.sum(1), -ifl[:, bnd].sum(1), rtol=1e-10,
                       atol=1e-10 * np.abs(ifl).sum(1).max())
    qu = np.zeros_like(q)
    qu[0] = 300.0; qu[1] = prm["u0"][0]; qu[4] = 101325.0
    src_u, ifl_u, _, _ = o.residual(qu)
    assert np.abs(src_u).max() <= 1e-12 * np.abs(ifl_u).max()


def test_lsq_gradient_exact_for_linear_fields():
    g, q, prm = workloads.channel(nx=6, ny=8, nz=6)
    o = oracle.Oracle(g, dict(prm, limiter="none"))
    coef = np.array([[300.0, 5e3, -3e3, 2e3], [10.0, 1e4, 2e3, -1e3], [1.0, 3e3, 1e3, 2e3],
                     [-2.0, 1e3, -2e3, 3e3], [101325.0, 1e6, 2e6, -1e6]])
    x = g.cc
    ql = coef[:, :1] + coef[:, 1:] @ x
    grad, phi = o.gradients(ql)
    # cells whose stencil holds no boundary-face state (periodic ghosts are cells)
    touch = np.zeros(g.nc, bool)
    b = (g.btype != grids.INTERIOR) & (g.btype != grids.PERIODIC)
    touch[g.cl[b]] = True
    per = g.btype == grids.PERIODIC
    touch[g.cl[per]] = True  # a linear field is not periodic
    inner = ~touch
    assert inner.sum() > 0
    for k in range(5):
        for d in range(3):
            assert np.allclose(grad[3 * k + d, inner], coef[k, 1 + d], rtol=1e-9)
    assert np.all(phi == 1.0)


@pytest.mark.parametrize("lim", ["barth", "venkatakrishnan", "zero"])
def test_limiter_bounds(lim):
    g, q, prm = workloads.box(6)
    o = oracle.Oracle(g, dict(prm, limiter=lim))
    qf = o.face_states(q)
    grad, phi = o.gradients(q, qf)
    assert np.all(phi >= 0) and np.all(phi <= 1)
    if lim == "zero":
        assert np.all(phi == 0)
        return
    if lim != "barth":
        return
    # Barth: every limited face extrapolation stays within the neighbour range
    nb = {c: [] for c in range(g.nc)}
    for f in range(g.nf):
        l, r = g.cl[f], g.cr[f]
        if r >= 0:
            nb[l].append(q[:, r]); nb[r].append(q[:, l])
        else:
            nb[l].append(qf[:, f])
    for f in range(g.nf):
        for c in ([g.cl[f], g.cr[f]] if g.cr[f] >= 0 else [g.cl[f]]):
            vals = np.array(nb[c] + [q[:, c]])
            lo, hi = vals.min(0), vals.max(0)
            d = g.fc[:, f] - g.cc[:, c]
            ext = q[:, c] + phi[:, c] * (grad[:, c].reshape(5, 3) @ d)
            span = hi - lo + 1e-300
            assert np.all(ext <= hi + 1e-9 * span) and np.all(ext >= lo - 1e-9 * span)


def test_viscous_flux_linear_couette():
    """Linear u(y) and T(y): tau_xy = mu du/dy and q_y = k dT/dy on interior faces."""
    g, q, prm = workloads.channel(nx=4, ny=8, nz=4)
    o = oracle.Oracle(g, dict(prm, limiter="none"))
    y = g.cc[1]
    dudy, dTdy = 3e5, 2e4
    ql = np.zeros_like(q)
    ql[0] = 300.0 + dTdy * y
    ql[1] = 50.0 + dudy * y
    ql[4] = 101325.0
    out = o.fluxes(ql)
    vf = out["vflux"]
    ny_faces = (g.btype == grids.INTERIOR) & (np.abs(g.n[1]) > 0.99)
    f = np.nonzero(ny_faces)[0]
    l, r = g.cl[f], g.cr[f]
    suth = lambda T: 1.458e-6 * T ** 1.5 / (T + 110.4)
    kth = lambda T: 2.495e-3 * T ** 1.5 / (T + 194.0)
    vl, vr = g.vol[l], g.vol[r]
    mu_f = (vl * suth(ql[0, r]) + vr * suth(ql[0, l])) / (vl + vr)
    k_f = (vl * kth(ql[0, r]) + vr * kth(ql[0, l])) / (vl + vr)
    A = g.area[f]
    assert np.allclose(vf[0, f], mu_f * dudy * A, rtol=1e-9)
    assert np.allclose(vf[1, f], 0.0, atol=1e-9 * np.abs(vf[0, f]).max())
    uf = (vl * ql[1, r] + vr * ql[1, l]) / (vl + vr)
    assert np.allclose(vf[3, f], uf * vf[0, f] + k_f * dTdy * A, rtol=1e-9)
    assert np.isclose(suth(300.0), 1.8460e-5, rtol=1e-4)


def test_newton_damping_clip():
    """alpha = min(1, 0.9 T/|dT|, 0.9 p/|dp|): drive a step that would make p < 0."""
    g, q, prm = workloads.box(4, cflmax=1e6)
    o = oracle.Oracle(g, prm)
    qn, _ = o.newton(q, 1)
    assert np.all(qn[0] > 0) and np.all(qn[4] > 0)


def test_f32_build_is_float():
    assert oracle.lib("f32").og_real_fj_bytes() == 4
    assert oracle.lib("f64").og_real_fj_bytes() == 8
