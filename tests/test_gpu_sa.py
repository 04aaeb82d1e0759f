"""Parity of the Spalart-Allmaras path (turb = FPSI_TURB_SA, through the C
ABI) against the oracle: the fluid residual and system with the eddy
viscosity, the nuTilde system and its scalar SGS, and whole Newton steps of
the segregated fluid + nuTilde solve, under both linear solvers (scalar SGS
with block SGS, scalar LSGS with block LSGS).  Tolerances as for the laminar path:
1e-12 relative per face/cell quantity, 1e-10 on the state after N steps
(the only differences come from pow() in Sutherland's law and the SA
functions, which the device and the host library round differently in the
last place)."""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu


def _renumbered():
    g, q, prm = workloads.channel_sa(nx=5, ny=8, nz=4)
    perm = np.random.default_rng(12).permutation(g.nc)
    g2, q2 = grids.renumber_cells(g, perm, q)
    return g2, q2, prm


def _cases():
    return {
        "channel_sa": lambda: workloads.channel_sa(nx=6, ny=10, nz=4),
        "channel_sa_lowmach": lambda: workloads.channel_sa(nx=5, ny=8, nz=6, Minf=0.2),
        "channel_sa_renumbered": _renumbered,
    }


@pytest.fixture(scope="module", params=list(_cases()))
def case(request):
    g, q, prm = _cases()[request.param]()
    print(f"\n[sa case {request.param}] nc={g.nc} nf={g.nf}", flush=True)
    return request.param, g, q, prm


def rel(a, b):
    scale = np.abs(b).max() + 1e-300
    return float(np.abs(a - b).max() / scale)


def test_fluid_residual_with_eddy_viscosity(case):
    name, g, q, prm = case
    o = oracle.Oracle(g, prm)
    src_r, if_r, vf_r, _ = o.residual(q)
    s = Solver(g, prm)
    s.set_state(q)
    src, ifl, vfl, _ = s.residual()
    e = [rel(ifl, if_r), rel(vfl, vf_r), rel(src, src_r)]
    print(f"iflux {e[0]:.2e} vflux {e[1]:.2e} src {e[2]:.2e}")
    assert max(e) < 1e-11


def test_nutilde_system(case):
    name, g, q, prm = case
    o = oracle.Oracle(g, prm)
    ref = o.sa_assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    s.assemble()
    sy = s.sa_get_system()
    errs = {k: rel(sy[k], ref[k]) for k in ("src", "fjp", "fjm", "D", "B", "tmu")}
    print("nuTilde system errors", {k: f"{v:.1e}" for k, v in errs.items()})
    assert max(errs.values()) < 1e-11
    for nsw in (0, 3, 10):
        s.assemble()
        dX = s.sa_solve(nsw)
        dX_r = o.sa_sgs(ref, nsw)
        e = rel(dX, dX_r)
        print(f"scalar SGS sweeps={nsw} err {e:.2e}")
        assert e < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_newton_steps(case, prec):
    name, g, q, prm = case
    p = dict(prm, precision=prec)
    nsteps = 5
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
    print(f"{prec} SA state err after {nsteps} steps {err:.2e} (nuTilde row "
          f"{float((np.abs(q_g[5] - q_r[5]) / scale[5]).max()):.2e}); drop {np.sqrt(tot[-1] / tot[0]):.3e}")
    assert err < (1e-10 if prec == "fp64" else 1e-6)
    assert np.allclose(tot, tot_r, rtol=1e-8)
    assert np.all(q_g[5] >= 1e-15)


def test_nutilde_lsgs(case):
    """Scalar line SGS: the update norm is a tree sum here and a sequential
    sum in the oracle, so w and X agree to rounding, not bitwise."""
    name, g, q, prm = case
    prm = dict(prm, solver="lsgs")
    o = oracle.Oracle(g, prm)
    ref = o.sa_assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    for maxiter, reltol in ((0, 0.1), (1, 0.1), (3, 0.1), (10, 0.1), (25, 1e-6)):
        s.set_params(dict(prm, lsgs_maxiter=maxiter, lsgs_reltol=reltol))
        s.assemble()
        dX = s.sa_solve(maxiter)
        st = s.sa_lsgs_stats()
        dX_r, info = o.sa_lsgs(ref, maxiter=maxiter, reltol=reltol)
        e = rel(dX, dX_r)
        print(f"scalar lsgs maxiter={maxiter} reltol={reltol}: iters {st['iters']}/{info['iters']} "
              f"w {st['w']:.6g}/{info['w']:.6g} err {e:.2e}")
        assert st["iters"] == info["iters"]
        assert st["w"] == pytest.approx(info["w"], rel=1e-13)
        assert e < 1e-10


def test_newton_steps_lsgs(case):
    name, g, q, prm = case
    p = dict(prm, solver="lsgs")
    nsteps = 5
    o = oracle.Oracle(g, p)
    q_r, h_r = o.newton(q, nsteps)
    s = Solver(g, p)
    s.set_state(q)
    s.newton(nsteps)
    q_g = s.get_state()
    scale = np.abs(q_r).max(axis=1, keepdims=True)
    err = float((np.abs(q_g - q_r) / scale).max())
    print(f"lsgs SA state err after {nsteps} steps {err:.2e}")
    assert err < 1e-10
    assert np.all(q_g[5] >= 1e-15)


def test_zero_nutilde_is_laminar():
    g, q, prm = workloads.channel_sa(nx=5, ny=8, nz=4)
    q0 = q.copy()
    q0[5] = 0.0
    s = Solver(g, prm)
    s.set_state(q0)
    s.assemble()
    sy = s.get_system()
    sl = Solver(g, dict(prm, turb="none"))
    sl.set_state(q0[:5])
    sl.assemble()
    syl = sl.get_system()
    for k in ("fjp", "fjm", "D", "B", "dtau"):
        assert np.array_equal(sy[k], syl[k])


def test_scalar_rows_fused_in_dataflow_sweep(case):
    """The dataflow sweep carries the turbulence scalar rows in the fluid's
    launches (the block sweep with all fluid-scalar couplings zero): the
    state after Newton steps equals that of the separate fluid and scalar
    sweeps bit for bit, FP64 and FP32 blocks."""
    name, g, q, prm = case
    for prec in ("fp64", "fp32"):
        outs = []
        for mode in ("per_level", "persistent"):
            s = Solver(g, dict(prm, precision=prec, sweep=mode))
            s.set_state(q)
            s.newton(4)
            outs.append(s.get_state())
        print(f"SA fused vs separate sweeps, {prec}: max diff {np.abs(outs[0] - outs[1]).max():.1e}")
        assert np.array_equal(outs[0], outs[1])
