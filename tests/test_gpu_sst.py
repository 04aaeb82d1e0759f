"""Parity of the Menter SST path (turb = FPSI_TURB_SST, through the C ABI)
against the oracle: the fluid residual with the SST eddy viscosity, the k
and omega systems and their scalar SGS / LSGS solves, and whole Newton
steps of the segregated fluid + k + omega solve under both linear solvers.
Tolerances: 1e-11 relative per face/cell quantity, 1e-10 on the state after
N steps (pow, tanh and sqrt may round differently on the device and in the
host library)."""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu


def _renumbered():
    g, q, prm = workloads.channel_sst(nx=5, ny=8, nz=4)
    perm = np.random.default_rng(13).permutation(g.nc)
    g2, q2 = grids.renumber_cells(g, perm, q)
    return g2, q2, prm


def _cases():
    return {
        "channel_sst": lambda: workloads.channel_sst(nx=6, ny=10, nz=4),
        "channel_sst_lowmach": lambda: workloads.channel_sst(nx=5, ny=8, nz=6, Minf=0.2),
        "channel_sst_renumbered": _renumbered,
    }


@pytest.fixture(scope="module", params=list(_cases()))
def case(request):
    g, q, prm = _cases()[request.param]()
    print(f"\n[sst case {request.param}] nc={g.nc} nf={g.nf}", flush=True)
    return request.param, g, q, prm


def rel(a, b):
    scale = np.abs(b).max() + 1e-300
    return float(np.abs(a - b).max() / scale)


def test_fluid_residual_with_sst_viscosity(case):
    name, g, q, prm = case
    src_r, if_r, vf_r, _ = oracle.Oracle(g, prm).residual(q)
    s = Solver(g, prm)
    s.set_state(q)
    src, ifl, vfl, _ = s.residual()
    e = [rel(ifl, if_r), rel(vfl, vf_r), rel(src, src_r)]
    print(f"residual errors iflux {e[0]:.2e} vflux {e[1]:.2e} src {e[2]:.2e}")
    assert max(e) < 1e-11


@pytest.mark.parametrize("which", [0, 1])
def test_scalar_system_and_sgs(case, which):
    name, g, q, prm = case
    o = oracle.Oracle(g, prm)
    full = o.sst_assemble(q)
    ref = oracle.Oracle.scalar(full, which)
    s = Solver(g, prm)
    s.set_state(q)
    s.assemble()
    sy = s.turb_get_system(which)
    errs = {k: rel(sy[k], ref[k]) for k in ("src", "fjp", "fjm", "D", "B")}
    errs["tmu"] = rel(sy["tmu"], full["tmu"])
    print(f"scalar {which} system errors", {k: f"{v:.1e}" for k, v in errs.items()})
    assert max(errs.values()) < 1e-11
    for nsw in (0, 3, 10):
        s.assemble()
        dX = s.turb_solve(which, nsw)
        e = rel(dX, o.sa_sgs(ref, nsw))
        print(f"scalar {which} SGS sweeps={nsw} err {e:.2e}")
        assert e < 1e-10


@pytest.mark.parametrize("which", [0, 1])
def test_scalar_lsgs(case, which):
    name, g, q, prm = case
    prm = dict(prm, solver="lsgs")
    o = oracle.Oracle(g, prm)
    ref = oracle.Oracle.scalar(o.sst_assemble(q), which)
    s = Solver(g, prm)
    s.set_state(q)
    for maxiter, reltol in ((0, 0.1), (1, 0.1), (3, 0.1), (25, 1e-6)):
        s.set_params(dict(prm, lsgs_maxiter=maxiter, lsgs_reltol=reltol))
        s.assemble()
        dX = s.turb_solve(which, maxiter)
        st = s.turb_lsgs_stats(which)
        dX_r, info = o.sa_lsgs(ref, maxiter=maxiter, reltol=reltol)
        e = rel(dX, dX_r)
        print(f"scalar {which} lsgs maxiter={maxiter}: iters {st['iters']}/{info['iters']} "
              f"w {st['w']:.6g}/{info['w']:.6g} err {e:.2e}")
        assert st["iters"] == info["iters"]
        assert st["w"] == pytest.approx(info["w"], rel=1e-13)
        assert e < 1e-10


@pytest.mark.parametrize("solver", ["sgs", "lsgs"])
def test_newton_steps(case, solver):
    name, g, q, prm = case
    p = dict(prm, solver=solver)
    nsteps = 5
    q_r, h_r = oracle.Oracle(g, p).newton(q, nsteps)
    s = Solver(g, p)
    s.set_state(q)
    hist = s.newton(nsteps)
    q_g = s.get_state()
    scale = np.abs(q_r).max(axis=1, keepdims=True)
    err = (np.abs(q_g - q_r) / scale).max(axis=1)
    tot_r = h_r[:, 0] + h_r[:, 1] + h_r[:, 2]
    tot = np.array([h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist])
    print(f"{solver} SST state err after {nsteps} steps per row {np.array2string(err, precision=1)}")
    assert err.max() < 1e-10
    assert np.allclose(tot, tot_r, rtol=1e-8)
    assert np.all(q_g[5:] >= 1e-15)


def test_zero_k_is_laminar():
    g, q, prm = workloads.channel_sst(nx=5, ny=8, nz=4)
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
        print(f"SST fused vs separate sweeps, {prec}: max diff {np.abs(outs[0] - outs[1]).max():.1e}")
        assert np.array_equal(outs[0], outs[1])
