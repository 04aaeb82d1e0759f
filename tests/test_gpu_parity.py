"""Parity of the HIP Newton step (through the C ABI) against the CPU oracle.

Tolerances: per-face fluxes and per-cell residuals to
1e-12 relative, the state after N Newton steps to 1e-10 relative, in FP64.
A per-cell residual is a sum of nearly cancelling face fluxes, so its error
is measured against the sum of the magnitudes of the fluxes that enter it.
With FP32 Jacobians (real_fj = float) the kernels evaluate the same
operations in the same order as the oracle's float build, so the assembled
blocks, the swept updates and the states are compared bit for bit.
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu

RTOL_FLUX = 1e-12
RTOL_STATE = 1e-10


def _cases():
    return {
        "box_roe": lambda: workloads.box(8),
        "box_hllc_barth": lambda: workloads.box(6, flux="hllc", limiter="barth"),
        "channel_visc": lambda: workloads.channel(nx=6, ny=10, nz=4),
        "channel_visc_nolim": lambda: workloads.channel(nx=5, ny=8, nz=6, limiter="none"),
        # low-Mach preconditioning active (Minf < 0.5 makes eta vary per cell)
        "box_roe_lowmach": lambda: workloads.box(8, Minf=0.1),
        "channel_visc_lowmach": lambda: workloads.channel(nx=6, ny=10, nz=4, Minf=0.2),
        # temporalDamping set: damping_fjp / damping_fjm in the face Jacobians
        "box_roe_damped": lambda: workloads.box(6, temporal_damping=0.1),
        "channel_visc_damped": lambda: workloads.channel(nx=5, ny=8, nz=4, temporal_damping=0.3),
    }


@pytest.fixture(scope="module", params=list(_cases()))
def case(request):
    g, q, prm = _cases()[request.param]()
    print(f"\n[case {request.param}] nc={g.nc} nf={g.nf}", flush=True)
    return request.param, g, q, prm


def face_rel_err(a, b):
    """max over faces/components of |a-b| / (|b| + 1e-6 max|b_k|)."""
    scale = np.abs(b).max(axis=1, keepdims=True) + 1e-300
    return float((np.abs(a - b) / (np.abs(b) + 1e-6 * scale)).max())


def cell_res_err(src_gpu, src_ref, g, iflux, vflux):
    """|r_gpu - r_ref| / sum_{f in c} |flux_f| per component."""
    nc = g.nc
    mag = np.zeros((5, nc))
    af = np.abs(iflux)
    av = np.abs(vflux)
    for k in range(5):
        np.add.at(mag[k], g.cl, af[k])
        inter = g.cr >= 0
        np.add.at(mag[k], g.cr[inter], af[k][inter])
        if k >= 1:
            np.add.at(mag[k], g.cl, av[k - 1])
            np.add.at(mag[k], g.cr[inter], av[k - 1][inter])
    return float((np.abs(src_gpu - src_ref) / (mag + 1e-300)).max())


def test_gradients(case):
    name, g, q, prm = case
    o = oracle.Oracle(g, prm)
    grad_r, phi_r = o.gradients(q)
    s = Solver(g, prm)
    s.set_state(q)
    grad, phi = s.gradients()
    scale = np.abs(grad_r).max(axis=1, keepdims=True) + 1e-300
    eg = float((np.abs(grad - grad_r) / (np.abs(grad_r) + 1e-9 * scale)).max())
    ep = float(np.abs(phi - phi_r).max())
    print(f"grad err {eg:.2e} phi err {ep:.2e}")
    assert eg < 1e-10 and ep < 1e-11


def test_fluxes_and_residual(case):
    name, g, q, prm = case
    o = oracle.Oracle(g, prm)
    src_r, if_r, vf_r, n_r = o.residual(q)
    s = Solver(g, prm)
    s.set_state(q)
    src, ifl, vfl, n = s.residual()
    e_if = face_rel_err(ifl, if_r)
    e_vf = face_rel_err(vfl, vf_r) if prm["viscous"] else 0.0
    e_src = cell_res_err(src, src_r, g, if_r, vf_r)
    print(f"iflux {e_if:.2e} vflux {e_vf:.2e} src {e_src:.2e}")
    assert e_if < RTOL_FLUX
    assert e_vf < RTOL_FLUX
    assert e_src < RTOL_FLUX
    for k in ("rtrms", "etrms", "mtrms"):
        assert abs(getattr(n, k) - getattr(n_r, k)) <= 1e-12 * abs(getattr(n_r, k))
    for k in ("rtmax", "etmax", "mtmax"):
        assert abs(getattr(n, k) - getattr(n_r, k)) <= 1e-10 * abs(getattr(n_r, k)) + 1e-300
    assert np.allclose(list(n.ccmaxe), list(n_r.ccmaxe)) and np.allclose(list(n.ccmaxr), list(n_r.ccmaxr))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_assembled_system(case, prec):
    name, g, q, prm = case
    prm = dict(prm, precision=prec)
    fj = "f64" if prec == "fp64" else "f32"
    o = oracle.Oracle(g, prm, fj)
    ref = o.assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    s.assemble()
    sy = s.get_system()
    tol = 1e-12 if prec == "fp64" else 0.0  # FP32: bit for bit
    inner = (g.btype == 0) | (g.btype == 4)
    for key in ("fjp", "fjm"):
        a = sy[key][inner].astype(np.float64); b = ref[key][inner].astype(np.float64)
        scale = np.abs(b).max(axis=1, keepdims=True) + 1e-300
        err = float((np.abs(a - b) / scale).max())
        print(f"{prec} {key} err {err:.2e}")
        assert err <= tol
    a = sy["D"].astype(np.float64); b = ref["D"].astype(np.float64)
    scale = np.abs(b).max(axis=1, keepdims=True)
    errD = float((np.abs(a - b) / scale).max())
    print(f"{prec} LU(D) err {errD:.2e}")
    assert errD <= (1e-11 if prec == "fp64" else 0.0)
    assert np.array_equal(sy["piv"], ref["piv"])
    eB = float(np.abs(sy["B"].astype(np.float64) - ref["B"].astype(np.float64)).max() /
               np.abs(ref["B"]).max())
    edt = float((np.abs(sy["dtau"] - ref["dtau"]) / ref["dtau"]).max())
    print(f"{prec} B err {eB:.2e} dtau err {edt:.2e}")
    assert eB <= tol and edt < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_sgs_sweeps(case, prec):
    name, g, q, prm = case
    prm = dict(prm, precision=prec)
    fj = "f64" if prec == "fp64" else "f32"
    o = oracle.Oracle(g, prm, fj)
    ref = o.assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    for nsw in (0, 3):
        s.assemble()
        dq = s.linear_solve(nsw)
        dq_r = o.sgs(ref, nsw)
        scale = np.abs(dq_r).max(axis=0, keepdims=True)
        err = float((np.abs(dq - dq_r) / scale).max())
        print(f"{prec} sweeps={nsw} dq err {err:.2e}")
        assert err <= (1e-11 if prec == "fp64" else 0.0)


def test_newton_steps_fp64(case):
    name, g, q, prm = case
    nsteps = 6
    o = oracle.Oracle(g, prm)
    q_r, h_r = o.newton(q, nsteps)
    s = Solver(g, prm)
    s.set_state(q)
    hist = s.newton(nsteps)
    q_g = s.get_state()
    scale = np.abs(q_r).max(axis=1, keepdims=True)
    err = float((np.abs(q_g - q_r) / scale).max())
    tot_r = h_r[:, 0] + h_r[:, 1] + h_r[:, 2]
    tot = np.array([h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist])
    print(f"state err after {nsteps} steps {err:.2e}; residual drop {np.sqrt(tot[-1] / tot[0]):.3e}")
    assert err < RTOL_STATE
    assert np.allclose(tot, tot_r, rtol=1e-9)


def test_newton_fp32_vs_fp64_oracle(case):
    """Mixed precision (real_fj = float) tracks the FP64 oracle; the gap is reported."""
    name, g, q, prm = case
    nsteps = 6
    o64 = oracle.Oracle(g, prm)
    q64, _ = o64.newton(q, nsteps)
    o32 = oracle.Oracle(g, prm, "f32")
    q32r, _ = o32.newton(q, nsteps)
    s = Solver(g, dict(prm, precision="fp32"))
    s.set_state(q)
    s.newton(nsteps)
    q32 = s.get_state()
    scale = np.abs(q64).max(axis=1, keepdims=True)
    e_vs_f32_oracle = float((np.abs(q32 - q32r) / scale).max())
    e_vs_f64 = float((np.abs(q32 - q64) / scale).max())
    print(f"fp32 path: vs fp32 oracle {e_vs_f32_oracle:.2e}, vs fp64 oracle {e_vs_f64:.2e}")
    assert e_vs_f32_oracle <= 1e-15  # the same float Jacobians, FP64 residual and update
    assert e_vs_f64 < 1e-4


def test_bitwise_reproducible():
    g, q, prm = workloads.channel(nx=6, ny=8, nz=4)
    outs = []
    for _ in range(2):
        s = Solver(g, prm)
        s.set_state(q)
        s.newton(3)
        outs.append(s.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_sweep_modes_agree():
    """One launch per level, the cooperative dataflow kernel and the tiled
    kernel give the same bits."""
    g, q, prm = workloads.channel(nx=12, ny=16, nz=10)
    outs = []
    for mode in ("persistent", "per_level", "tiled"):
        s = Solver(g, dict(prm, sweep=mode))
        assert s.info()["sweep_tiles"] > 0
        s.set_state(q)
        s.newton(2)
        outs.append(s.get_state())
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[1], outs[2])


@pytest.mark.parametrize("mk,prec", [("box40", "fp64"), ("channel64", "fp64"), ("channel64", "fp32")])
def test_cooperative_sweeps_bitwise_at_size(mk, prec):
    """More tiles / chunks than co-resident workgroups (each workgroup walks
    many): the tiled and the dataflow sweeps equal the level-per-launch sweep
    bit for bit."""
    if mk == "box40":
        g, q, prm = workloads.box(40)
    else:
        g, q, prm = workloads.channel(64)
    prm = dict(prm, precision=prec)
    outs = []
    for mode in ("per_level", "tiled", "persistent"):
        s = Solver(g, dict(prm, sweep=mode))
        info = s.info()
        s.set_state(q)
        s.assemble()
        outs.append(s.linear_solve(3))
    print(f"{mk} {prec}: {info['sweep_tiles']} tiles, {info['nlevels']} levels")
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


def test_c1_box_full_size_fifty_steps():
    """The 32^3 Euler box of the Roe/farfield workload at its full size and
    step count: 50 Newton steps of the HIP path against the oracle."""
    g, q, prm = workloads.box(32)
    q_r, h_r = oracle.Oracle(g, prm).newton(q, 50)
    s = Solver(g, prm)
    s.set_state(q)
    hist = s.newton(50)
    q_g = s.get_state()
    scale = np.abs(q_r).max(axis=1, keepdims=True)
    err = float((np.abs(q_g - q_r) / scale).max())
    tot_r = h_r[:, 0] + h_r[:, 1] + h_r[:, 2]
    tot = np.array([h["rtrms"] + h["etrms"] + h["mtrms"] for h in hist])
    print(f"box32 x50: state err {err:.2e}, residual drop {np.sqrt(tot[-1] / tot[0]):.3e}")
    assert err < RTOL_STATE
    assert np.allclose(tot, tot_r, rtol=1e-9)
