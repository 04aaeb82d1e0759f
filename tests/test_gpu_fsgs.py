"""Parity of FSGS (linear_solver = FPSI_SOLVER_FSGS, through the C ABI)
against the oracle's restatement of src/linearSolverFSGS.loci: the solve
from zero for 0-6 sweeps (same arithmetic in the same order, so bitwise),
Newton steps, and the turbulence scalars' FSGS under SA and SST."""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads
from flowpsi_amd.capi import Solver

pytestmark = pytest.mark.gpu


def _shuffled_box():
    g, q, prm = workloads.box(6, flux="hllc")
    perm = np.random.default_rng(17).permutation(g.nc)
    g2, q2 = grids.renumber_cells(g, perm, q)
    return g2, q2, prm


def _cases():
    return {
        "box_roe": lambda: workloads.box(8),
        "channel_visc": lambda: workloads.channel(nx=6, ny=10, nz=4),
        "box_renumbered": _shuffled_box,
    }


@pytest.fixture(scope="module", params=list(_cases()))
def case(request):
    g, q, prm = _cases()[request.param]()
    print(f"\n[fsgs case {request.param}] nc={g.nc} nf={g.nf}", flush=True)
    return request.param, g, q, dict(prm, solver="fsgs")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_solve(case, prec):
    name, g, q, prm = case
    prm = dict(prm, precision=prec)
    o = oracle.Oracle(g, prm, "f64" if prec == "fp64" else "f32")
    ref = o.assemble(q)
    s = Solver(g, prm)
    s.set_state(q)
    for nsw in (0, 1, 3, 6):
        s.assemble()
        dq = s.linear_solve(nsw)
        dq_r = o.fsgs(ref, nsw)
        print(f"{prec} sweeps={nsw}: max diff {np.abs(dq - dq_r).max():.1e}")
        assert np.array_equal(dq, dq_r)


def test_newton_steps(case):
    name, g, q, prm = case
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-6)):
        p = dict(prm, precision=prec)
        q_r, h_r = oracle.Oracle(g, p, "f64" if prec == "fp64" else "f32").newton(q, 5)
        s = Solver(g, p)
        s.set_state(q)
        s.newton(5)
        q_g = s.get_state()
        err = float((np.abs(q_g - q_r) / np.abs(q_r).max(axis=1, keepdims=True)).max())
        print(f"{prec} fsgs state err after 5 steps {err:.2e}")
        assert err < tol


@pytest.mark.parametrize("turb", ["sa", "sst"])
def test_turbulence_scalars(turb):
    g, q, prm = (workloads.channel_sa if turb == "sa" else workloads.channel_sst)(nx=5, ny=8, nz=4)
    prm = dict(prm, solver="fsgs")
    o = oracle.Oracle(g, prm)
    if turb == "sa":
        systems = [o.sa_assemble(q)]
    else:
        full = o.sst_assemble(q)
        systems = [oracle.Oracle.scalar(full, j) for j in (0, 1)]
    s = Solver(g, prm)
    s.set_state(q)
    s.assemble()
    for j, sy in enumerate(systems):
        for nsw in (1, 4):
            dX = s.turb_solve(j, nsw)
            ref = o.sa_fsgs(sy, nsw)
            e = float(np.abs(dX - ref).max() / np.abs(ref).max())
            print(f"{turb} scalar {j} fsgs sweeps={nsw} err {e:.1e}")
            assert e < 1e-10
    q_r, _ = o.newton(q, 4)
    s.set_state(q)
    s.newton(4)
    q_g = s.get_state()
    err = float((np.abs(q_g - q_r) / np.abs(q_r).max(axis=1, keepdims=True)).max())
    print(f"{turb} fsgs Newton state err {err:.2e}")
    assert err < 1e-10
