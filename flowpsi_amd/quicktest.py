"""flowPsi's quickTest regression cases on this path.

quickTest/inviscid/tests.sh runs a case, extracts nodal values with Loci's
`extract -ascii <case> 0 <vars>` and compares them with the committed
`<case>.dat` through tools/ndiff.cc at the tolerances of `<case>.tol.dat`.
This module holds the same cases (options restated from each
quickTest/inviscid/<case>.vars), the cell-to-node extract of the state and
the comparison.

Cases whose boundary conditions and options this path builds:
  inflow      testGrid, inflow + outflow + symmetry, 500 steady steps
  isentropic  testGrid, isentropicInflow + outflow + symmetry, 400 steps
  superSonic  testGrid, supersonicInflow + supersonicOutflow + symmetry, 400
  fixedMass   testGrid, laminar, fixedMass + outflow + reflecting, 300
  shock_tube  shock_tube, BDF2 dt 1e-5, 600 time steps of 12 Newton
              iterations, reflecting ends, Venkatakrishnan limiter
Not built: cylinder (its .vog is not among the fixtures), nozzle
(gridCoordinates: axisymmetric), wedge (no fixture grid).
"""
from __future__ import annotations

import os

import numpy as np

from . import grids, vog

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# src/solverEoS.loci:26-36: air, molecularMass 28.96
RT_AIR = 8314.4621 / 28.96
GAMMA = 1.4

S, R = grids.SYMMETRY, grids.REFLECTING

# name -> options of quickTest/inviscid/<name>.vars; boundary values are
# (type, option dict) with the options' SI values; initialConditions as
# (p, T, u or M); the .tol.dat tolerances (ndiff tol, min_error)
CASES = {
    "inflow": dict(grid="testGrid", viscous=0, limiter="none", steps=500, dtmax=5e-4, urelax=0.2,
                   bcs={"BC_1": S, "BC_2": S, "BC_3": S, "BC_4": S,
                        "BC_5": (grids.INFLOW, dict(p=1e5, T=300.0, u=100.0)),
                        "BC_6": (grids.OUTFLOW, dict(p=1e5))},
                   ic=dict(p=1e5, T=300.0, u=90.0), tol=(1e-5, 1e-5)),
    "isentropic": dict(grid="testGrid", viscous=0, limiter="none", steps=400, dtmax=1e-3, urelax=0.2,
                       bcs={"BC_1": S, "BC_2": S, "BC_3": S, "BC_4": S,
                            "BC_5": (grids.ISENTROPIC_INFLOW, dict(p0=1e5, T0=300.0)),
                            "BC_6": (grids.OUTFLOW, dict(p=9e4))},
                       ic=dict(p=9e4, T=300.0, M=0.4), tol=(1e-5, 1e-5)),
    "superSonic": dict(grid="testGrid", viscous=0, limiter="none", steps=400, dtmax=1e-3, urelax=0.2,
                       bcs={"BC_1": S, "BC_2": S, "BC_3": S, "BC_4": S,
                            "BC_5": (grids.SUPERSONIC_INFLOW, dict(p=1e5, T=300.0, u=500.0)),
                            "BC_6": (grids.SUPERSONIC_OUTFLOW, {})},
                       ic=dict(p=9e4, T=300.0, u=0.0), tol=(1e-5, 1e-5)),
    "fixedMass": dict(grid="testGrid", viscous=1, limiter="none", steps=300, dtmax=5e-4, urelax=0.2,
                      bcs={"BC_1": R, "BC_2": R, "BC_3": R, "BC_4": R,
                           "BC_5": (grids.FIXED_MASS, dict(mdot=144.0, T0=400.0)),
                           "BC_6": (grids.OUTFLOW, dict(p=9e4))},
                      ic=dict(p=9e4, T=300.0, M=0.4), tol=(1e-5, 1e-5)),
    "shock_tube": dict(grid="shock_tube", viscous=0, limiter="venkatakrishnan", steps=600, dtmax=1e-5,
                       urelax=0.9, unsteady=True, newton_iter=12,
                       bcs={"BC_1": S, "BC_2": S, "BC_3": S, "BC_4": S, "BC_5": R, "BC_6": R},
                       ic=dict(regions=(dict(p=1e6, T=300.0, u=0.0), dict(p=1e5, T=300.0, u=0.0))),
                       tol=(1e-3, 1e-3)),
}


def _velocity(d, T):
    """initialConditions / BC velocity: u (a scalar is the x component) or
    M times the sound speed along x"""
    if "M" in d:
        return (d["M"] * np.sqrt(GAMMA * RT_AIR * T), 0.0, 0.0)
    u = d.get("u", 0.0)
    return (float(u), 0.0, 0.0) if np.isscalar(u) else tuple(float(x) for x in u)


def setup(name):
    """(grid, state q[5][nc], parameter dict, the VOG read result) of a case"""
    c = CASES[name]
    v = vog.read(os.path.join(ROOT, "tests", "golden", "vog", c["grid"] + ".vog"))
    codes, sets = {}, []
    for sname, bc in c["bcs"].items():
        if isinstance(bc, tuple):
            t, opt = bc
            b = {}
            if "T" in opt:
                b["T"] = opt["T"]
                b["u"] = _velocity(opt, opt["T"])
            if "p" in opt:
                b["p"] = opt["p"]
            for k in ("p0", "T0", "mdot"):
                if k in opt:
                    b[k] = opt[k]
            codes[sname] = grids.bc_code(t, len(sets))
            sets.append(b)
        else:
            codes[sname] = bc
    g = vog.to_grid(v, codes, None, name)
    g.check()
    q = np.zeros((5, g.nc))
    ic = c["ic"]
    if "regions" in ic:  # leftPlane(point 0, normal +x): the driver state at x < 0
        drv, dfl = ic["regions"]
        left = g.cc[0] < 0.0
        for m, st in ((left, drv), (~left, dfl)):
            u = _velocity(st, st["T"])
            q[0, m] = st["T"]
            q[1, m], q[2, m], q[3, m] = u
            q[4, m] = st["p"]
    else:
        u = _velocity(ic, ic["T"])
        q[0] = ic["T"]
        q[1], q[2], q[3] = (np.full(g.nc, x) for x in u)
        q[4] = ic["p"]
    u0 = _velocity(ic, ic["T"]) if "T" in ic else (0.0, 0.0, 0.0)
    prm = dict(gamma=GAMMA, Rt=RT_AIR, Pambient=0.0, flux="hllc", limiter=c["limiter"], Kl=1.0,
               viscous=c["viscous"], cflmax=0.0, dtmax=c["dtmax"], urelax=c["urelax"], solver="lsgs",
               u0=u0, pg0=ic.get("p", 1e5), T0=ic.get("T", 300.0), bc_sets=sets)
    if c.get("unsteady"):
        prm.update(unsteady=True, time_scheme="bdf2", newton_iter=c["newton_iter"])
    return g, q, prm, v


def cell2node(v, vals, cc):
    """Nodal values of cell fields vals[k][nc]: each node takes the
    inverse-distance-weighted mean of the cells that hold it (a convex
    combination, so a uniform field extracts to itself).  Loci's own
    cell2node is not in the reference; only uniform fields and fields
    whose node neighbourhoods are symmetric pin this choice."""
    pos, off, fn, cl, cr = v["pos"], v["face_off"], v["face_nodes"], v["cl"], v["cr"]
    nn = pos.shape[0]
    cnt = np.diff(off)
    fidx = np.repeat(np.arange(len(cnt)), cnt)
    pairs = [np.stack([fn, cl[fidx]], 1)]
    inter = cr[fidx] >= 0
    pairs.append(np.stack([fn[inter], cr[fidx][inter]], 1))
    nc_pairs = np.unique(np.concatenate(pairs), axis=0)
    node, cell = nc_pairs[:, 0], nc_pairs[:, 1]
    w = 1.0 / np.linalg.norm(pos[node] - cc.T[cell], axis=1)
    den = np.bincount(node, w, nn)
    return np.stack([np.bincount(node, w * f[cell], nn) / den for f in vals])


def extract(v, g, q, prm, comps=False):
    """`extract -ascii <case> 0 P t u` (with comps: also the velocity
    components 0 1 2): P = pg + Pambient, t, and |u| of the nodal velocity"""
    nodal = cell2node(v, [q[4] + prm.get("Pambient", 0.0), q[0], q[1], q[2], q[3]], g.cc)
    cols = [nodal[0], nodal[1], np.sqrt(nodal[2] ** 2 + nodal[3] ** 2 + nodal[4] ** 2)]
    if comps:
        cols += [nodal[2], nodal[3], nodal[4]]
    return np.stack(cols, 1)


def golden(name):
    """the reference's quickTest/inviscid/<name>.dat (kept as a fixture)"""
    return np.loadtxt(os.path.join(ROOT, "tests", "golden", "quicktest", name + ".dat"), ndmin=2)


def ndiff(a, b, tol, min_error=0.0):
    """tools/ndiff.cc:84-119: per column the rms of the difference over the
    rms of max(a^2, b^2); a column fails when that relative error exceeds
    tol and the rms difference exceeds min_error.  Returns (ok, relative
    errors)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if a.shape != b.shape:
        return False, None
    delta = np.sqrt(((a - b) ** 2).sum(0))
    rms = np.sqrt(np.maximum(a * a, b * b).sum(0))
    rel = delta / (rms + min_error)
    ok = not np.any((delta > min_error) & (rel > tol))
    return ok, rel


def run_gpu(name, **over):
    """The case on the GPU through the C ABI, line SGS as one dataflow
    launch per half-sweep where the grid allows it: (nodal columns, state)"""
    from .capi import Solver
    g, q, prm, v = setup(name)
    prm.update(over)
    c = CASES[name]
    s = Solver(g, dict(prm, sweep="per_level"))
    try:
        if s.info()["lsgs_dataflow"]:
            s.set_params(dict(prm, sweep="persistent_flags"))
        s.set_state(q)
        if c.get("unsteady"):
            s.unsteady(c["steps"])
        else:
            s.newton(c["steps"])
        qs = s.get_state()
    finally:
        s.close()
    return extract(v, g, qs, prm, comps=False), qs
