"""The Newton-step workloads: grid, initial state and options facts.

box32   -- inviscid Euler, 32^3 uniform hex box on the unit cube, farfield on
           all sides, Mach 0.3 freestream (p = 101325 Pa, T = 300 K) plus
           seeded 1% Gaussian perturbations of rho and p, Roe flux with
           Venkatakrishnan-limited second-order reconstruction, CFL 10,
           5 SGS sweeps, 50 Newton steps.
channel128 -- laminar Navier-Stokes, 128^3 channel, periodic x/z, isothermal
           no-slip walls at y = +-h with 1.1 geometric stretching, Mach 0.2
           Poiseuille flow plus seeded 2% velocity noise, Re = 500 on h,
           Sutherland viscosity, HLLC + viscous fluxes, CFL 10, 10 SGS sweeps.
"""
from __future__ import annotations

import numpy as np

from . import grids

A300 = float(np.sqrt(grids.GAMMA * grids.R_AIR * 300.0))

BASE = dict(gamma=grids.GAMMA, Rt=grids.R_AIR, Pambient=0.0, limiter="venkatakrishnan", Kl=1.0,
            Twall=300.0, pg0=101325.0, T0=300.0, dtmax=1e30, urelax=1e30, solver="sgs",
            precision="fp64")


def box(n=32, seed=1, **over):
    g = grids.box_grid(n)
    q = grids.box_state(g, seed=seed)
    prm = dict(BASE, flux="roe", viscous=0, u0=(0.3 * A300, 0.0, 0.0), cflmax=10.0, sgs_iter=5)
    prm.update(over)
    return g, q, prm


def channel(n=128, seed=2, nx=None, ny=None, nz=None, **over):
    g = grids.channel_grid(nx or n, ny or n, nz or n)
    q = grids.channel_state(g, seed=seed)
    prm = dict(BASE, flux="hllc", viscous=1, u0=(0.0, 0.0, 0.0), cflmax=10.0, sgs_iter=10)
    prm.update(over)
    return g, q, prm


def channel_sa(n=32, seed=2, nx=None, ny=None, nz=None, nut_ratio=3.0, **over):
    """The channel with the Spalart-Allmaras model: nuTilde = nut_ratio x the
    laminar kinematic viscosity plus 10 % seeded noise, wall distance in
    g.meta["dist"]."""
    g, q5, prm = channel(n, seed, nx, ny, nz)
    rng = __import__("numpy").random.default_rng(seed + 100)
    np_ = __import__("numpy")
    rho = q5[4] / (grids.R_AIR * q5[0])
    nu = grids.sutherland_mu(q5[0]) / rho
    q = np_.vstack([q5, (nut_ratio * nu * (1.0 + 0.1 * rng.standard_normal(g.nc)))[None, :]])
    g.meta["dist"] = grids.wall_distance(g)
    prm = dict(prm, turb="sa")
    prm.update(over)
    return g, q, prm


def channel_sst(n=32, seed=2, nx=None, ny=None, nz=None, tmu_ratio=10.0, **over):
    """The channel with Menter's SST model: k = 1.5 (5 % of the peak speed)^2
    and omega = k / (tmu_ratio nu) (each with seeded noise) plus the
    near-wall 6 nu / (beta1 d^2), wall distance in g.meta["dist"]."""
    import numpy as np_
    g, q5, prm = channel(n, seed, nx, ny, nz)
    rng = np_.random.default_rng(seed + 200)
    rho = q5[4] / (grids.R_AIR * q5[0])
    nu = grids.sutherland_mu(q5[0]) / rho
    dist = grids.wall_distance(g)
    umax = float(np_.abs(q5[1:4]).max())
    k = 1.5 * (0.05 * umax) ** 2 * (1.0 + 0.2 * rng.uniform(-1, 1, g.nc))
    w = k / (tmu_ratio * nu) * (1.0 + 0.1 * rng.uniform(-1, 1, g.nc)) + 6.0 * nu / (0.075 * dist ** 2)
    q = np_.vstack([q5, k[None, :], w[None, :]])
    g.meta["dist"] = dist
    prm = dict(prm, turb="sst")
    prm.update(over)
    return g, q, prm


WORKLOADS = {"box32": dict(fn=box, steps=50), "channel128": dict(fn=channel, steps=None)}


def _freestream(mach, re_per_m=None, T=300.0, p=101325.0):
    """(pressure, speed) of a freestream at `mach`; with `re_per_m` the
    pressure is set so that rho U / mu = re_per_m (Sutherland mu at T)."""
    import numpy as np_
    a = np_.sqrt(grids.GAMMA * grids.R_AIR * T)
    U = mach * a
    if re_per_m is not None:
        rho = re_per_m * grids.sutherland_mu(T) / U
        p = rho * grids.R_AIR * T
    return p, U


def c3_tets(n=88, seed=3, **over):
    """Unit cube of n^3 hexes cut into 6 jittered tetrahedra each (interior
    nodes moved by U(-0.1, 0.1) h), viscous wall at y = 0, farfield elsewhere;
    Spalart-Allmaras, M = 0.5 along x, Re = 1e6 per metre, nuTilde = 3 nu;
    FP32 Jacobians (the mixed-precision path)."""
    import numpy as np_
    g = grids.tet_cube(n, 0.1, seed)
    g.meta["dist"] = grids.wall_distance(g)
    p, U = _freestream(0.5, 1e6)
    q5 = grids.freestream_state(g, 0.5, p=p, seed=seed)
    rho = q5[4] / (grids.R_AIR * q5[0])
    nu = grids.sutherland_mu(q5[0]) / rho
    q = np_.vstack([q5, (3.0 * nu)[None, :]])
    prm = dict(BASE, flux="hllc", viscous=1, u0=(U, 0.0, 0.0), pg0=p, cflmax=10.0, sgs_iter=10, turb="sa",
               precision="fp32")
    prm.update(over)
    return g, q, prm


def c4_hybrid(nx=8, nz=8, n_prism=8, n_hex=8, seed=4, **over):
    """Hex core (interior nodes jittered by 10 %) over prism layers at a wall
    (first height 1e-4, growth 1.2); Spalart-Allmaras, M = 0.3 along x,
    Re = 1e6 per metre, nuTilde = 3 nu."""
    import numpy as np_
    g = grids.prism_hex_box(nx, nz, n_prism, n_hex, first=1e-4, growth=1.2, jitter=0.1, seed=seed)
    g.meta["dist"] = grids.wall_distance(g)
    p, U = _freestream(0.3, 1e6)
    q5 = grids.freestream_state(g, 0.3, p=p, seed=seed)
    rho = q5[4] / (grids.R_AIR * q5[0])
    nu = grids.sutherland_mu(q5[0]) / rho
    q = np_.vstack([q5, (3.0 * nu)[None, :]])
    prm = dict(BASE, flux="hllc", viscous=1, u0=(U, 0.0, 0.0), pg0=p, cflmax=10.0, sgs_iter=10, turb="sa")
    prm.update(over)
    return g, q, prm


def c5_bl(nx=16, ny=60, nz=16, aspect=1e3, seed=5, side=grids.SYMMETRY, **over):
    """Boundary-layer hex grid (wall at y = 0, first cell aspect ratio
    `aspect`, geometric growth to y = 1); Menter SST, M = 0.6 along x,
    Re = 5e6 per metre; k and omega from the inflow defaults with seeded
    noise plus the near-wall 6 nu / (beta1 d^2); line SGS; Venkatakrishnan
    limiter (Kl = 1), CFL 100 with flowPsi's default urelax = 0.2
    (src/solverTimestep.loci:281-283).  The spanwise planes z = 0, 1 are
    symmetry planes (`side`): with farfield there, the characteristic
    boundary switches between its inflow and outflow states where the wall
    meets it (w ~ 0) and the residual stalls near 3e-6; with symmetry the
    grid converges to machine level (with urelax unset, CFL 10 crawls: a
    3e-4 drop after 800 steps)."""
    import numpy as np_
    g = grids.bl_box(nx, ny, nz, aspect, side=side)
    g.meta["dist"] = grids.wall_distance(g)
    p, U = _freestream(0.6, 5e6)
    q5 = grids.freestream_state(g, 0.6, p=p, seed=seed)
    rho = q5[4] / (grids.R_AIR * q5[0])
    nu = grids.sutherland_mu(q5[0]) / rho
    rng = np_.random.default_rng(seed + 300)
    kb = max(1.5 * (1e-2 * U) ** 2, 1e-3)
    k = kb * (1.0 + 0.1 * rng.uniform(-1, 1, g.nc))
    w = 0.09 * kb / 1e-8 * (1.0 + 0.1 * rng.uniform(-1, 1, g.nc)) + 6.0 * nu / (0.075 * g.meta["dist"] ** 2)
    q = np_.vstack([q5, k[None, :], w[None, :]])
    prm = dict(BASE, flux="hllc", viscous=1, u0=(U, 0.0, 0.0), pg0=p, cflmax=100.0, urelax=0.2, sgs_iter=10,
               turb="sst", solver="lsgs")
    prm.update(over)
    return g, q, prm


# ---- the reference's own grids (quickTest/*.vog, kept under tests/golden/vog) ----

def vog_path(name):
    import os
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden", "vog",
                        name + ".vog")


def vog_case(name, seed=6, **over):
    """A Newton-step workload on one of the reference's VOG grids, with the
    boundary types this path implements ('reflecting' walls of the
    reference's cases become symmetry planes, inflow/outflow conditions
    farfield): testGrid (inviscid, farfield box), nozzle (inviscid,
    axisymmetric-slice nozzle), periodic (laminar, periodic channel with a
    blockage wall), shock_tube (inviscid; the diaphragm state p = 1e6 / 1e5
    Pa at x = 0 of the reference's case)."""
    import numpy as np_
    from . import vog
    F, W, S, P = grids.FARFIELD, grids.WALL, grids.SYMMETRY, grids.PERIODIC
    spec = {
        "testGrid": (dict(BC_1=F, BC_2=F, BC_3=F, BC_4=F, BC_5=F, BC_6=F), None, 0, 0.3),
        "nozzle": (dict(axis=S, **{"in": F}, nozzle=S, out=F, symmetry1=S, symmetry2=S), None, 0, 0.3),
        "periodic": (dict(blockage=W, bottom=P, top=P, **{"in": F}, out=F, BC_8=S, BC_9=S), {"top": "bottom"}, 1,
                     0.1),
        "shock_tube": (dict(BC_1=S, BC_2=S, BC_3=S, BC_4=S, BC_5=S, BC_6=S), None, 0, 0.0),
    }[name]
    bcs, per, visc, mach = spec
    g = vog.load(vog_path(name), bcs, per, name)
    g.check()
    if name == "shock_tube":
        q = np_.zeros((5, g.nc))
        q[0] = 300.0
        q[4] = np_.where(g.cc[0] < 0.0, 1e6, 1e5)
        u0 = (0.0, 0.0, 0.0)
    else:
        q = grids.freestream_state(g, mach, seed=seed)
        u0 = (mach * A300, 0.0, 0.0)
    if visc:
        g.meta["dist"] = grids.wall_distance(g)
    prm = dict(BASE, flux="hllc" if visc else "roe", viscous=visc, u0=u0, cflmax=10.0, sgs_iter=5)
    prm.update(over)
    return g, q, prm
