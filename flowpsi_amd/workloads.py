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
