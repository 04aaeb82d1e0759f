"""The general polyhedral grids of the turbulent workloads (no GPU): jittered
tetrahedra (6 per hex), prism layers under a jittered hex core (with 7-faced
transition cells), and a high-aspect-ratio boundary-layer box.

* every cell is closed (sum of outward area vectors 0), volumes positive and
  summing to the box volume, cl < cr on interior faces;
* cell and face counts of each construction;
* the projected wall distance of cells over a flat wall is their height;
* LSQ gradients of linear fields are exact on every cell (face values of
  the same field on the boundary);
* a few oracle Newton steps stay finite with positive turbulence scalars.
"""
import numpy as np
import pytest

import oracle
from flowpsi_amd import grids, workloads


def test_tet_cube():
    n = 5
    g = grids.tet_cube(n)
    g.check()
    assert g.nc == 6 * n ** 3
    assert np.all(g.vol > 0) and g.vol.sum() == pytest.approx(1.0, rel=1e-12)
    assert np.sum(g.btype != grids.INTERIOR) == 12 * n * n
    assert np.sum(g.btype == grids.WALL) == 2 * n * n
    d = grids.wall_distance(g)
    assert np.allclose(d, g.cc[1], rtol=0, atol=1e-14)


def test_prism_hex_box():
    nx, nz, npr, nh = 4, 3, 5, 4
    g = grids.prism_hex_box(nx, nz, npr, nh, first=1e-3, growth=1.2, H=1.0)
    g.check()
    assert g.nc == 2 * nx * nz * npr + nx * nz * nh
    assert np.all(g.vol > 0) and g.vol.sum() == pytest.approx(1.0, rel=1e-12)
    faces_per_cell = np.bincount(g.cl, minlength=g.nc) + np.bincount(g.cr[g.cr >= 0], minlength=g.nc)
    assert sorted(set(faces_per_cell)) == [5, 6, 7]
    assert np.sum(faces_per_cell == 7) == nx * nz          # the transition layer
    prism = g.cc[1] < 1e-3 * (1.2 ** npr - 1) / 0.2
    d = grids.wall_distance(g)
    assert np.allclose(d[prism], g.cc[1][prism], rtol=1e-12)


def test_bl_box_aspect():
    g = grids.bl_box(8, 30, 8, aspect=1e3)
    g.check()
    dy0 = 2 * g.cc[1].min()
    assert (1.0 / 8) / dy0 == pytest.approx(1e3, rel=1e-9)


@pytest.mark.parametrize("mk", ["tets", "prism_hex"])
def test_lsq_gradient_exact_for_linear_fields(mk):
    g = grids.tet_cube(4) if mk == "tets" else grids.prism_hex_box(4, 4, 3, 3)
    g.meta["dist"] = grids.wall_distance(g)
    prm = dict(workloads.BASE, flux="hllc", viscous=0, limiter="none", u0=(0, 0, 0))
    a = np.array([[1.0, 2.0, -3.0], [0.5, -1.0, 2.0], [2.0, 0.3, 1.0], [-0.7, 1.1, 0.2], [100.0, -50.0, 25.0]])
    base = np.array([300.0, 10.0, -5.0, 3.0, 1e5])
    q = base[:, None] + a @ g.cc
    qf = base[:, None] + a @ g.fc
    grad, phi = oracle.Oracle(g, prm).gradients(q, qf)
    for k in range(5):
        for d in range(3):
            assert np.allclose(grad[3 * k + d], a[k, d], rtol=1e-9, atol=1e-9 * np.abs(a[k]).max())


@pytest.mark.parametrize("mk", ["c3", "c4", "c5"])
def test_oracle_newton_on_the_turbulent_grids(mk):
    fn = {"c3": lambda: workloads.c3_tets(4), "c4": lambda: workloads.c4_hybrid(4, 4, 4, 3),
          "c5": lambda: workloads.c5_bl(4, 12, 4)}[mk]
    g, q, prm = fn()
    o = oracle.Oracle(g, prm, "f32" if prm["precision"] == "fp32" else "f64")
    q2, h = o.newton(q, 3)
    assert np.all(np.isfinite(q2)) and np.all(q2[5:] >= 1e-15)
    assert h[-1, :3].sum() < h[0, :3].sum()
