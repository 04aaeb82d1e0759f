"""Seeded synthetic unstructured face grids for the Newton-step workloads.

Every grid is returned as a general unstructured face list, in the form the
flowPsi rules consume (Loci FVM facts, used at src/fluxInviscid.loci:1170-1188
and src/solverLinearSystem.loci:106-120):

* faces carry ``cl``/``cr`` (interior faces have ``cl < cr``; boundary faces
  have ``cr = -1`` and a boundary type), a unit normal pointing from ``cl`` to
  ``cr`` (out of the domain on boundaries), an area and a centre;
* cells carry a volume and a centre;
* periodic faces (src/solverPeriodic.loci:25-87) name their partner face
  ``pface`` and the shift that carries the partner's cell centre to the
  ghost position.

Because every interior face has ``cl < cr``, cell index order is a
topological order of the ``cl -> cr`` face graph that block Gauss-Seidel
follows (src/linearSolverSGS.loci:178-182).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

INTERIOR, FARFIELD, WALL, SYMMETRY, PERIODIC = 0, 1, 2, 3, 4
BC_NAMES = {INTERIOR: "interior", FARFIELD: "farfield", WALL: "viscousWall",
            SYMMETRY: "symmetry", PERIODIC: "periodic"}


@dataclass
class Grid:
    nc: int
    nf: int
    cl: np.ndarray          # int32 [nf]
    cr: np.ndarray          # int32 [nf]
    btype: np.ndarray       # int32 [nf]
    pface: np.ndarray       # int32 [nf]
    n: np.ndarray           # float64 [3, nf]
    area: np.ndarray        # float64 [nf]
    fc: np.ndarray          # float64 [3, nf]
    vol: np.ndarray         # float64 [nc]
    cc: np.ndarray          # float64 [3, nc]
    pshift: np.ndarray      # float64 [3, nf]
    name: str = ""
    meta: dict = field(default_factory=dict)

    def check(self) -> None:
        """Structural invariants: orientation, closure of every cell."""
        assert self.cl.shape == (self.nf,) and self.cr.shape == (self.nf,)
        inter = self.btype == INTERIOR
        assert np.all(self.cr[inter] > self.cl[inter]), "interior faces must have cl < cr"
        assert np.all(self.cr[~inter] == -1)
        per = self.btype == PERIODIC
        assert np.all(self.pface[per] >= 0)
        assert np.all(self.pface[self.pface[per]] == np.nonzero(per)[0])
        # closure: sum of outward area vectors per cell = 0
        s = np.zeros((3, self.nc))
        av = self.n * self.area
        for d in range(3):
            np.add.at(s[d], self.cl, av[d])
            np.subtract.at(s[d], self.cr[inter], av[d][inter])
        scale = np.sqrt(self.area.max() * self.area.min())
        assert np.abs(s).max() < 1e-9 * scale, "cells are not closed"


def renumber_cells(g: Grid, perm, q=None):
    """The same grid with cell i renamed perm[i]; interior faces whose new cl
    exceeds the new cr are turned round (cl and cr swapped, normal negated)
    so that cl < cr still holds.  Returns (grid, q renumbered or None)."""
    perm = np.asarray(perm, np.int64)
    assert np.array_equal(np.sort(perm), np.arange(g.nc))
    inter = g.cr >= 0
    cl = perm[g.cl]
    cr = np.where(inter, perm[np.where(inter, g.cr, 0)], -1)
    flip = inter & (cl > cr)
    n = g.n.copy()
    n[:, flip] *= -1.0
    cl2 = np.where(flip, cr, cl).astype(np.int32)
    cr2 = np.where(flip, cl, cr).astype(np.int32)
    vol = np.empty_like(g.vol)
    vol[perm] = g.vol
    cc = np.empty_like(g.cc)
    cc[:, perm] = g.cc
    meta = dict(g.meta)
    if "dist" in meta:  # per-cell wall distance follows its cell
        d = np.empty_like(meta["dist"])
        d[perm] = meta["dist"]
        meta["dist"] = d
    out = Grid(g.nc, g.nf, cl2, cr2, g.btype.copy(), g.pface.copy(), n, g.area.copy(), g.fc.copy(), vol, cc,
               g.pshift.copy(), name=g.name + "_renumbered", meta=meta)
    q2 = None
    if q is not None:
        q2 = np.empty_like(q)
        q2[:, perm] = q
    return out, q2


def _axis_coords(n, lo, hi, kind="uniform", ratio=1.0):
    if kind == "uniform":
        return np.linspace(lo, hi, n + 1)
    if kind == "two_sided_geometric":
        # n cells, n/2 per half, geometric growth `ratio` away from each end
        half = n // 2
        w = ratio ** np.arange(half)
        w = w / w.sum() * (hi - lo) / 2.0
        left = lo + np.concatenate([[0.0], np.cumsum(w)])
        right = hi - np.concatenate([[0.0], np.cumsum(w)])[::-1]
        x = np.concatenate([left, right[1:]])
        x[half] = 0.5 * (lo + hi)
        return x
    raise ValueError(kind)


def _quad(p0, p1, p2, p3):
    """Area vector (1/2 of the diagonal cross product), centre (node mean)."""
    d1 = p2 - p0
    d2 = p3 - p1
    av = 0.5 * np.cross(d1, d2, axis=0)
    ctr = 0.25 * (p0 + p1 + p2 + p3)
    return av, ctr


def structured_hex(nodes, bc, periodic_shift=None, name=""):
    """Hex grid from a node lattice ``nodes[3, nx+1, ny+1, nz+1]``.

    ``bc`` maps each side ("xmin", "xmax", ...) to a boundary type; a pair of
    opposite sides both set to PERIODIC are joined, with ``periodic_shift``
    giving the period vector of that axis.
    """
    _, NX, NY, NZ = nodes.shape
    nx, ny, nz = NX - 1, NY - 1, NZ - 1
    nc = nx * ny * nz

    def cid(i, j, k):
        return i + nx * (j + ny * k)

    X = nodes
    cc = 0.125 * (X[:, :-1, :-1, :-1] + X[:, 1:, :-1, :-1] + X[:, :-1, 1:, :-1] + X[:, 1:, 1:, :-1]
                  + X[:, :-1, :-1, 1:] + X[:, 1:, :-1, 1:] + X[:, :-1, 1:, 1:] + X[:, 1:, 1:, 1:])
    cc = cc.transpose(0, 3, 2, 1).reshape(3, nc)  # i fastest

    # face lists per axis: quad on the plane between cells, normal along +axis
    def xfaces(i):  # plane i, all j,k ; normal +x
        p0 = X[:, i, :-1, :-1]; p1 = X[:, i, 1:, :-1]; p2 = X[:, i, 1:, 1:]; p3 = X[:, i, :-1, 1:]
        return _quad(p0, p1, p2, p3)

    def yfaces(j):
        p0 = X[:, :-1, j, :-1]; p1 = X[:, :-1, j, 1:]; p2 = X[:, 1:, j, 1:]; p3 = X[:, 1:, j, :-1]
        return _quad(p0, p1, p2, p3)

    def zfaces(k):
        p0 = X[:, :-1, :-1, k]; p1 = X[:, 1:, :-1, k]; p2 = X[:, 1:, 1:, k]; p3 = X[:, :-1, 1:, k]
        return _quad(p0, p1, p2, p3)

    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cl_l, cr_l, av_l, fc_l, bt_l, side_l = [], [], [], [], [], []

    # interior faces: x, then y, then z
    if nx > 1:
        av, ctr = xfaces(slice(1, nx))                # [3, nx-1, ny, nz]
        cl_l.append(cid(I[:-1], J[:-1], K[:-1]).transpose(2, 1, 0).ravel())
        cr_l.append(cid(I[1:], J[1:], K[1:]).transpose(2, 1, 0).ravel())
        av_l.append(av.transpose(0, 3, 2, 1).reshape(3, -1)); fc_l.append(ctr.transpose(0, 3, 2, 1).reshape(3, -1))
    if ny > 1:
        av, ctr = yfaces(slice(1, ny))
        cl_l.append(cid(I[:, :-1], J[:, :-1], K[:, :-1]).transpose(2, 1, 0).ravel())
        cr_l.append(cid(I[:, 1:], J[:, 1:], K[:, 1:]).transpose(2, 1, 0).ravel())
        av_l.append(av.transpose(0, 3, 2, 1).reshape(3, -1)); fc_l.append(ctr.transpose(0, 3, 2, 1).reshape(3, -1))
    if nz > 1:
        av, ctr = zfaces(slice(1, nz))
        cl_l.append(cid(I[:, :, :-1], J[:, :, :-1], K[:, :, :-1]).transpose(2, 1, 0).ravel())
        cr_l.append(cid(I[:, :, 1:], J[:, :, 1:], K[:, :, 1:]).transpose(2, 1, 0).ravel())
        av_l.append(av.transpose(0, 3, 2, 1).reshape(3, -1)); fc_l.append(ctr.transpose(0, 3, 2, 1).reshape(3, -1))
    n_int = sum(len(c) for c in cl_l)
    for _ in cl_l:
        bt_l.append(None)

    bfaces = []  # (side, cl, av, fc)
    av, ctr = xfaces(0); bfaces.append(("xmin", cid(0, J[0], K[0]).T.ravel(), -av.transpose(0, 2, 1).reshape(3, -1), ctr.transpose(0, 2, 1).reshape(3, -1)))
    av, ctr = xfaces(nx); bfaces.append(("xmax", cid(nx - 1, J[0], K[0]).T.ravel(), av.transpose(0, 2, 1).reshape(3, -1), ctr.transpose(0, 2, 1).reshape(3, -1)))
    av, ctr = yfaces(0); bfaces.append(("ymin", cid(I[:, 0], 0, K[:, 0]).T.ravel(), -av.transpose(0, 2, 1).reshape(3, -1), ctr.transpose(0, 2, 1).reshape(3, -1)))
    av, ctr = yfaces(ny); bfaces.append(("ymax", cid(I[:, 0], ny - 1, K[:, 0]).T.ravel(), av.transpose(0, 2, 1).reshape(3, -1), ctr.transpose(0, 2, 1).reshape(3, -1)))
    av, ctr = zfaces(0); bfaces.append(("zmin", cid(I[:, :, 0], J[:, :, 0], 0).T.ravel(), -av.transpose(0, 2, 1).reshape(3, -1), ctr.transpose(0, 2, 1).reshape(3, -1)))
    av, ctr = zfaces(nz); bfaces.append(("zmax", cid(I[:, :, 0], J[:, :, 0], nz - 1).T.ravel(), av.transpose(0, 2, 1).reshape(3, -1), ctr.transpose(0, 2, 1).reshape(3, -1)))

    cl = np.concatenate(cl_l + [b[1] for b in bfaces]).astype(np.int32)
    cr = np.concatenate(cr_l + [np.full(len(b[1]), -1) for b in bfaces]).astype(np.int32)
    avec = np.concatenate(av_l + [b[2] for b in bfaces], axis=1)
    fc = np.concatenate(fc_l + [b[3] for b in bfaces], axis=1)
    nf = len(cl)
    btype = np.zeros(nf, np.int32)
    pface = np.full(nf, -1, np.int32)
    pshift = np.zeros((3, nf))
    off = n_int
    side_range = {}
    for side, bcl, _, _ in bfaces:
        side_range[side] = (off, off + len(bcl))
        btype[off:off + len(bcl)] = bc[side]
        off += len(bcl)
    for a, b, axis in (("xmin", "xmax", 0), ("ymin", "ymax", 1), ("zmin", "zmax", 2)):
        if bc[a] == PERIODIC or bc[b] == PERIODIC:
            assert bc[a] == PERIODIC and bc[b] == PERIODIC
            ra, rb = side_range[a], side_range[b]
            fa = np.arange(*ra); fb = np.arange(*rb)
            pface[fa] = fb; pface[fb] = fa
            L = np.asarray(periodic_shift[axis], float)
            pshift[:, fa] = -L[:, None]
            pshift[:, fb] = L[:, None]
    area = np.sqrt((avec ** 2).sum(0))
    n = avec / area
    # volumes by the divergence theorem: V = 1/3 sum_f x_f . A_f (outward)
    vol = np.zeros(nc)
    xa = (fc * avec).sum(0)
    np.add.at(vol, cl, xa)
    inter = cr >= 0
    np.subtract.at(vol, cr[inter], xa[inter])
    # periodic/boundary faces are outward for cl; the divergence theorem needs
    # the face-centre of the face as seen from the cell, which it is.
    vol /= 3.0
    g = Grid(nc=nc, nf=nf, cl=cl, cr=cr, btype=btype, pface=pface, n=np.ascontiguousarray(n),
             area=area, fc=np.ascontiguousarray(fc), vol=vol, cc=np.ascontiguousarray(cc),
             pshift=np.ascontiguousarray(pshift), name=name,
             meta={"shape": (nx, ny, nz)})
    return g


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------

GAMMA = 1.4
R_AIR = 287.0


def box_grid(n=32, bc_type=FARFIELD):
    """Uniform n^3 hex box on the unit cube, one boundary type on all sides."""
    x = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"))
    bc = {s: bc_type for s in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    return structured_hex(X, bc, name=f"box{n}")


def box_state(g, mach=0.3, p=101325.0, T=300.0, amp=0.01, seed=1):
    """Uniform Mach `mach` x-flow plus seeded N(0, amp) relative noise on rho and p."""
    rng = np.random.default_rng(seed)
    rho0 = p / (R_AIR * T)
    a = np.sqrt(GAMMA * R_AIR * T)
    rho = rho0 * (1.0 + amp * rng.standard_normal(g.nc))
    pp = p * (1.0 + amp * rng.standard_normal(g.nc))
    q = np.zeros((5, g.nc))
    q[0] = pp / (rho * R_AIR)
    q[1] = mach * a
    q[4] = pp
    return q


def sutherland_mu(T, a1=1.458e-6, a2=1.5, a3=110.4):
    return a1 * T ** a2 / (T + a3)


def channel_grid(nx=128, ny=128, nz=128, ratio=1.1, re=500.0, mach=0.2, T=300.0, p=101325.0):
    """Periodic-x/z channel with viscous walls at y = +-h, geometric stretching
    `ratio` toward each wall; h is chosen so that Re = rho U_c h / mu = re."""
    a = np.sqrt(GAMMA * R_AIR * T)
    Uc = mach * a
    rho = p / (R_AIR * T)
    h = re * sutherland_mu(T) / (rho * Uc)
    xs = _axis_coords(nx, 0.0, 2 * np.pi * h)
    ys = _axis_coords(ny, -h, h, "two_sided_geometric", ratio)
    zs = _axis_coords(nz, 0.0, np.pi * h)
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"))
    bc = {"xmin": PERIODIC, "xmax": PERIODIC, "ymin": WALL, "ymax": WALL,
          "zmin": PERIODIC, "zmax": PERIODIC}
    shift = [(2 * np.pi * h, 0, 0), (0, 0, 0), (0, 0, np.pi * h)]
    g = structured_hex(X, bc, periodic_shift=shift, name=f"channel{nx}x{ny}x{nz}")
    g.meta.update(h=h, Uc=Uc)
    return g


def wall_distance(g: Grid) -> np.ndarray:
    """dist_noslip: projected distance |(x_f - x_c) . n_f| to the no-slip face
    whose centre is nearest the cell centre (src/solverTurbulence.loci:133-137,
    the default "projected" distance function)."""
    from scipy.spatial import cKDTree
    w = np.flatnonzero(g.btype == WALL)
    if len(w) == 0:
        return np.full(g.nc, 1e30)
    tree = cKDTree(g.fc[:, w].T)
    _, idx = tree.query(g.cc.T)
    f = w[idx]
    d = g.fc[:, f] - g.cc
    return np.abs(np.einsum("ij,ij->j", d, g.n[:, f]))


def channel_state(g, T=300.0, p=101325.0, amp=0.02, seed=2):
    """Poiseuille u = U_c (1 - y^2/h^2) plus seeded N(0, amp U_c) velocity noise."""
    rng = np.random.default_rng(seed)
    h, Uc = g.meta["h"], g.meta["Uc"]
    y = g.cc[1]
    q = np.zeros((5, g.nc))
    q[0] = T
    q[1] = Uc * (1.0 - (y / h) ** 2)
    q[1:4] += amp * Uc * rng.standard_normal((3, g.nc))
    q[4] = p
    return q
