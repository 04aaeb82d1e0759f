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
# boundary types with a parameter set of their own (include/fpsi.h)
REFLECTING, INFLOW, SUPERSONIC_INFLOW, OUTFLOW, SUPERSONIC_OUTFLOW, EXTRAPOLATE = 5, 6, 7, 8, 9, 10
ISENTROPIC_INFLOW, FIXED_MASS = 11, 12


def bc_code(btype, bc_set=0):
    """btype of a boundary face of type `btype` reading parameter set
    `bc_set` (bits 8-15)"""
    return int(btype) | (int(bc_set) << 8)


def bc_type(btype):
    return np.asarray(btype) & 0xFF
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


# ---------------------------------------------------------------------------
# general polyhedral grids (tetrahedra, prisms, hybrid cells)
# ---------------------------------------------------------------------------

def polyhedral_grid(points, tris, tri_cell, quads, quad_cell, bc_of_face, name=""):
    """Face grid from cells given as polygons.

    ``points[npts, 3]``; ``tris[m3, 3]`` / ``quads[m4, 4]`` node ids of each
    cell's faces (any orientation), ``tri_cell`` / ``quad_cell`` the cell each
    polygon belongs to.  A polygon listed by two cells is an interior face, by
    one cell a boundary face whose type is ``bc_of_face(centre[3, nb],
    normal[3, nb])``.  Face area vectors are the polygon area vectors (a quad
    is split along its first diagonal), face centres their area-weighted
    centroids; cell volumes follow from the divergence theorem and cell
    centres from the pyramids on each face.  Interior faces get ``cl < cr``
    and the normal from cl to cr.
    """
    P = np.asarray(points, float)
    nc = int(max(np.max(tri_cell, initial=-1), np.max(quad_cell, initial=-1))) + 1

    def tri_geom(a, b, c):
        av = 0.5 * np.cross(b - a, c - a)
        return av, (a + b + c) / 3.0

    polys = []
    for arr, cells in ((np.asarray(tris, np.int64).reshape(-1, 3), np.asarray(tri_cell, np.int64)),
                       (np.asarray(quads, np.int64).reshape(-1, 4), np.asarray(quad_cell, np.int64))):
        if len(arr) == 0:
            continue
        key = np.sort(arr, axis=1)
        _, first, inv, cnt = np.unique(key, axis=0, return_index=True, return_inverse=True, return_counts=True)
        inv = inv.ravel()
        assert cnt.max() <= 2, "a face is shared by more than two cells"
        nu = len(first)
        # geometry from the first occurrence of each face
        nodes = arr[first]
        if arr.shape[1] == 3:
            av, fc = tri_geom(P[nodes[:, 0]], P[nodes[:, 1]], P[nodes[:, 2]])
        else:
            a1, c1 = tri_geom(P[nodes[:, 0]], P[nodes[:, 1]], P[nodes[:, 2]])
            a2, c2 = tri_geom(P[nodes[:, 0]], P[nodes[:, 2]], P[nodes[:, 3]])
            av = a1 + a2
            w1 = np.linalg.norm(a1, axis=1)[:, None]
            w2 = np.linalg.norm(a2, axis=1)[:, None]
            fc = (w1 * c1 + w2 * c2) / (w1 + w2)
        owners = np.full((nu, 2), -1, np.int64)
        order = np.argsort(inv, kind="stable")
        sinv = inv[order]
        starts = np.searchsorted(sinv, np.arange(nu))
        owners[:, 0] = cells[order[starts]]
        two = cnt == 2
        owners[two, 1] = cells[order[starts[two] + 1]]
        polys.append((av, fc, owners))
    av = np.concatenate([p[0] for p in polys])
    fc = np.concatenate([p[1] for p in polys])
    own = np.concatenate([p[2] for p in polys])
    # cell centre estimate: mean of the face centres of the cell
    est = np.zeros((nc, 3))
    cnt = np.zeros(nc)
    for col in (0, 1):
        m = own[:, col] >= 0
        np.add.at(est, own[m, col], fc[m])
        np.add.at(cnt, own[m, col], 1.0)
    est /= cnt[:, None]
    inter = own[:, 1] >= 0
    cl = np.where(inter, np.minimum(own[:, 0], own[:, 1]), own[:, 0])
    cr = np.where(inter, np.maximum(own[:, 0], own[:, 1]), -1)
    # orient: outward from cl
    sgn = np.sign(np.einsum("ij,ij->i", av, fc - est[cl]))
    sgn[sgn == 0] = 1.0
    av = av * sgn[:, None]
    # volumes and centroids (pyramids on each face from the estimate)
    vol = np.zeros(nc)
    cen = np.zeros((nc, 3))
    for side, cells in ((1.0, cl), (-1.0, np.where(inter, cr, -1))):
        m = cells >= 0
        c = cells[m]
        vf = side * np.einsum("ij,ij->i", av[m], fc[m] - est[c]) / 3.0
        np.add.at(vol, c, vf)
        np.add.at(cen, c, vf[:, None] * (est[c] + 0.75 * (fc[m] - est[c])))
    assert np.all(vol > 0), "non-positive cell volume"
    cen /= vol[:, None]
    # faces: interior first (by cl, cr), then boundary
    idx_i = np.flatnonzero(inter)
    idx_i = idx_i[np.lexsort((cr[idx_i], cl[idx_i]))]
    idx_b = np.flatnonzero(~inter)
    idx_b = idx_b[np.argsort(cl[idx_b], kind="stable")]
    perm = np.concatenate([idx_i, idx_b])
    av, fc, cl, cr = av[perm], fc[perm], cl[perm], cr[perm]
    area = np.linalg.norm(av, axis=1)
    n = (av / area[:, None]).T
    nf = len(cl)
    btype = np.zeros(nf, np.int32)
    nb = len(idx_b)
    btype[len(idx_i):] = np.asarray(bc_of_face(fc[len(idx_i):].T, n[:, len(idx_i):]), np.int32)
    g = Grid(nc=nc, nf=nf, cl=cl.astype(np.int32), cr=cr.astype(np.int32), btype=btype,
             pface=np.full(nf, -1, np.int32), n=np.ascontiguousarray(n), area=area,
             fc=np.ascontiguousarray(fc.T), vol=vol, cc=np.ascontiguousarray(cen.T),
             pshift=np.zeros((3, nf)), name=name, meta={})
    return g


def _side_bc(lo, hi, walls, tol=1e-9):
    """Boundary types of a box [lo, hi]^3: WALL on the sides named in
    `walls` ("ymin", ...), FARFIELD elsewhere."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)

    def bc(fc, n):
        t = np.full(fc.shape[1], FARFIELD, np.int32)
        for s in walls:
            ax = "xyz".index(s[0])
            at = lo[ax] if s.endswith("min") else hi[ax]
            t[np.abs(fc[ax] - at) < tol * max(1.0, abs(at) + 1.0)] = WALL
        return t
    return bc


def _jitter(X, amp, rng):
    """Move interior lattice nodes by U(-amp, amp) x the local spacing per axis."""
    Y = X.copy()
    for ax in range(3):
        d = np.diff(X[ax], axis=ax)
        h = np.minimum(np.take(d, range(0, d.shape[ax] - 1), axis=ax), np.take(d, range(1, d.shape[ax]), axis=ax))
        sl = [slice(1, -1)] * 3
        Y[ax][tuple(sl)] += amp * rng.uniform(-1.0, 1.0, Y[ax][tuple(sl)].shape) * h[tuple(
            slice(None) if a == ax else slice(1, -1) for a in range(3))]
    return Y


def tet_cube(n=8, jitter=0.1, seed=3, walls=("ymin",)):
    """n^3 hexes on the unit cube, each cut into 6 tetrahedra along its main
    diagonal (conforming across hexes), interior nodes jittered by
    U(-jitter, jitter) x h per coordinate."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(x, x, x, indexing="ij"))
    X = _jitter(X, jitter, rng)
    P = X.reshape(3, -1).T
    N = n + 1

    def nid(i, j, k):
        return (i * N + j) * N + k
    I, J, K = [a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")]
    v = {(a, b, c): nid(I + a, J + b, K + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)}
    paths = [((1, 0, 0), (1, 1, 0)), ((1, 0, 0), (1, 0, 1)), ((0, 1, 0), (1, 1, 0)),
             ((0, 1, 0), (0, 1, 1)), ((0, 0, 1), (1, 0, 1)), ((0, 0, 1), (0, 1, 1))]
    tets = []
    for p1, p2 in paths:
        tets.append(np.stack([v[(0, 0, 0)], v[p1], v[p2], v[(1, 1, 1)]], axis=1))
    T = np.concatenate(tets)                      # [6 n^3, 4]
    nct = len(T)
    faces = np.concatenate([T[:, [1, 2, 3]], T[:, [0, 2, 3]], T[:, [0, 1, 3]], T[:, [0, 1, 2]]])
    fcell = np.tile(np.arange(nct), 4)
    g = polyhedral_grid(P, faces, fcell, np.zeros((0, 4), np.int64), np.zeros(0, np.int64),
                        _side_bc((0, 0, 0), (1, 1, 1), walls), name=f"tets{n}")
    g.meta.update(kind="tets", n=n)
    return g


def prism_hex_box(nx=8, nz=8, n_prism=6, n_hex=6, first=1e-3, growth=1.2, H=1.0, jitter=0.1, seed=4):
    """Wall at y = 0 under `n_prism` layers of prisms (each x-z quad cut into
    two triangles, first layer height `first`, geometric `growth`) and a core
    of `n_hex` hex layers up to y = H with interior nodes jittered by
    U(-jitter, jitter) x h.  The first core layer sits on the prisms'
    triangles, so its cells are 7-faced polyhedra."""
    rng = np.random.default_rng(seed)
    hs = first * growth ** np.arange(n_prism)
    yp = np.concatenate([[0.0], np.cumsum(hs)])
    ytop = yp[-1]
    assert ytop < H
    yh = np.linspace(ytop, H, n_hex + 1)
    ys = np.concatenate([yp, yh[1:]])
    xs = np.linspace(0.0, 1.0, nx + 1)
    zs = np.linspace(0.0, 1.0, nz + 1)
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"))
    # jitter x and z of core nodes strictly inside, and y of core nodes above the prism top
    Y = X.copy()
    ny = len(ys) - 1
    core = slice(n_prism + 1, ny)
    hx, hz = 1.0 / nx, 1.0 / nz
    hy = (H - ytop) / n_hex
    Y[0][1:-1, core, 1:-1] += jitter * hx * rng.uniform(-1, 1, Y[0][1:-1, core, 1:-1].shape)
    Y[2][1:-1, core, 1:-1] += jitter * hz * rng.uniform(-1, 1, Y[2][1:-1, core, 1:-1].shape)
    Y[1][1:-1, core, 1:-1] += jitter * hy * rng.uniform(-1, 1, Y[1][1:-1, core, 1:-1].shape)
    P = Y.reshape(3, -1).T
    NY, NZ = ny + 1, nz + 1

    def nid(i, j, k):
        return (i * NY + j) * NZ + k
    tris, tcell, quads, qcell = [], [], [], []
    cid = 0
    I, K = [a.ravel() for a in np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")]
    m = len(I)
    for j in range(ny):
        a0, b0, c0, d0 = nid(I, j, K), nid(I + 1, j, K), nid(I + 1, j, K + 1), nid(I, j, K + 1)
        a1, b1, c1, d1 = nid(I, j + 1, K), nid(I + 1, j + 1, K), nid(I + 1, j + 1, K + 1), nid(I, j + 1, K + 1)
        if j < n_prism:   # two prisms per quad: (a, b, c) and (a, c, d)
            for (p, q_, r), (p1, q1, r1) in (((a0, b0, c0), (a1, b1, c1)), ((a0, c0, d0), (a1, c1, d1))):
                cells = cid + np.arange(m)
                cid += m
                tris += [np.stack([p, q_, r], 1), np.stack([p1, q1, r1], 1)]
                tcell += [cells, cells]
                quads += [np.stack([p, q_, q1, p1], 1), np.stack([q_, r, r1, q1], 1), np.stack([r, p, p1, r1], 1)]
                qcell += [cells, cells, cells]
        else:             # hex; on the prism top its bottom is the two triangles
            cells = cid + np.arange(m)
            cid += m
            if j == n_prism:
                tris += [np.stack([a0, b0, c0], 1), np.stack([a0, c0, d0], 1)]
                tcell += [cells, cells]
            else:
                quads.append(np.stack([a0, b0, c0, d0], 1))
                qcell.append(cells)
            quads += [np.stack([a1, b1, c1, d1], 1), np.stack([a0, b0, b1, a1], 1), np.stack([b0, c0, c1, b1], 1),
                      np.stack([c0, d0, d1, c1], 1), np.stack([d0, a0, a1, d1], 1)]
            qcell += [cells] * 5
    g = polyhedral_grid(P, np.concatenate(tris), np.concatenate(tcell), np.concatenate(quads),
                        np.concatenate(qcell), _side_bc((0, 0, 0), (1, H, 1), ("ymin",)),
                        name=f"prismhex{nx}x{n_prism}+{n_hex}x{nz}")
    g.meta.update(kind="prism_hex", first=first)
    return g


def bl_box(nx=8, ny=12, nz=8, aspect=1e3, L=1.0, seed=5, side=None):
    """Boundary-layer hex grid: wall at y = 0, geometric stretching in y from a
    first cell of height L/nx / aspect, farfield elsewhere; `side` sets the
    type of the spanwise planes z = 0, L (default farfield)."""
    dx = L / nx
    h1 = dx / aspect
    # growth r with h1 (r^ny - 1)/(r - 1) = L
    lo, hi = 1.0 + 1e-9, 3.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        tot = h1 * (r ** ny - 1.0) / (r - 1.0)
        lo, hi = (r, hi) if tot < L else (lo, r)
    ys = np.concatenate([[0.0], np.cumsum(h1 * r ** np.arange(ny))])
    ys *= L / ys[-1]
    xs = np.linspace(0.0, L, nx + 1)
    zs = np.linspace(0.0, L, nz + 1)
    X = np.stack(np.meshgrid(xs, ys, zs, indexing="ij"))
    zs_bc = FARFIELD if side is None else side
    bc = {"xmin": FARFIELD, "xmax": FARFIELD, "ymin": WALL, "ymax": FARFIELD, "zmin": zs_bc, "zmax": zs_bc}
    g = structured_hex(X, bc, name=f"bl{nx}x{ny}x{nz}")
    g.meta.update(kind="bl", aspect=aspect, growth=r)
    return g


def freestream_state(g, mach, T=300.0, p=101325.0, amp=0.01, seed=3):
    """Uniform Mach `mach` x-flow plus seeded N(0, amp) relative noise on rho
    and p; velocity zero at cells next to a wall ramps in linearly over the
    wall distance d < 0.05 (keeps the start near a no-slip state)."""
    rng = np.random.default_rng(seed)
    a = np.sqrt(GAMMA * R_AIR * T)
    rho0 = p / (R_AIR * T)
    rho = rho0 * (1.0 + amp * rng.standard_normal(g.nc))
    pp = p * (1.0 + amp * rng.standard_normal(g.nc))
    q = np.zeros((5, g.nc))
    q[0] = pp / (rho * R_AIR)
    q[1] = mach * a
    q[4] = pp
    d = g.meta.get("dist")
    if d is not None:
        q[1] *= np.minimum(1.0, d / 0.05)
    return q
