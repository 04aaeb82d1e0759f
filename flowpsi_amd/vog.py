"""Loci VOG grid files: reader and conversion to the solver's face grid.

A VOG file (the grid format flowPsi reads through Loci, quickTest/*.vog) is
HDF5 with
  file_info            attributes numNodes, numFaces, numCells
  node_info/positions  compound (x, y, z) float64 per node
  face_info/cluster_sizes, cluster_info
                       the faces in clusters of bytes (below)
  surface_info/<name>  one group per boundary surface, attribute Ident
Faces are stored in clusters.  A cluster is
  repeated { nsz, count, count x (nsz local node ids, local cl, local cr) },
  a 0 byte,
  the node table, then the cell table.
A table is its size (a byte, 0 meaning 256), the first entry as a signed
varint (6 value bits, 0x40 sign, 0x80 continuation) and the remaining
entries as unsigned varint increments (7 value bits, 0x80 continuation).
Node and cell ids are 0-based; a boundary face has cr = -Ident of its
surface; the node order of a face gives the normal pointing from cl to cr.
This layout was recovered from the files themselves: the reference ships
the grids but not Loci's reader.  Every decoded grid is checked for closed
cells and positive volumes.
"""
from __future__ import annotations

import numpy as np

from . import grids, hdf5


class VogError(ValueError):
    pass


def _signed(c, p):
    b = int(c[p]); p += 1
    neg = b & 0x40
    v = b & 0x3F
    shift = 6
    while b & 0x80:
        b = int(c[p]); p += 1
        v += (b & 0x7F) << shift
        shift += 7
    return (-v if neg else v), p


def _unsigned(c, p):
    b = int(c[p]); p += 1
    v = b & 0x7F
    shift = 7
    while b & 0x80:
        b = int(c[p]); p += 1
        v += (b & 0x7F) << shift
        shift += 7
    return v, p


def _table(c, p):
    n = int(c[p]) or 256
    p += 1
    out = np.empty(n, np.int64)
    out[0], p = _signed(c, p)
    for i in range(1, n):
        d, p = _unsigned(c, p)
        out[i] = out[i - 1] + d
    return out, p


def decode_cluster(c):
    """(face sizes, local node ids, local cl, local cr, node table, cell
    table, bytes used) of one cluster."""
    c = np.asarray(c, np.uint8)
    p = 0
    sizes, nodes, lcl, lcr = [], [], [], []
    while c[p] != 0:
        nsz, cnt = int(c[p]), int(c[p + 1])
        p += 2
        rec = c[p:p + cnt * (nsz + 2)].reshape(cnt, nsz + 2)
        p += cnt * (nsz + 2)
        sizes.append(np.full(cnt, nsz, np.int64))
        nodes.append(rec[:, :nsz].ravel())
        lcl.append(rec[:, nsz])
        lcr.append(rec[:, nsz + 1])
    p += 1
    ntab, p = _table(c, p)
    ctab, p = _table(c, p)
    return (np.concatenate(sizes), np.concatenate(nodes).astype(np.int64), np.concatenate(lcl).astype(np.int64),
            np.concatenate(lcr).astype(np.int64), ntab, ctab, p)


def read(path):
    """The raw contents of a VOG file: dict with pos [nn, 3], face_off
    [nf + 1], face_nodes, cl, cr [nf] (cr < 0: -Ident of a boundary
    surface), surfaces {name: Ident}, nc."""
    f = hdf5.File(path)
    info = f.attrs("file_info")
    nn, nf, nc = (int(np.asarray(info[k]).ravel()[0]) for k in ("numNodes", "numFaces", "numCells"))
    P = f["node_info/positions"]
    pos = np.stack([P["x"], P["y"], P["z"]], axis=1).astype(np.float64)
    if pos.shape[0] != nn:
        raise VogError("position count differs from numNodes")
    sizes = np.asarray(f["face_info/cluster_sizes"], np.int64)
    info_b = np.asarray(f["face_info/cluster_info"], np.uint8)
    off = 0
    fs, fn, cl, cr = [], [], [], []
    for s in sizes:
        c = info_b[off:off + s]
        sz, ln, lcl, lcr, ntab, ctab, used = decode_cluster(c)
        if used != s:
            raise VogError(f"cluster at byte {off}: decoded {used} of {s} bytes")
        fs.append(sz)
        fn.append(ntab[ln])
        cl.append(ctab[lcl])
        cr.append(ctab[lcr])
        off += s
    if off != len(info_b):
        raise VogError("cluster sizes do not cover cluster_info")
    fsz = np.concatenate(fs)
    if len(fsz) != nf:
        raise VogError(f"decoded {len(fsz)} faces, file says {nf}")
    face_off = np.concatenate([[0], np.cumsum(fsz)])
    face_nodes = np.concatenate(fn)
    cl = np.concatenate(cl)
    cr = np.concatenate(cr)
    if face_nodes.min() < 0 or face_nodes.max() >= nn or cl.min() < 0 or cl.max() >= nc or cr.max() >= nc:
        raise VogError("decoded ids out of range")
    surfaces = {}
    if "surface_info" in f.keys("/"):
        for name in f.keys("surface_info"):
            a = f.attrs("surface_info/" + name)
            surfaces[name] = int(np.asarray(a["Ident"]).ravel()[0])
    for t in np.unique(-cr[cr < 0]):
        if int(t) not in surfaces.values():
            surfaces[f"BC_{int(t)}"] = int(t)
    return dict(pos=pos, face_off=face_off, face_nodes=face_nodes, cl=cl, cr=cr, surfaces=surfaces, nc=nc)


def _face_geometry(pos, face_off, face_nodes):
    """Area vectors and centroids of polygons: triangles fanned from the
    node mean, centroid weighted by triangle area along the face normal."""
    nf = len(face_off) - 1
    fsz = np.diff(face_off)
    fid = np.repeat(np.arange(nf), fsz)
    X = pos[face_nodes]
    xm = np.zeros((nf, 3))
    np.add.at(xm, fid, X)
    xm /= fsz[:, None]
    nxt = np.arange(len(face_nodes)) + 1
    last = face_off[1:] - 1
    nxt[last] = face_off[:-1]
    a = X - xm[fid]
    b = X[nxt] - xm[fid]
    tri = 0.5 * np.cross(a, b)
    av = np.zeros((nf, 3))
    np.add.at(av, fid, tri)
    nrm = av / np.linalg.norm(av, axis=1)[:, None]
    w = np.einsum("ij,ij->i", tri, nrm[fid])
    tc = (X + X[nxt] + xm[fid]) / 3.0
    fc = np.zeros((nf, 3))
    np.add.at(fc, fid, w[:, None] * tc)
    wsum = np.zeros(nf)
    np.add.at(wsum, fid, w)
    fc /= wsum[:, None]
    return av, fc


def to_grid(v, bcs: dict, periodic: dict | None = None, name="") -> grids.Grid:
    """The solver's face grid of a read() result.

    bcs maps each surface name to a grids type (FARFIELD, WALL, SYMMETRY,
    PERIODIC); periodic maps a PERIODIC surface to its partner surface (one
    entry per pair).  Interior faces are oriented cl < cr (cell order is the
    file's), interior faces come first sorted by (cl, cr), then boundary
    faces sorted by cl.  Cell centres are volume centroids from pyramids on
    the face centroids."""
    pos, cl, cr, nc = v["pos"], v["cl"].copy(), v["cr"].copy(), v["nc"]
    av, fc = _face_geometry(pos, v["face_off"], v["face_nodes"])
    inter = cr >= 0
    flip = inter & (cl > cr)
    cl[flip], cr[flip] = cr[flip], cl[flip]
    av[flip] *= -1.0
    # cell centre estimate: mean of the face centres
    est = np.zeros((nc, 3))
    cnt = np.zeros(nc)
    np.add.at(est, cl, fc)
    np.add.at(cnt, cl, 1.0)
    np.add.at(est, cr[inter], fc[inter])
    np.add.at(cnt, cr[inter], 1.0)
    est /= cnt[:, None]
    vol = np.zeros(nc)
    cen = np.zeros((nc, 3))
    for side, cells, m in ((1.0, cl, np.ones(len(cl), bool)), (-1.0, cr, inter)):
        c = cells[m]
        vf = side * np.einsum("ij,ij->i", av[m], fc[m] - est[c]) / 3.0
        np.add.at(vol, c, vf)
        np.add.at(cen, c, vf[:, None] * (est[c] + 0.75 * (fc[m] - est[c])))
    if not np.all(vol > 0):
        raise VogError("non-positive cell volume: face orientation or decoding is wrong")
    cen /= vol[:, None]
    idx_i = np.flatnonzero(inter)
    idx_i = idx_i[np.lexsort((cr[idx_i], cl[idx_i]))]
    idx_b = np.flatnonzero(~inter)
    idx_b = idx_b[np.argsort(cl[idx_b], kind="stable")]
    perm = np.concatenate([idx_i, idx_b])
    av, fc, cl, tag = av[perm], fc[perm], cl[perm], -cr[perm]
    nf = len(cl)
    ni = len(idx_i)
    cr_out = np.full(nf, -1, np.int64)
    cr_out[:ni] = cr[perm][:ni]
    area = np.linalg.norm(av, axis=1)
    n = (av / area[:, None]).T
    btype = np.zeros(nf, np.int32)
    pface = np.full(nf, -1, np.int32)
    pshift = np.zeros((3, nf))
    ident = v["surfaces"]
    for sname, t in ident.items():
        m = np.flatnonzero(tag[ni:] == t) + ni
        if len(m) == 0:
            continue
        if sname not in bcs:
            raise VogError(f"no boundary condition for surface {sname}")
        btype[m] = bcs[sname]
    if np.any(btype[ni:] == grids.INTERIOR):
        raise VogError("boundary faces without a boundary condition")
    for a, b in (periodic or {}).items():
        fa = np.flatnonzero(tag[ni:] == ident[a]) + ni
        fb = np.flatnonzero(tag[ni:] == ident[b]) + ni
        if len(fa) != len(fb):
            raise VogError(f"periodic surfaces {a} and {b} differ in face count")
        # pair by centre after removing the mean offset between the surfaces
        shift = fc[fa].mean(0) - fc[fb].mean(0)
        d = np.linalg.norm((fc[fa][:, None, :] - shift) - fc[fb][None, :, :], axis=2)
        j = d.argmin(axis=1)
        if len(set(j.tolist())) != len(fa) or d[np.arange(len(fa)), j].max() > 1e-6 * np.sqrt(area.max()):
            raise VogError(f"periodic surfaces {a} and {b} do not match")
        pface[fa] = fb[j]
        pface[fb[j]] = fa
        pshift[:, fa] = (fc[fa] - fc[fb[j]]).T
        pshift[:, fb[j]] = (fc[fb[j]] - fc[fa]).T
    g = grids.Grid(nc=nc, nf=nf, cl=cl.astype(np.int32), cr=cr_out.astype(np.int32), btype=btype, pface=pface,
                   n=np.ascontiguousarray(n), area=area, fc=np.ascontiguousarray(fc.T), vol=vol,
                   cc=np.ascontiguousarray(cen.T), pshift=pshift, name=name,
                   meta={"surfaces": dict(ident), "face_tag": tag.astype(np.int32)})
    return g


def load(path, bcs: dict, periodic: dict | None = None, name=None) -> grids.Grid:
    import os
    return to_grid(read(path), bcs, periodic, name or os.path.basename(path))
