"""A reader for the subset of HDF5 that Loci's grid files use.

Written from the published HDF5 file-format specification (version 0/1
superblocks, version 1 object headers with continuation blocks, symbol-table
groups held in version 1 B-trees, symbol-table nodes and local heaps;
dataspace, datatype, data-layout, filter-pipeline and attribute messages;
contiguous, compact and chunked storage with deflate).  Fixed-point,
floating-point, fixed-length string and compound datatypes are decoded to
numpy; anything else raises.

    f = File(path)
    f.keys("/")                  -> names in the root group
    f["face_info/cluster_sizes"] -> numpy array
    f.attrs("file_info")         -> {name: value}
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


class HDF5Error(ValueError):
    pass


class _Obj:
    """Parsed object header: messages as (type, bytes)."""

    def __init__(self, msgs):
        self.msgs = msgs

    def first(self, t):
        for mt, data in self.msgs:
            if mt == t:
                return data
        return None

    def all(self, t):
        return [d for mt, d in self.msgs if mt == t]


class File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.b = fh.read()
        b = self.b
        if b[:8] != b"\x89HDF\r\n\x1a\n":
            raise HDF5Error(f"{path}: not an HDF5 file")
        ver = b[8]
        if ver not in (0, 1):
            raise HDF5Error(f"superblock version {ver} not supported")
        self.so, self.sl = b[13], b[14]
        if self.so != 8 or self.sl != 8:
            raise HDF5Error("only 8-byte offsets and lengths are supported")
        p = 24 + (4 if ver == 1 else 0)
        self.base = self._u(p, 8)
        p += 32  # base, free space, end of file, driver info
        # root group symbol table entry
        self.root = self._u(p + 8, 8)
        self._cache = {}

    # -- primitive reads
    def _u(self, p, n):
        return int.from_bytes(self.b[p:p + n], "little")

    def _addr(self, a):
        return self.base + a

    # -- object headers
    def _header(self, addr):
        if addr in self._cache:
            return self._cache[addr]
        b = self.b
        p = self._addr(addr)
        if b[p] != 1:
            raise HDF5Error(f"object header version {b[p]} at {addr:#x} not supported")
        nmsg = self._u(p + 2, 2)
        size = self._u(p + 8, 4)
        blocks = [(p + 16, size)]
        msgs = []
        while blocks and len(msgs) < nmsg:
            q, n = blocks.pop(0)
            end = q + n
            while q + 8 <= end and len(msgs) < nmsg:
                mt, ms = self._u(q, 2), self._u(q + 2, 2)
                data = b[q + 8:q + 8 + ms]
                if mt == 0x10:  # continuation
                    blocks.append((self._addr(int.from_bytes(data[:8], "little")), int.from_bytes(data[8:16], "little")))
                msgs.append((mt, data))
                q += 8 + ms
        obj = _Obj(msgs)
        self._cache[addr] = obj
        return obj

    # -- groups
    def _heap_name(self, heap, off):
        b = self.b
        p = self._addr(heap)
        if b[p:p + 4] != b"HEAP":
            raise HDF5Error("bad local heap")
        data = self._addr(self._u(p + 24, 8))
        e = b.index(b"\0", data + off)
        return b[data + off:e].decode()

    def _group_entries(self, addr):
        st = self._header(addr).first(0x11)
        if st is None:
            raise HDF5Error("not a group (only symbol-table groups are supported)")
        btree, heap = struct.unpack("<QQ", st[:16])
        out = {}

        def walk(node):
            p = self._addr(node)
            b = self.b
            if b[p:p + 4] != b"TREE":
                raise HDF5Error("bad B-tree node")
            level, used = b[p + 5], self._u(p + 6, 2)
            q = p + 24
            for i in range(used):
                child = self._u(q + 8, 8)  # key (length), child (offset)
                q += 16
                if level > 0:
                    walk(child)
                else:
                    s = self._addr(child)
                    if b[s:s + 4] != b"SNOD":
                        raise HDF5Error("bad symbol table node")
                    n = self._u(s + 6, 2)
                    e = s + 8
                    for _ in range(n):
                        name = self._heap_name(heap, self._u(e, 8))
                        out[name] = self._u(e + 8, 8)
                        e += 40

        walk(btree)
        return out

    def _resolve(self, path):
        addr = self.root
        for part in [x for x in path.split("/") if x]:
            ents = self._group_entries(addr)
            if part not in ents:
                raise KeyError(path)
            addr = ents[part]
        return addr

    def keys(self, path="/"):
        return list(self._group_entries(self._resolve(path)))

    def is_group(self, path):
        return self._header(self._resolve(path)).first(0x11) is not None

    # -- datatypes and dataspaces
    @classmethod
    def _dtype(cls_, t):
        return cls_._dtype_at(t, 0)[0]

    @classmethod
    def _dtype_at(cls_, t, p):
        """(numpy dtype, bytes consumed) of the datatype message at t[p:]."""
        cls, ver = t[p] & 0x0F, t[p] >> 4
        bits = t[p + 1] | (t[p + 2] << 8) | (t[p + 3] << 16)
        size = int.from_bytes(t[p + 4:p + 8], "little")
        order = ">" if bits & 1 else "<"
        if cls == 0:
            signed = bool(bits & 0x08)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}"), 12
        if cls == 1:
            return np.dtype(f"{order}f{size}"), 20
        if cls == 3:
            return np.dtype(f"S{size}"), 8
        if cls == 6 and ver in (1, 3):
            nmem = bits & 0xFFFF
            q = p + 8
            names, formats, offsets = [], [], []
            for _ in range(nmem):
                e = t.index(b"\0", q)
                names.append(t[q:e].decode())
                if ver == 1:
                    q += ((e - q) // 8 + 1) * 8
                    off = int.from_bytes(t[q:q + 4], "little")
                    nd = t[q + 4]
                    dims = [int.from_bytes(t[q + 16 + 4 * i:q + 20 + 4 * i], "little") for i in range(nd)]
                    q += 32
                else:
                    q = e + 1
                    nb = 1 if size < 256 else 2 if size < 65536 else 3 if size < (1 << 24) else 4
                    off = int.from_bytes(t[q:q + nb], "little")
                    q += nb
                    dims = []
                mdt, used = cls_._dtype_at(t, q)
                q += used
                formats.append((mdt, tuple(dims)) if dims else mdt)
                offsets.append(off)
            return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size}), q - p
        raise HDF5Error(f"datatype class {cls} (version {ver}) not supported")

    @staticmethod
    def _dims(s):
        ver, nd, flags = s[0], s[1], s[2]
        p = 8 if ver == 1 else 4
        if ver == 2 and s[3] == 2:  # null dataspace
            return None
        dims = [int.from_bytes(s[p + 8 * i:p + 8 * i + 8], "little") for i in range(nd)]
        return tuple(dims)

    def _read_data(self, obj, dt, dims):
        lay = obj.first(0x08)
        n = int(np.prod(dims)) if dims else 1
        nbytes = n * dt.itemsize
        filters = obj.first(0x0B)
        ver = lay[0]
        if ver in (1, 2):
            nd, cls = lay[1], lay[2]
            p = 8
            if cls == 0:  # compact
                p += 4 * nd
                size = int.from_bytes(lay[p:p + 4], "little")
                raw = lay[p + 4:p + 4 + size]
            else:
                addr = int.from_bytes(lay[p:p + 8], "little")
                p += 8
                cdims = [int.from_bytes(lay[p + 4 * i:p + 4 * i + 4], "little") for i in range(nd)]
                if cls == 1:
                    raw = self._contig(addr, nbytes)
                else:
                    raw = self._chunked(addr, cdims[:-1], dims, dt, filters)
        elif ver == 3:
            cls = lay[1]
            if cls == 0:
                size = int.from_bytes(lay[2:4], "little")
                raw = lay[4:4 + size]
            elif cls == 1:
                addr = int.from_bytes(lay[2:10], "little")
                raw = self._contig(addr, nbytes)
            elif cls == 2:
                nd = lay[2]
                addr = int.from_bytes(lay[3:11], "little")
                cdims = [int.from_bytes(lay[11 + 4 * i:15 + 4 * i], "little") for i in range(nd)]
                raw = self._chunked(addr, cdims[:-1], dims, dt, filters)
            else:
                raise HDF5Error(f"layout class {cls} not supported")
        else:
            raise HDF5Error(f"layout version {ver} not supported")
        a = np.frombuffer(raw[:nbytes], dt, count=n).copy()
        return a.reshape(dims) if dims else a[0]

    def _contig(self, addr, nbytes):
        if addr == UNDEF:
            return b"\0" * nbytes
        p = self._addr(addr)
        return self.b[p:p + nbytes]

    def _chunked(self, btree, cdims, dims, dt, filters):
        deflate = False
        if filters is not None:
            ver, nf = filters[0], filters[1]
            q = 8 if ver == 1 else 2
            for _ in range(nf):
                fid = int.from_bytes(filters[q:q + 2], "little")
                if ver == 1:
                    nl = int.from_bytes(filters[q + 2:q + 4], "little")
                    nv = int.from_bytes(filters[q + 6:q + 8], "little")
                    q += 8 + ((nl + 7) // 8) * 8 + 4 * nv + (4 if nv % 2 else 0)
                else:
                    nv = int.from_bytes(filters[q + 4:q + 6], "little")
                    q += 6 + 4 * nv
                if fid == 1:
                    deflate = True
                else:
                    raise HDF5Error(f"filter {fid} not supported")
        out = np.zeros(dims, dt)
        nd = len(dims)

        def walk(node):
            p = self._addr(node)
            b = self.b
            if b[p:p + 4] != b"TREE" or b[p + 4] != 1:
                raise HDF5Error("bad chunk B-tree node")
            level, used = b[p + 5], self._u(p + 6, 2)
            ksize = 8 + 8 * (nd + 1)
            q = p + 24
            for i in range(used):
                csize = self._u(q, 4)
                offs = [self._u(q + 8 + 8 * d, 8) for d in range(nd)]
                child = self._u(q + ksize, 8)
                q += ksize + 8
                if level > 0:
                    walk(child)
                    continue
                raw = self.b[self._addr(child):self._addr(child) + csize]
                if deflate:
                    raw = zlib.decompress(raw)
                chunk = np.frombuffer(raw, dt, count=int(np.prod(cdims))).reshape(cdims)
                sl = tuple(slice(o, min(o + c, n)) for o, c, n in zip(offs, cdims, dims))
                out[sl] = chunk[tuple(slice(0, s.stop - s.start) for s in sl)]

        walk(btree)
        return out.tobytes()

    def __getitem__(self, path):
        obj = self._header(self._resolve(path))
        dt = self._dtype(obj.first(0x03))
        dims = self._dims(obj.first(0x01))
        return self._read_data(obj, dt, dims)

    def attrs(self, path="/"):
        obj = self._header(self._resolve(path))
        out = {}
        for a in obj.all(0x0C):
            ver = a[0]
            if ver != 1:
                raise HDF5Error(f"attribute message version {ver} not supported")
            nlen, tlen, slen = struct.unpack("<HHH", a[2:8])
            p = 8
            name = a[p:p + nlen].split(b"\0")[0].decode()
            p += ((nlen + 7) // 8) * 8
            dt = self._dtype(a[p:p + tlen])
            p += ((tlen + 7) // 8) * 8
            dims = self._dims(a[p:p + slen])
            p += ((slen + 7) // 8) * 8
            n = int(np.prod(dims)) if dims else 1
            v = np.frombuffer(a[p:p + n * dt.itemsize], dt, count=n)
            out[name] = v.reshape(dims) if dims else v[0]
        return out
