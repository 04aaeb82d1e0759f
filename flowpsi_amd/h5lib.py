"""A thin ctypes binding of the HDF5 C library (libhdf5, 1.10 API) for
writing and reading the restart files.  The library is looked up as
libhdf5.so on the loader path, then under $CONDA_PREFIX/lib and /opt/conda/lib;
without it, writing raises (reading grids needs only flowpsi_amd/hdf5.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

hid_t = C.c_int64
hsize_t = C.c_uint64
H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5T_COMPOUND = 6

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    cands = ["libhdf5.so", "libhdf5.so.103"]
    for d in (os.environ.get("CONDA_PREFIX"), "/opt/conda"):
        if d:
            cands += [os.path.join(d, "lib", "libhdf5.so"), os.path.join(d, "lib", "libhdf5.so.103")]
    err = None
    for c in cands:
        try:
            L = C.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"libhdf5 not found ({err})")
    for name in ("H5Fcreate", "H5Fopen", "H5Gcreate2", "H5Gopen2", "H5Screate_simple", "H5Dcreate2", "H5Dopen2",
                 "H5Dget_space", "H5Tcreate", "H5Acreate2", "H5Dget_type", "H5Tget_size"):
        getattr(L, name).restype = hid_t
    L.H5Tget_size.restype = C.c_size_t
    if L.H5open() < 0:
        raise OSError("H5open failed")
    _lib = L
    return L


def _native(name):
    return hid_t.in_dll(lib(), name + "_g").value


def _chk(rc, what):
    if rc < 0:
        raise OSError(f"HDF5 call {what} failed")
    return rc


class Writer:
    """Create a file; write groups, datasets (1-d or 2-d float64/int64, or
    an [n, 3] float64 array as compound {x, y, z}) and scalar attributes."""

    def __init__(self, path):
        L = lib()
        self.L = L
        self.f = _chk(L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, hid_t(H5P_DEFAULT), hid_t(H5P_DEFAULT)), "H5Fcreate")

    def group(self, name):
        g = _chk(self.L.H5Gcreate2(hid_t(self.f), name.encode(), hid_t(H5P_DEFAULT), hid_t(H5P_DEFAULT),
                                   hid_t(H5P_DEFAULT)), "H5Gcreate2")
        _chk(self.L.H5Gclose(hid_t(g)), "H5Gclose")

    def dataset(self, name, a, vect3d=False):
        L = self.L
        a = np.ascontiguousarray(a)
        ctype = None
        if vect3d:
            assert a.ndim == 2 and a.shape[1] == 3 and a.dtype == np.float64
            dims = (hsize_t * 1)(a.shape[0])
            ctype = _chk(L.H5Tcreate(H5T_COMPOUND, C.c_size_t(24)), "H5Tcreate")
            dbl = hid_t(_native("H5T_NATIVE_DOUBLE"))
            for i, m in enumerate("xyz"):
                _chk(L.H5Tinsert(hid_t(ctype), m.encode(), C.c_size_t(8 * i), dbl), "H5Tinsert")
            t, rank = ctype, 1
        else:
            t = _native({np.dtype(np.float64): "H5T_NATIVE_DOUBLE", np.dtype(np.int64): "H5T_NATIVE_LLONG",
                         np.dtype(np.int32): "H5T_NATIVE_INT"}[a.dtype])
            rank = a.ndim
            dims = (hsize_t * rank)(*a.shape)
        sp = _chk(L.H5Screate_simple(rank, dims, None), "H5Screate_simple")
        d = _chk(L.H5Dcreate2(hid_t(self.f), name.encode(), hid_t(t), hid_t(sp), hid_t(H5P_DEFAULT),
                              hid_t(H5P_DEFAULT), hid_t(H5P_DEFAULT)), "H5Dcreate2")
        _chk(L.H5Dwrite(hid_t(d), hid_t(t), hid_t(H5S_ALL), hid_t(H5S_ALL), hid_t(H5P_DEFAULT),
                        a.ctypes.data_as(C.c_void_p)), "H5Dwrite")
        L.H5Dclose(hid_t(d))
        L.H5Sclose(hid_t(sp))
        if ctype is not None:
            L.H5Tclose(hid_t(ctype))

    def close(self):
        if self.f is not None:
            self.L.H5Fclose(hid_t(self.f))
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_dataset(path, name, dtype):
    """Read a whole dataset through libhdf5 into `dtype` (a numpy dtype;
    structured dtypes map field by field to a compound memory type)."""
    L = lib()
    f = _chk(L.H5Fopen(path.encode(), H5F_ACC_RDONLY, hid_t(H5P_DEFAULT)), "H5Fopen")
    try:
        d = _chk(L.H5Dopen2(hid_t(f), name.encode(), hid_t(H5P_DEFAULT)), "H5Dopen2")
        sp = _chk(L.H5Dget_space(hid_t(d)), "H5Dget_space")
        rank = L.H5Sget_simple_extent_ndims(hid_t(sp))
        dims = (hsize_t * max(rank, 1))()
        L.H5Sget_simple_extent_dims(hid_t(sp), dims, None)
        shape = tuple(dims[i] for i in range(rank))
        dtype = np.dtype(dtype)
        natives = {np.dtype(np.float64): "H5T_NATIVE_DOUBLE", np.dtype(np.int64): "H5T_NATIVE_LLONG",
                   np.dtype(np.int32): "H5T_NATIVE_INT", np.dtype(np.uint8): "H5T_NATIVE_UCHAR",
                   np.dtype(np.uint16): "H5T_NATIVE_USHORT"}
        mtype, own = None, False
        if dtype.names:
            mtype = _chk(L.H5Tcreate(H5T_COMPOUND, C.c_size_t(dtype.itemsize)), "H5Tcreate")
            own = True
            for n in dtype.names:
                ft, off = dtype.fields[n][:2]
                _chk(L.H5Tinsert(hid_t(mtype), n.encode(), C.c_size_t(off), hid_t(_native(natives[ft]))), "H5Tinsert")
        else:
            mtype = _native(natives[dtype])
        out = np.zeros(shape, dtype)
        _chk(L.H5Dread(hid_t(d), hid_t(mtype), hid_t(H5S_ALL), hid_t(H5S_ALL), hid_t(H5P_DEFAULT),
                       out.ctypes.data_as(C.c_void_p)), "H5Dread")
        if own:
            L.H5Tclose(hid_t(mtype))
        L.H5Sclose(hid_t(sp))
        L.H5Dclose(hid_t(d))
        return out
    finally:
        L.H5Fclose(hid_t(f))
