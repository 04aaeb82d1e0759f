"""ctypes binding of libfpsi_gpu.so (include/fpsi.h).

The names follow the flowPsi facts each call produces (iflux, vflux, src,
residual, fjp/fjm, fluid_D/B, dtau, dfluid).  There is no CPU fallback: a
missing library or a failing HIP call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import grids as _grids

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfpsi_gpu.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# enum values of include/fpsi.h
FLUX = {"hllc": 0, "roe": 1}
LIMITER = {"none": 0, "venkatakrishnan": 1, "zero": 2, "barth": 3}
PRECISION = {"fp64": 0, "fp32": 1}
SOLVER = {"sgs": 0, "lsgs": 1, "fsgs": 2}
SWEEP = {"per_level": 0, "persistent": 1, "tiled": 2, "persistent_flags": 3}
TURB = {"none": 0, "sa": 1, "sst": 2}
TIME_SCHEME = {"bdf2": 1, "kec": 2}
NSCALAR = {"none": 0, "sa": 1, "sst": 2}  # turbulence scalars after the 5 fluid variables
NOALLOC_MAX_STEPS = 4096  # FPSI_NOALLOC_MAX_STEPS
ERRORS = {-1: "FPSI_ERR_ARG", -2: "FPSI_ERR_HIP", -3: "FPSI_ERR_NAN", -4: "FPSI_ERR_STATE", -5: "FPSI_ERR_LU"}

# every entry point declared in include/fpsi.h
EXPORTS = ["fpsi_create", "fpsi_destroy", "fpsi_last_error", "fpsi_get_info", "fpsi_set_params",
           "fpsi_set_state", "fpsi_get_state", "fpsi_residual", "fpsi_gradients", "fpsi_assemble",
           "fpsi_get_system", "fpsi_linear_solve", "fpsi_update", "fpsi_newton_step",
           "fpsi_newton_step_noalloc",
           "fpsi_synchronize", "fpsi_profile_enable", "fpsi_profile_read", "fpsi_linear_residual",
           "fpsi_linear_solve_tol", "fpsi_newton_adaptive", "fpsi_mlp_num_params", "fpsi_mlp_init",
           "fpsi_mlp_get_params", "fpsi_mlp_set_params", "fpsi_mlp_harvest", "fpsi_mlp_train",
           "fpsi_mlp_eval", "fpsi_mlp_get_sample", "fpsi_lsgs_stats", "fpsi_set_wall_distance",
           "fpsi_sa_get_system", "fpsi_sa_solve", "fpsi_sa_lsgs_stats", "fpsi_turb_get_system",
           "fpsi_turb_solve", "fpsi_turb_lsgs_stats", "fpsi_set_time_levels", "fpsi_get_time_levels",
           "fpsi_unsteady_step", "fpsi_finish_time_step"]
KGROUPS = ["residual", "face_jac", "assemble", "sgs_sweep", "update"]


class GridDesc(C.Structure):
    _fields_ = [("nc", C.c_int), ("nf", C.c_int), ("cl", _ip), ("cr", _ip), ("btype", _ip),
                ("pface", _ip), ("n", _dp), ("area", _dp), ("fc", _dp), ("vol", _dp), ("cc", _dp),
                ("pshift", _dp)]


MAX_BC_SETS = 8  # FPSI_MAX_BC_SETS


class BcSet(C.Structure):
    _fields_ = [("T", C.c_double), ("pg", C.c_double), ("u", C.c_double * 3), ("p0", C.c_double),
                ("T0", C.c_double), ("mdot", C.c_double)]


class Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("Rtilde", C.c_double), ("Pambient", C.c_double),
                ("flux", C.c_int), ("limiter", C.c_int), ("Kl", C.c_double), ("viscous", C.c_int),
                ("suth", C.c_double * 6), ("Twall", C.c_double), ("pg0", C.c_double),
                ("T0", C.c_double), ("u0", C.c_double * 3), ("cflmax", C.c_double),
                ("dtmax", C.c_double), ("urelax", C.c_double), ("linear_solver", C.c_int),
                ("sgs_iter", C.c_int), ("precision", C.c_int), ("minPg", C.c_double),
                ("sweep_mode", C.c_int), ("lsgs_maxiter", C.c_int), ("lsgs_reltol", C.c_double),
                ("lsgs_abstol", C.c_double), ("lsgs_relax", C.c_double), ("Minf", C.c_double),
                ("etaT", C.c_double), ("PLimPC", C.c_double), ("turb", C.c_int), ("Prt", C.c_double),
                ("nut_bc", C.c_double), ("k_bc", C.c_double), ("w_bc", C.c_double),
                ("time_mode", C.c_int), ("newton_iter", C.c_int), ("pseudo_cflmax", C.c_double),
                ("pbias_control", C.c_double), ("sweep_spin_limit", C.c_int),
                ("temporal_damping", C.c_double), ("sweep_waves_per_cu", C.c_int),
                ("bc", BcSet * MAX_BC_SETS)]


class Norms(C.Structure):
    _fields_ = [("rtrms", C.c_double), ("etrms", C.c_double), ("mtrms", C.c_double),
                ("rtmax", C.c_double), ("etmax", C.c_double), ("mtmax", C.c_double),
                ("ccmaxr", C.c_double * 3), ("ccmaxe", C.c_double * 3)]

    def as_dict(self):
        return {k: getattr(self, k) for k in ("rtrms", "etrms", "mtrms", "rtmax", "etmax", "mtmax")}


class Info(C.Structure):
    _fields_ = [("nlevels", C.c_int), ("max_level_width", C.c_int), ("nnz_blocks", C.c_int),
                ("device_bytes", C.c_double), ("lsgs_nlevels", C.c_int), ("lsgs_contiguous", C.c_int),
                ("sweep_tiles", C.c_int), ("lsgs_dataflow", C.c_int)]


class LinStats(C.Structure):
    _fields_ = [("r0", C.c_double), ("rguess", C.c_double), ("rfinal", C.c_double),
                ("sweeps", C.c_int), ("used_guess", C.c_int), ("alpha", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Profile(C.Structure):
    _fields_ = [("ms", C.c_double * 5), ("launches", C.c_long * 5), ("bytes", C.c_double * 5)]


class FpsiError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FpsiError(f"{path} is missing: build it with flowpsi_amd.build.build()")
        L = C.CDLL(path)
        L.fpsi_last_error.restype = C.c_char_p
        L.fpsi_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(GridDesc)]
        for name in EXPORTS:
            if name not in ("fpsi_last_error", "fpsi_destroy"):
                getattr(L, name).restype = C.c_int
        L.fpsi_destroy.restype = None
        L.fpsi_mlp_init.argtypes = [C.c_void_p, C.c_ulonglong]
        L.fpsi_mlp_train.argtypes = [C.c_void_p, C.c_int, C.c_float, C.POINTER(C.c_double)]
        L.fpsi_linear_solve_tol.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(LinStats)]
        L.fpsi_newton_adaptive.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p,
                                           C.c_void_p]
        _lib = L
    return _lib


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def make_params(prm: dict) -> Params:
    s = Params()
    s.gamma = prm.get("gamma", 1.4)
    s.Rtilde = prm.get("Rt", 287.0)
    s.Pambient = prm.get("Pambient", 0.0)
    s.flux = FLUX[prm.get("flux", "hllc")]
    s.limiter = LIMITER[prm.get("limiter", "venkatakrishnan")]
    s.Kl = prm.get("Kl", 1.0)
    s.viscous = int(prm.get("viscous", 0))
    s.suth[:] = prm.get("suth", (1.458e-6, 1.5, 110.4, 2.495e-3, 1.5, 194.0))
    s.Twall = prm.get("Twall", 300.0)
    s.pg0 = prm.get("pg0", 101325.0)
    s.T0 = prm.get("T0", 300.0)
    s.u0[:] = prm.get("u0", (0.0, 0.0, 0.0))
    s.cflmax = prm.get("cflmax", 0.0)
    s.dtmax = prm.get("dtmax", 1e30)
    s.urelax = prm.get("urelax", 1e30)
    s.linear_solver = SOLVER[prm.get("solver", "sgs")]
    s.sgs_iter = prm.get("sgs_iter", 5)
    s.precision = PRECISION[prm.get("precision", "fp64")]
    s.minPg = prm.get("minPg", -s.Pambient)
    s.sweep_mode = SWEEP[prm.get("sweep", "per_level")]
    s.lsgs_maxiter = prm.get("lsgs_maxiter", 10)
    s.lsgs_reltol = prm.get("lsgs_reltol", 0.1)
    s.lsgs_abstol = prm.get("lsgs_abstol", 1e-10)
    s.lsgs_relax = prm.get("lsgs_relax", 0.25)
    s.Minf = prm.get("Minf", 1.0)
    s.etaT = prm.get("etaT", 0.25)
    s.PLimPC = prm.get("PLimPC", 10.0)
    s.turb = TURB[prm.get("turb", "none")]
    s.Prt = prm.get("Prt", 0.9)
    s.nut_bc = prm.get("nut_bc", 5e-9)
    s.k_bc = prm.get("k_bc", 0.0)  # <= 0: the library's default from |u0|
    s.w_bc = prm.get("w_bc", 0.0)
    # time-accurate: unsteady=True, time_scheme "bdf2" (default) or "kec",
    # time step dtmax, newton_iter iterations per step (<= 0: 6),
    # pseudo_cflmax (<= 0: 10), pbias_control (KEC's timePressureBiasControl)
    s.time_mode = TIME_SCHEME[prm.get("time_scheme", "bdf2")] if prm.get("unsteady", False) else 0
    s.newton_iter = int(prm.get("newton_iter", 6))
    s.pseudo_cflmax = float(prm.get("pseudo_cflmax", 10.0))
    s.pbias_control = float(prm.get("pbias_control", 0.0))
    s.sweep_spin_limit = int(prm.get("sweep_spin_limit", 0))  # <= 0: the library's bound
    s.temporal_damping = float(prm.get("temporal_damping", 0.0))  # temporalDamping (<= 0: not set)
    s.sweep_waves_per_cu = int(prm.get("sweep_waves_per_cu", 0))  # <= 0: every co-resident wave
    # boundary parameter sets: bc_sets[k] = dict(T, pg or p, u, p0, T0, mdot)
    # for the faces whose btype carries set k (grids.bc_code)
    for k, b in enumerate(prm.get("bc_sets", ())):
        d = s.bc[k]
        d.T = float(b.get("T", 0.0))
        d.pg = float(b["pg"]) if "pg" in b else float(b.get("p", 0.0)) - s.Pambient
        u = b.get("u", (0.0, 0.0, 0.0))
        d.u[:] = (float(u), 0.0, 0.0) if np.isscalar(u) else tuple(float(x) for x in u)
        d.p0 = float(b.get("p0", 0.0))
        d.T0 = float(b.get("T0", 0.0))
        d.mdot = float(b.get("mdot", 0.0))
    return s


class Solver:
    """One device context for one grid (the drop-in for the Newton-step rules)."""

    def __init__(self, grid: "_grids.Grid", prm: dict | None = None):
        self.L = load()
        self.grid = grid
        self._keep = dict(
            cl=np.ascontiguousarray(grid.cl, np.int32), cr=np.ascontiguousarray(grid.cr, np.int32),
            btype=np.ascontiguousarray(grid.btype, np.int32),
            pface=np.ascontiguousarray(grid.pface, np.int32),
            n=np.ascontiguousarray(grid.n, np.float64), area=np.ascontiguousarray(grid.area, np.float64),
            fc=np.ascontiguousarray(grid.fc, np.float64), vol=np.ascontiguousarray(grid.vol, np.float64),
            cc=np.ascontiguousarray(grid.cc, np.float64),
            pshift=np.ascontiguousarray(grid.pshift, np.float64))
        k = self._keep
        gd = GridDesc(grid.nc, grid.nf, _ptr(k["cl"], C.c_int), _ptr(k["cr"], C.c_int),
                      _ptr(k["btype"], C.c_int), _ptr(k["pface"], C.c_int), _ptr(k["n"]),
                      _ptr(k["area"]), _ptr(k["fc"]), _ptr(k["vol"]), _ptr(k["cc"]), _ptr(k["pshift"]))
        h = C.c_void_p()
        rc = self.L.fpsi_create(C.byref(h), C.byref(gd))
        self.h = h
        if rc != 0:
            msg = self.L.fpsi_last_error(h).decode() if h.value else "fpsi_create failed"
            self.close()
            raise FpsiError(f"fpsi_create: {ERRORS.get(rc, rc)}: {msg}")
        self.prm = None
        if prm is not None:
            self.set_params(prm)
        dist = getattr(grid, "meta", {}).get("dist")
        if dist is not None:
            self.set_wall_distance(dist)

    # -- plumbing
    def _chk(self, rc, what):
        if rc != 0:
            raise FpsiError(f"{what}: {ERRORS.get(rc, rc)}: {self.L.fpsi_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.fpsi_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def fj(self):
        return np.float32 if self.prm and self.prm.get("precision") == "fp32" else np.float64

    def info(self) -> dict:
        i = Info()
        self._chk(self.L.fpsi_get_info(self.h, C.byref(i)), "fpsi_get_info")
        return dict(nlevels=i.nlevels, max_level_width=i.max_level_width, nnz_blocks=i.nnz_blocks,
                    device_bytes=i.device_bytes, lsgs_nlevels=i.lsgs_nlevels,
                    lsgs_contiguous=bool(i.lsgs_contiguous), sweep_tiles=i.sweep_tiles,
                    lsgs_dataflow=i.lsgs_dataflow)

    # -- rules
    def set_params(self, prm: dict):
        self.prm = dict(prm)
        p = make_params(prm)
        self._chk(self.L.fpsi_set_params(self.h, C.byref(p)), "fpsi_set_params")

    @property
    def nvar(self):
        return 5 + NSCALAR[(self.prm or {}).get("turb", "none")]

    def set_state(self, q):
        q = np.ascontiguousarray(q, np.float64)
        assert q.shape == (self.nvar, self.grid.nc)
        self._chk(self.L.fpsi_set_state(self.h, _ptr(q)), "fpsi_set_state")

    def get_state(self):
        q = np.zeros((self.nvar, self.grid.nc))
        self._chk(self.L.fpsi_get_state(self.h, _ptr(q)), "fpsi_get_state")
        return q

    def set_wall_distance(self, dist):
        d = np.ascontiguousarray(dist, np.float64)
        self._keep["dist"] = d
        self._chk(self.L.fpsi_set_wall_distance(self.h, _ptr(d)), "fpsi_set_wall_distance")

    def sa_get_system(self):
        nc, nf = self.grid.nc, self.grid.nf
        out = {k: np.zeros(n) for k, n in (("src", nc), ("fjp", nf), ("fjm", nf), ("D", nc), ("B", nc),
                                            ("tmu", nc))}
        self._chk(self.L.fpsi_sa_get_system(self.h, _ptr(out["src"]), _ptr(out["fjp"]), _ptr(out["fjm"]),
                                            _ptr(out["D"]), _ptr(out["B"]), _ptr(out["tmu"])),
                  "fpsi_sa_get_system")
        return out

    def sa_solve(self, nsweeps):
        dX = np.zeros(self.grid.nc)
        self._chk(self.L.fpsi_sa_solve(self.h, int(nsweeps), _ptr(dX)), "fpsi_sa_solve")
        return dX

    def sa_lsgs_stats(self) -> dict:
        it = C.c_int(); w = C.c_double(); r = C.c_double()
        self._chk(self.L.fpsi_sa_lsgs_stats(self.h, C.byref(it), C.byref(w), C.byref(r)), "fpsi_sa_lsgs_stats")
        return dict(iters=it.value, w=w.value, resid=r.value)

    def turb_get_system(self, which):
        """Scalar system `which` (0 nuTilde or k, 1 omega): src, fjp, fjm, D, B, tmu."""
        nc, nf = self.grid.nc, self.grid.nf
        out = {k: np.zeros(n) for k, n in (("src", nc), ("fjp", nf), ("fjm", nf), ("D", nc), ("B", nc),
                                            ("tmu", nc))}
        self._chk(self.L.fpsi_turb_get_system(self.h, int(which), _ptr(out["src"]), _ptr(out["fjp"]),
                                              _ptr(out["fjm"]), _ptr(out["D"]), _ptr(out["B"]),
                                              _ptr(out["tmu"])), "fpsi_turb_get_system")
        return out

    def turb_solve(self, which, nsweeps):
        dX = np.zeros(self.grid.nc)
        self._chk(self.L.fpsi_turb_solve(self.h, int(which), int(nsweeps), _ptr(dX)), "fpsi_turb_solve")
        return dX

    def turb_lsgs_stats(self, which) -> dict:
        it = C.c_int(); w = C.c_double(); r = C.c_double()
        self._chk(self.L.fpsi_turb_lsgs_stats(self.h, int(which), C.byref(it), C.byref(w), C.byref(r)),
                  "fpsi_turb_lsgs_stats")
        return dict(iters=it.value, w=w.value, resid=r.value)

    def residual(self, want_faces=True):
        nc, nf = self.grid.nc, self.grid.nf
        src = np.zeros((5, nc))
        iflux = np.zeros((5, nf)) if want_faces else None
        vflux = np.zeros((4, nf)) if want_faces else None
        nrm = Norms()
        self._chk(self.L.fpsi_residual(self.h, _ptr(src), _ptr(iflux) if want_faces else None,
                                       _ptr(vflux) if want_faces else None, C.byref(nrm)),
                  "fpsi_residual")
        return src, iflux, vflux, nrm

    def gradients(self):
        nc = self.grid.nc
        grad = np.zeros((15, nc)); phi = np.zeros((5, nc))
        self._chk(self.L.fpsi_gradients(self.h, _ptr(grad), _ptr(phi)), "fpsi_gradients")
        return grad, phi

    def assemble(self):
        self._chk(self.L.fpsi_assemble(self.h), "fpsi_assemble")

    def get_system(self):
        nc, nf = self.grid.nc, self.grid.nf
        fj = self.fj
        fjp = np.zeros((nf, 25), fj); fjm = np.zeros((nf, 25), fj)
        D = np.zeros((nc, 25), fj); piv = np.zeros((nc, 5), np.int32); B = np.zeros((nc, 5), fj)
        dtau = np.zeros(nc)
        self._chk(self.L.fpsi_get_system(self.h, fjp.ctypes.data_as(C.c_void_p),
                                         fjm.ctypes.data_as(C.c_void_p), D.ctypes.data_as(C.c_void_p),
                                         _ptr(piv, C.c_int), B.ctypes.data_as(C.c_void_p), _ptr(dtau)),
                  "fpsi_get_system")
        return dict(fjp=fjp, fjm=fjm, D=D, piv=piv, B=B, dtau=dtau)

    def linear_solve(self, nsweeps):
        dq = np.zeros((self.grid.nc, 5))
        self._chk(self.L.fpsi_linear_solve(self.h, int(nsweeps), _ptr(dq)), "fpsi_linear_solve")
        return dq

    def lsgs_stats(self) -> dict:
        it = C.c_int(); w = C.c_double(); r = C.c_double()
        self._chk(self.L.fpsi_lsgs_stats(self.h, C.byref(it), C.byref(w), C.byref(r)), "fpsi_lsgs_stats")
        return dict(iters=it.value, w=w.value, resid=r.value)

    def update(self):
        self._chk(self.L.fpsi_update(self.h), "fpsi_update")

    def newton(self, nsteps):
        hist = (Norms * max(nsteps, 1))()
        self._chk(self.L.fpsi_newton_step(self.h, int(nsteps), hist), "fpsi_newton_step")
        return [hist[i].as_dict() for i in range(nsteps)]

    def newton_noalloc(self, nsteps):
        """newton() through the entry point that allocates nothing: the
        norm history sits in the library's fixed per-GPU array
        (at most NOALLOC_MAX_STEPS steps per call)."""
        hist = (Norms * max(nsteps, 1))()
        self._chk(self.L.fpsi_newton_step_noalloc(self.h, int(nsteps), hist), "fpsi_newton_step_noalloc")
        return [hist[i].as_dict() for i in range(nsteps)]

    # -- time-accurate BDF2
    def unsteady(self, nsteps):
        """nsteps time steps (BDF2 or KEC); the norms of every Newton
        iteration as a list of nsteps lists of newton_iter dicts."""
        ni = int(self.prm.get("newton_iter", 6))
        ni = ni if ni > 0 else 6
        hist = (Norms * max(nsteps * ni, 1))()
        self._chk(self.L.fpsi_unsteady_step(self.h, int(nsteps), hist), "fpsi_unsteady_step")
        return [[hist[s * ni + i].as_dict() for i in range(ni)] for s in range(nsteps)]

    def finish_time_step(self):
        self._chk(self.L.fpsi_finish_time_step(self.h), "fpsi_finish_time_step")

    def set_time_levels(self, qn=None, qnm1=None):
        arrs = []
        for a in (qn, qnm1):
            if a is not None:
                a = np.ascontiguousarray(a, np.float64)
                assert a.shape == (self.nvar, self.grid.nc)
            arrs.append(a)
        self._chk(self.L.fpsi_set_time_levels(self.h, *(None if a is None else _ptr(a) for a in arrs)),
                  "fpsi_set_time_levels")

    def get_time_levels(self):
        qn, qnm1 = np.zeros((self.nvar, self.grid.nc)), np.zeros((self.nvar, self.grid.nc))
        self._chk(self.L.fpsi_get_time_levels(self.h, _ptr(qn), _ptr(qnm1)), "fpsi_get_time_levels")
        return qn, qnm1

    # -- linear tolerance and learned guess
    def linear_residual(self):
        v = C.c_double()
        self._chk(self.L.fpsi_linear_residual(self.h, C.byref(v)), "fpsi_linear_residual")
        return v.value

    def linear_solve_tol(self, rtol, max_sweeps, use_guess=False):
        st = LinStats()
        self._chk(self.L.fpsi_linear_solve_tol(self.h, rtol, int(max_sweeps), int(use_guess), C.byref(st)),
                  "fpsi_linear_solve_tol")
        return st.as_dict()

    def newton_adaptive(self, nsteps, rtol, max_sweeps, use_guess=False):
        hist = (Norms * max(nsteps, 1))()
        lin = (LinStats * max(nsteps, 1))()
        self._chk(self.L.fpsi_newton_adaptive(self.h, int(nsteps), rtol, int(max_sweeps), int(use_guess), hist, lin),
                  "fpsi_newton_adaptive")
        return [hist[i].as_dict() for i in range(nsteps)], [lin[i].as_dict() for i in range(nsteps)]

    def mlp_init(self, seed=1):
        self._chk(self.L.fpsi_mlp_init(self.h, seed), "fpsi_mlp_init")

    def mlp_num_params(self):
        return self.L.fpsi_mlp_num_params()

    def mlp_params(self):
        p = np.zeros(self.mlp_num_params(), np.float32)
        self._chk(self.L.fpsi_mlp_get_params(self.h, _ptr(p, C.c_float)), "fpsi_mlp_get_params")
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(self.L.fpsi_mlp_set_params(self.h, _ptr(p, C.c_float)), "fpsi_mlp_set_params")

    def mlp_harvest(self, target_sweeps):
        slot = C.c_int()
        self._chk(self.L.fpsi_mlp_harvest(self.h, int(target_sweeps), C.byref(slot)), "fpsi_mlp_harvest")
        return slot.value

    def mlp_train(self, epochs, lr=1e-3):
        loss = C.c_double()
        self._chk(self.L.fpsi_mlp_train(self.h, int(epochs), float(lr), C.byref(loss)), "fpsi_mlp_train")
        return loss.value

    def mlp_eval(self, slot):
        nc = self.grid.nc
        out = np.zeros((nc, 16), np.float32)
        grad = np.zeros(self.mlp_num_params(), np.float32)
        loss = C.c_double()
        self._chk(self.L.fpsi_mlp_eval(self.h, int(slot), _ptr(out, C.c_float), _ptr(grad, C.c_float),
                                       C.byref(loss)), "fpsi_mlp_eval")
        return out, grad, loss.value

    def mlp_sample(self, slot):
        nc = self.grid.nc
        x = np.zeros((nc, 32), np.float32)
        y = np.zeros((nc, 8), np.float32)
        off = np.zeros(nc + 1, np.int32)
        self._chk(self.L.fpsi_mlp_get_sample(self.h, int(slot), _ptr(x, C.c_float), _ptr(y, C.c_float),
                                             _ptr(off, C.c_int), None), "fpsi_mlp_get_sample")
        idx = np.zeros(int(off[-1]), np.int32)
        self._chk(self.L.fpsi_mlp_get_sample(self.h, int(slot), None, None, _ptr(off, C.c_int),
                                             _ptr(idx, C.c_int)), "fpsi_mlp_get_sample")
        return x, y, off, idx

    def profile(self, on: bool):
        self._chk(self.L.fpsi_profile_enable(self.h, int(on)), "fpsi_profile_enable")

    def profile_read(self) -> dict:
        p = Profile()
        self._chk(self.L.fpsi_profile_read(self.h, C.byref(p)), "fpsi_profile_read")
        return {k: dict(ms=p.ms[i], launches=p.launches[i], bytes=p.bytes[i]) for i, k in enumerate(KGROUPS)}

    def synchronize(self):
        self._chk(self.L.fpsi_synchronize(self.h), "fpsi_synchronize")
