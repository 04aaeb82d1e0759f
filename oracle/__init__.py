"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OgGrid(C.Structure):
    _fields_ = [("nc", C.c_int), ("nf", C.c_int), ("cl", _ip), ("cr", _ip), ("btype", _ip),
                ("pface", _ip), ("n", _dp), ("area", _dp), ("fc", _dp), ("vol", _dp),
                ("cc", _dp), ("pshift", _dp), ("dist", _dp)]


class OgParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("Rt", C.c_double), ("Pambient", C.c_double),
                ("flux", C.c_int), ("limiter", C.c_int), ("Kl", C.c_double),
                ("viscous", C.c_int), ("suth", C.c_double * 6), ("Twall", C.c_double),
                ("pg0", C.c_double), ("T0", C.c_double), ("u0", C.c_double * 3),
                ("cflmax", C.c_double), ("dtmax", C.c_double), ("urelax", C.c_double),
                ("sgs_iter", C.c_int), ("minPg", C.c_double), ("linear_solver", C.c_int),
                ("lsgs_maxiter", C.c_int), ("lsgs_reltol", C.c_double), ("lsgs_abstol", C.c_double),
                ("lsgs_relax", C.c_double), ("Minf", C.c_double), ("etaT", C.c_double),
                ("PLimPC", C.c_double), ("turb", C.c_int), ("sa_cb1", C.c_double), ("sa_k", C.c_double),
                ("sa_cv1", C.c_double), ("sa_cb2", C.c_double), ("sa_cw2", C.c_double), ("sa_cw3", C.c_double),
                ("sa_sigma", C.c_double), ("Prt", C.c_double), ("nut_bc", C.c_double), ("k_bc", C.c_double),
                ("w_bc", C.c_double), ("unsteady", C.c_int), ("newton_iter", C.c_int),
                ("pseudo_cflmax", C.c_double), ("pbias_control", C.c_double),
                ("temporal_damping", C.c_double), ("bcs", (C.c_double * 8) * 8)]


class OgNorms(C.Structure):
    _fields_ = [("rtrms", C.c_double), ("etrms", C.c_double), ("mtrms", C.c_double),
                ("rtmax", C.c_double), ("etmax", C.c_double), ("mtmax", C.c_double),
                ("ccmaxr", C.c_double * 3), ("ccmaxe", C.c_double * 3)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


_libs = {}


def lib(fj="f64"):
    if fj not in _libs:
        path = os.path.join(BUILD, f"liboracle_{fj}.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.og_real_fj_bytes.restype = C.c_int
        L.og_lu_decompose.restype = C.c_int
        L.og_newton.restype = C.c_int
        L.og_eta_face.restype = C.c_double
        _libs[fj] = L
    return _libs[fj]


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class OGrid:
    """Keeps the numpy arrays alive behind an og_grid struct."""

    def __init__(self, g):
        self.g = g
        self.arrs = dict(
            cl=np.ascontiguousarray(g.cl, np.int32), cr=np.ascontiguousarray(g.cr, np.int32),
            btype=np.ascontiguousarray(g.btype, np.int32), pface=np.ascontiguousarray(g.pface, np.int32),
            n=np.ascontiguousarray(g.n, np.float64), area=np.ascontiguousarray(g.area, np.float64),
            fc=np.ascontiguousarray(g.fc, np.float64), vol=np.ascontiguousarray(g.vol, np.float64),
            cc=np.ascontiguousarray(g.cc, np.float64), pshift=np.ascontiguousarray(g.pshift, np.float64))
        dist = g.meta.get("dist") if getattr(g, "meta", None) else None
        self.arrs["dist"] = np.ascontiguousarray(dist if dist is not None else np.ones(g.nc), np.float64)
        a = self.arrs
        self.s = OgGrid(g.nc, g.nf, _ptr(a["cl"], C.c_int), _ptr(a["cr"], C.c_int),
                        _ptr(a["btype"], C.c_int), _ptr(a["pface"], C.c_int), _ptr(a["n"]),
                        _ptr(a["area"]), _ptr(a["fc"]), _ptr(a["vol"]), _ptr(a["cc"]),
                        _ptr(a["pshift"]), _ptr(a["dist"]))

    @property
    def ref(self):
        return C.byref(self.s)


def params(p: dict) -> OgParams:
    s = OgParams()
    s.gamma = p.get("gamma", 1.4)
    s.Rt = p.get("Rt", 287.0)
    s.Pambient = p.get("Pambient", 0.0)
    s.flux = {"hllc": 0, "roe": 1}[p.get("flux", "hllc")]
    s.limiter = {"none": 0, "venkatakrishnan": 1, "zero": 2, "barth": 3}[p.get("limiter", "venkatakrishnan")]
    s.Kl = p.get("Kl", 1.0)
    s.viscous = int(p.get("viscous", 0))
    s.suth[:] = p.get("suth", (1.458e-6, 1.5, 110.4, 2.495e-3, 1.5, 194.0))
    s.Twall = p.get("Twall", 300.0)
    s.pg0 = p.get("pg0", 101325.0)
    s.T0 = p.get("T0", 300.0)
    s.u0[:] = p.get("u0", (0.0, 0.0, 0.0))
    s.cflmax = p.get("cflmax", 0.0)
    s.dtmax = p.get("dtmax", 1e30)
    s.urelax = p.get("urelax", 1e30)
    s.sgs_iter = p.get("sgs_iter", 5)
    s.minPg = p.get("minPg", -s.Pambient)
    s.linear_solver = {"sgs": 0, "lsgs": 1, "fsgs": 2}[p.get("solver", "sgs")]
    s.lsgs_maxiter = p.get("lsgs_maxiter", 10)
    s.lsgs_reltol = p.get("lsgs_reltol", 0.1)
    s.lsgs_abstol = p.get("lsgs_abstol", 1e-10)
    s.lsgs_relax = p.get("lsgs_relax", 0.25)
    s.Minf = p.get("Minf", 1.0)
    s.etaT = p.get("etaT", 0.25)
    s.PLimPC = p.get("PLimPC", 10.0)
    s.turb = {"none": 0, "sa": 1, "sst": 2}[p.get("turb", "none")]
    # Spa_All_param defaults, turbulence/SA/SA_param.h:25-34
    s.sa_cb1, s.sa_k, s.sa_cv1, s.sa_cb2 = 0.1355, 0.41, 7.1, 0.622
    s.sa_cw2, s.sa_cw3, s.sa_sigma = 0.3, 2.0, 2. / 3.
    s.Prt = p.get("Prt", 0.9)
    s.nut_bc = p.get("nut_bc", 5e-9)
    kb, wb = sst_bc_defaults(s.u0)
    s.k_bc = p.get("k_bc", kb)
    s.w_bc = p.get("w_bc", wb)
    # timeStepMode unsteady with the BDF2 scheme (src/solverTimestep.loci),
    # newtonMaxIter (src/solverNewton.loci:197-199), pseudo_cflmax
    # (src/solverLinearSystem.loci:380-382)
    # timeStepScheme BDF2 (1) or KEC (2)
    s.unsteady = {"bdf2": 1, "kec": 2}[p.get("time_scheme", "bdf2")] if p.get("unsteady", False) else 0
    s.newton_iter = int(p.get("newton_iter", 6))
    s.pseudo_cflmax = float(p.get("pseudo_cflmax", 10.0))
    s.pbias_control = float(p.get("pbias_control", 0.0))  # timePressureBiasControl
    s.temporal_damping = float(p.get("temporal_damping", 0.0))  # temporalDamping (<= 0: not set)
    # boundary parameter sets: bc_sets[k] = dict(T, pg or p, u, p0, T0, mdot)
    for k, b in enumerate(p.get("bc_sets", ())):
        u = b.get("u", (0.0, 0.0, 0.0))
        u = (float(u), 0.0, 0.0) if np.isscalar(u) else tuple(float(x) for x in u)
        pg = float(b["pg"]) if "pg" in b else float(b.get("p", 0.0)) - s.Pambient
        s.bcs[k][:] = (float(b.get("T", 0.0)), pg, *u, float(b.get("p0", 0.0)), float(b.get("T0", 0.0)),
                       float(b.get("mdot", 0.0)))
    return s


def sst_bc_defaults(u0):
    """k and omega of an inflow boundary without k/w options
    (turbulence/KOmega/sst_viscous.loci:63-118): k = max(1.5 (0.01 |u|)^2,
    1e-3), omega = 0.09 k / (1e-3 * 1e-5)."""
    v = float(np.linalg.norm(np.asarray(list(u0), float)))
    k = max(1.5 * (1e-2 * v) * (1e-2 * v), 1e-3)
    return k, 0.09 * k / (1e-3 * 1e-5)


class Oracle:
    """Whole-grid stages of the oracle on one grid."""

    def __init__(self, g, prm: dict, fj="f64", omp=False):
        """omp: the OpenMP build (face/cell loops on all OMP_NUM_THREADS
        threads), used only to time the CPU baseline."""
        self.g = g
        self.og = OGrid(g)
        self.p = params(prm)
        self.L = lib(("omp_" if omp else "") + fj)
        self.fj = np.float64 if fj == "f64" else np.float32

    def eta(self, q):
        """Eta_p per cell."""
        q = np.ascontiguousarray(q, np.float64)
        e = np.zeros(self.g.nc)
        self.L.og_eta(self.og.ref, C.byref(self.p), _ptr(q), _ptr(e))
        return e

    def face_states(self, q):
        q = np.ascontiguousarray(q, np.float64)
        qf = np.zeros((7, self.g.nf))
        self.L.og_face_states(self.og.ref, C.byref(self.p), _ptr(q), _ptr(qf))
        return np.ascontiguousarray(qf[:self.nq])

    @property
    def nq(self):
        return {0: 5, 1: 6, 2: 7}[self.p.turb]

    def gradients(self, q, qf=None):
        q = np.ascontiguousarray(q, np.float64)
        if qf is None:
            qf = self.face_states(q)
        nq = self.nq
        grad = np.zeros((3 * nq, self.g.nc)); phi = np.zeros((nq, self.g.nc))
        self.L.og_gradients(self.og.ref, C.byref(self.p), _ptr(q), _ptr(qf), _ptr(grad), _ptr(phi))
        return grad, phi

    def fluxes(self, q):
        q = np.ascontiguousarray(q, np.float64)
        qf = self.face_states(q)
        grad, phi = self.gradients(q, qf)
        nf = self.g.nf
        iflux = np.zeros((5, nf)); vflux = np.zeros((4, nf)); maxev = np.zeros(nf)
        self.L.og_fluxes(self.og.ref, C.byref(self.p), _ptr(q), _ptr(qf), _ptr(grad), _ptr(phi),
                         _ptr(iflux), _ptr(vflux), _ptr(maxev))
        return dict(qf=qf, grad=grad, phi=phi, iflux=iflux, vflux=vflux, maxev=maxev)

    def residual(self, q):
        q = np.ascontiguousarray(q, np.float64)
        nc, nf = self.g.nc, self.g.nf
        src = np.zeros((5, nc)); iflux = np.zeros((5, nf)); vflux = np.zeros((4, nf))
        nrm = OgNorms()
        self.L.og_residual(self.og.ref, C.byref(self.p), _ptr(q), _ptr(src), _ptr(iflux),
                           _ptr(vflux), C.byref(nrm))
        return src, iflux, vflux, nrm

    def _levels(self, qn, qm1):
        if qn is None:
            return None, None
        return (np.ascontiguousarray(qn, np.float64),
                None if qm1 is None else np.ascontiguousarray(qm1, np.float64))

    def assemble(self, q, qn=None, qm1=None):
        """The fluid system at q; with the time levels qn, qm1 (unsteady
        parameters) the BDF2 or KEC system of a Newton iteration (KEC reads
        only qn; pbias holds its localPBias)."""
        q = np.ascontiguousarray(q, np.float64)
        qn, qm1 = self._levels(qn, qm1)
        nc, nf = self.g.nc, self.g.nf
        src = np.zeros((5, nc)); dtau = np.zeros(nc)
        fjp = np.zeros((nf, 25), self.fj); fjm = np.zeros((nf, 25), self.fj)
        D = np.zeros((nc, 25), self.fj); piv = np.zeros((nc, 5), np.int32)
        B = np.zeros((nc, 5), self.fj); Draw = np.zeros((nc, 25), self.fj)
        pbias = np.zeros(nc)
        nrm = OgNorms()
        self.L.og_assemble_t(self.og.ref, C.byref(self.p), _ptr(q),
                             None if qn is None else _ptr(qn), None if qm1 is None else _ptr(qm1),
                             _ptr(src), _ptr(dtau), fjp.ctypes.data_as(C.c_void_p), fjm.ctypes.data_as(C.c_void_p),
                             D.ctypes.data_as(C.c_void_p), _ptr(piv, C.c_int),
                             B.ctypes.data_as(C.c_void_p), C.byref(nrm), Draw.ctypes.data_as(C.c_void_p),
                             _ptr(pbias))
        return dict(src=src, dtau=dtau, fjp=fjp, fjm=fjm, D=D, piv=piv, B=B, norms=nrm, Draw=Draw, pbias=pbias)

    def sa_assemble(self, q, qn=None, qm1=None):
        """nuTilde system: src, fjp, fjm (per face), D, B, tmu."""
        q = np.ascontiguousarray(q, np.float64)
        qn, qm1 = self._levels(qn, qm1)
        nc, nf = self.g.nc, self.g.nf
        out = {k: np.zeros(n) for k, n in (("src", nc), ("fjp", nf), ("fjm", nf), ("D", nc), ("B", nc),
                                            ("tmu", nc))}
        self.L.og_sa_assemble_t(self.og.ref, C.byref(self.p), _ptr(q), None if qn is None else _ptr(qn),
                                None if qm1 is None else _ptr(qm1), _ptr(out["src"]), _ptr(out["fjp"]),
                              _ptr(out["fjm"]), _ptr(out["D"]), _ptr(out["B"]), _ptr(out["tmu"]))
        return out

    def sst_assemble(self, q, qn=None, qm1=None):
        """k and omega systems: src, D, B [2][nc]; fjp, fjm [2][nf]; tmu;
        aux F1, F2, gkgw (grad k . grad omega), CD (cross diffusion)."""
        q = np.ascontiguousarray(q, np.float64)
        qn, qm1 = self._levels(qn, qm1)
        nc, nf = self.g.nc, self.g.nf
        out = {k: np.zeros(n) for k, n in (("src", (2, nc)), ("fjp", (2, nf)), ("fjm", (2, nf)),
                                            ("D", (2, nc)), ("B", (2, nc)), ("tmu", nc), ("aux", (4, nc)))}
        self.L.og_sst_assemble_t(self.og.ref, C.byref(self.p), _ptr(q), None if qn is None else _ptr(qn),
                                 None if qm1 is None else _ptr(qm1), _ptr(out["src"]), _ptr(out["fjp"]),
                               _ptr(out["fjm"]), _ptr(out["D"]), _ptr(out["B"]), _ptr(out["tmu"]),
                               _ptr(out["aux"]))
        for i, k in enumerate(("F1", "F2", "gkgw", "CD")):
            out[k] = out["aux"][i]
        return out

    @staticmethod
    def scalar(sys_, j):
        """Scalar system j (0 = k, 1 = omega) of an SST system, in the layout
        sa_sgs / sa_lsgs take."""
        return {k: np.ascontiguousarray(sys_[k][j]) for k in ("src", "fjp", "fjm", "D", "B")}

    def sa_sgs(self, sys_, nsweeps):
        X = np.zeros(self.g.nc)
        self.L.og_sa_sgs(self.og.ref, int(nsweeps), _ptr(sys_["fjp"]), _ptr(sys_["fjm"]), _ptr(sys_["D"]),
                         _ptr(sys_["B"]), _ptr(X))
        return X

    def sa_lsgs(self, sys_, maxiter=10, abstol=1e-10, reltol=0.1, w=0.25):
        X = np.zeros(self.g.nc)
        wo = C.c_double(); ro = C.c_double()
        self.L.og_sa_lsgs.restype = C.c_int
        it = self.L.og_sa_lsgs(self.og.ref, int(maxiter), C.c_double(abstol), C.c_double(reltol), C.c_double(w),
                               _ptr(sys_["fjp"]), _ptr(sys_["fjm"]), _ptr(sys_["D"]), _ptr(sys_["B"]), _ptr(X),
                               C.byref(wo), C.byref(ro))
        return X, dict(iters=it, w=wo.value, resid=ro.value)

    def lsgs_order(self):
        order = np.zeros(self.g.nc, np.int32)
        n = self.L.og_lsgs_order(self.og.ref, _ptr(order, C.c_int))
        assert n == self.g.nc
        return order

    def lsgs(self, sys, maxiter=10, abstol=1e-10, reltol=0.1, w=0.25):
        X = np.zeros((self.g.nc, 5))
        wo = C.c_double(); ro = C.c_double()
        self.L.og_lsgs.restype = C.c_int
        it = self.L.og_lsgs(self.og.ref, int(maxiter), C.c_double(abstol), C.c_double(reltol), C.c_double(w),
                            sys["fjp"].ctypes.data_as(C.c_void_p), sys["fjm"].ctypes.data_as(C.c_void_p),
                            sys["Draw"].ctypes.data_as(C.c_void_p), sys["B"].ctypes.data_as(C.c_void_p),
                            _ptr(X), C.byref(wo), C.byref(ro))
        return X, dict(iters=it, w=wo.value, resid=ro.value)

    def fsgs(self, sys, nsweeps):
        """FSGS from zero: nsweeps symmetric sweeps in the line order."""
        X = np.zeros((self.g.nc, 5))
        self.L.og_fsgs(self.og.ref, int(nsweeps), sys["fjp"].ctypes.data_as(C.c_void_p),
                       sys["fjm"].ctypes.data_as(C.c_void_p), sys["Draw"].ctypes.data_as(C.c_void_p),
                       sys["B"].ctypes.data_as(C.c_void_p), _ptr(X))
        return X

    def sa_fsgs(self, sys_, nsweeps):
        X = np.zeros(self.g.nc)
        self.L.og_sa_fsgs(self.og.ref, int(nsweeps), _ptr(sys_["fjp"]), _ptr(sys_["fjm"]), _ptr(sys_["D"]),
                          _ptr(sys_["B"]), _ptr(X))
        return X

    def sgs(self, sys, nsweeps):
        X = np.zeros((self.g.nc, 5))
        self.L.og_sgs(self.og.ref, nsweeps, sys["fjp"].ctypes.data_as(C.c_void_p),
                      sys["fjm"].ctypes.data_as(C.c_void_p), sys["D"].ctypes.data_as(C.c_void_p),
                      _ptr(sys["piv"], C.c_int), sys["B"].ctypes.data_as(C.c_void_p), _ptr(X))
        return X

    def newton(self, q, nsteps):
        q = np.array(q, np.float64, copy=True, order="C")
        hist = np.zeros((nsteps, 6))
        self.L.og_newton(self.og.ref, C.byref(self.p), _ptr(q), nsteps, _ptr(hist))
        return q, hist

    def unsteady(self, q, qm1, nsteps):
        """nsteps BDF2 time steps from level n (q) and n-1 (qm1): returns the
        new (q, qm1) and hist [nsteps, newton_iter, 6]."""
        assert self.p.unsteady, "unsteady parameters required"
        q = np.array(q, np.float64, copy=True, order="C")
        qm1 = np.array(qm1, np.float64, copy=True, order="C")
        hist = np.zeros((nsteps, self.p.newton_iter, 6))
        self.L.og_unsteady.restype = C.c_int
        rc = self.L.og_unsteady(self.og.ref, C.byref(self.p), _ptr(q), _ptr(qm1), nsteps, _ptr(hist))
        assert rc == 0
        return q, qm1, hist


# ---- point functions -------------------------------------------------------

def sa_point(prm, T, pg, nut, vort, y, vol=1.0):
    """(tmu, production - destruction, destruction Jacobian) of one cell."""
    out = np.zeros(3)
    p = params(dict(prm, turb="sa"))
    lib().og_sa_point(C.byref(p), C.c_double(T), C.c_double(pg), C.c_double(nut), C.c_double(vort),
                      C.c_double(y), C.c_double(vol), _ptr(out))
    return out


def sst_point(prm, T, pg, k, w, gu, y, gkgw, vol=1.0):
    """dict F1, F2, tmu, ksrc, wsrc, kJ, wJ, Pk, CD of one SST cell
    (gu[3i+j] = d u_i / d x_j)."""
    out = np.zeros(9)
    p = params(dict(prm, turb="sst"))
    gu = np.ascontiguousarray(gu, np.float64).ravel()
    lib().og_sst_point(C.byref(p), C.c_double(T), C.c_double(pg), C.c_double(k), C.c_double(w), _ptr(gu),
                       C.c_double(y), C.c_double(gkgw), C.c_double(vol), _ptr(out))
    return dict(zip(("F1", "F2", "tmu", "ksrc", "wsrc", "kJ", "wJ", "Pk", "CD"), out))


def hllc_flux(qL, qR, n, area=1.0, pamb=0.0, Rt=287.0, gamma=1.4, eta=1.0, fj="f64"):
    f = np.zeros(5)
    n = np.asarray(n, float); UL = np.asarray(qL[1:4], float); UR = np.asarray(qR[1:4], float)
    lib(fj).og_hllc_flux(_ptr(f), C.c_double(qL[4]), C.c_double(qL[0]), _ptr(UL), C.c_double(qR[4]),
                         C.c_double(qR[0]), _ptr(UR), _ptr(n), C.c_double(area), C.c_double(pamb),
                         C.c_double(Rt), C.c_double(gamma), C.c_double(0.0), C.c_double(eta))
    return f


def roe_flux(qL, qR, n, area=1.0, pamb=0.0, Rt=287.0, gamma=1.4, eta=1.0, fj="f64"):
    f = np.zeros(5)
    n = np.asarray(n, float); UL = np.asarray(qL[1:4], float); UR = np.asarray(qR[1:4], float)
    lib(fj).og_roe_flux(_ptr(f), C.c_double(qL[4]), C.c_double(qL[0]), _ptr(UL), C.c_double(qR[4]),
                        C.c_double(qR[0]), _ptr(UR), _ptr(n), C.c_double(area), C.c_double(pamb),
                        C.c_double(Rt), C.c_double(gamma), C.c_double(0.0), C.c_double(eta))
    return f


def inviscid_flux(q, n, area=1.0, pamb=0.0, Rt=287.0, gamma=1.4):
    f = np.zeros(5)
    n = np.asarray(n, float); U = np.asarray(q[1:4], float)
    lib().og_inviscid_flux(_ptr(f), C.c_double(q[4]), C.c_double(q[0]), _ptr(U), _ptr(n),
                           C.c_double(area), C.c_double(pamb), C.c_double(Rt), C.c_double(gamma),
                           C.c_double(0.0))
    return f


def roe_jacobian(qL, qR, n, dir, area=1.0, pamb=0.0, Rt=287.0, gamma=1.4, eta=1.0, fj="f64"):
    """5x5 block as a numpy [row, col] matrix."""
    L = lib(fj)
    dt = np.float64 if fj == "f64" else np.float32
    blk = np.zeros(25, dt)
    n = np.asarray(n, float); UL = np.asarray(qL[1:4], float); UR = np.asarray(qR[1:4], float)
    L.og_roe_jacobian(blk.ctypes.data_as(C.c_void_p), C.c_double(qL[4]), C.c_double(qL[0]), _ptr(UL),
                      C.c_double(qR[4]), C.c_double(qR[0]), _ptr(UR), _ptr(n), C.c_double(area),
                      C.c_double(pamb), C.c_double(Rt), C.c_double(gamma), C.c_double(0.0),
                      C.c_double(eta), C.c_int(dir))
    return blk.reshape(5, 5).T.astype(np.float64)  # column-major -> [row, col]


def lu_decompose(A, fj="f64"):
    dt = np.float64 if fj == "f64" else np.float32
    a = np.ascontiguousarray(np.asarray(A, dt).T).ravel().copy()  # column-major
    piv = np.zeros(5, np.int32)
    rc = lib(fj).og_lu_decompose(a.ctypes.data_as(C.c_void_p), _ptr(piv, C.c_int), 5)
    return a, piv, rc


def lu_solve(a, piv, b, fj="f64"):
    x = np.zeros(5)
    b = np.asarray(b, np.float64)
    lib(fj).og_lu_solve(a.ctypes.data_as(C.c_void_p), _ptr(piv, C.c_int), _ptr(b), _ptr(x), 5)
    return x


def cv_state(kind, qi, q0, n, pamb=0.0, Rt=287.0, gamma=1.4, beta=1.0):
    pgb = C.c_double(); Tb = C.c_double(); ub = np.zeros(3)
    ui = np.asarray(qi[1:4], float); u0 = np.asarray(q0[1:4], float); n = np.asarray(n, float)
    fn = lib().og_cv_inflow if kind == "inflow" else lib().og_cv_outflow
    fn(C.byref(pgb), C.byref(Tb), _ptr(ub), C.c_double(qi[4]), C.c_double(qi[0]), _ptr(ui),
       C.c_double(q0[4]), C.c_double(q0[0]), _ptr(u0), _ptr(n), C.c_double(0.0), C.c_double(pamb),
       C.c_double(gamma), C.c_double(Rt), C.c_double(beta))
    return np.array([Tb.value, *ub, pgb.value])
