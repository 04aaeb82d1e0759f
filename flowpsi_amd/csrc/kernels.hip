// Newton-step kernels for gfx950.  One lane per face or per cell; no atomics:
// every cell quantity is gathered through the cell's face list in a fixed
// order, so results are bitwise reproducible run to run.
#include "kernels.h"
#include "blocks.h"
#include "physics.h"

#include <algorithm>

namespace fpsi {

static constexpr int TPB = 256;
static inline int nblk(long n) { return (int)((n + TPB - 1) / TPB); }

// ---------------------------------------------------------------------------
// boundary-face states: farfield src/bcInflow.loci:311-340, viscous wall
// src/bcViscousWall.loci:205-214 (stationary), symmetry src/bcImpermeable.loci:48-64
// ---------------------------------------------------------------------------
// Eta_pf of a boundary face from its interior cell, src/fluxInviscid.loci:151-155
__device__ __forceinline__ double eta_face(const Prm &p, double T, const V3 &U) {
  if (!p.pre) return 1.0;
  const double a = sqrt(p.gamma * p.Rt * T);
  const double M = sqrt(dot(U, U)) / a;
  const double M2 = 2. * dmin(0.25, M * M);
  return dmin(1.0, dmax(M2 / (1. - M2), p.Mlim2));
}

__global__ void k_bface(DevGrid g, Prm p, const double *__restrict__ q, double *__restrict__ qf) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nb) return;
  const int f = g.bnd_begin + b;
  const int c = g.fcl[f];
  const int nc = g.nc;
  const V3 n = {g.fnx[f], g.fny[f], g.fnz[f]};
  Q5 i;
  i.T = q[c];
  i.u = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
  i.pg = q[4 * nc + c];
  Q5 o;
  const int t = g.fbt[f] & 0xFF;
  const double *bs = p.bcs + 8 * (g.fbt[f] >> 8);  // the face's parameter set (types 5-12)
  if (t == 1) {  // farfield
    const double us_n = 0.0, eta_pf = eta_face(p, i.T, i.u);
    const double uit = dot(i.u, n) - us_n;
    const double ai = sqrt(p.gamma * p.Rt * i.T);
    const double a0 = sqrt(p.gamma * p.Rt * p.T0);
    Q5 r0;
    r0.T = p.T0;
    r0.u = {p.u0x, p.u0y, p.u0z};
    r0.pg = p.pg0;
    if (uit > 0) {
      if (uit > ai) o = i;
      else cv_outflow(o, i, r0, n, us_n, p.Pamb, p.gamma, p.Rt, eta_pf);
    } else {
      if (-uit > a0) o = r0;
      else cv_inflow(o, i, r0, n, us_n, p.Pamb, p.gamma, p.Rt, eta_pf);
    }
  } else if (t == 2) {  // viscous wall, uwall = us = 0
    o.u = {0.0, 0.0, 0.0};
    o.pg = i.pg;
    o.T = p.Twall;
  } else if (t == 3) {  // symmetry
    const double uit = dot(i.u, n);
    o.u = {i.u.x - uit * n.x, i.u.y - uit * n.y, i.u.z - uit * n.z};
    const double rho = (i.pg + p.Pamb) / (p.Rt * i.T);
    const double U = sqrt(dot(i.u, i.u));
    o.pg = i.pg + 0.5 * rho * U * uit;
    o.T = i.T;
  } else if (t == 5) {  // reflecting, src/bcImpermeable.loci:102-117
    reflecting_state(o, i, (i.pg + p.Pamb) / (p.Rt * i.T), n, 0.0, p.Pamb);
  } else if (t == 6 || t == 7) {  // inflow (src/bcInflow.loci:377-402), supersonic inflow (:270-278)
    Q5 r0;
    r0.T = bs[0];
    r0.pg = bs[1];
    r0.u = {bs[2], bs[3], bs[4]};
    o = r0;
    const double uit = dot(i.u, n) - 0.0;
    const double a0 = sqrt(p.gamma * p.Rt * r0.T);
    if (t == 6 && !(uit > 0) && !(-uit > a0)) cv_inflow(o, i, r0, n, 0.0, p.Pamb, p.gamma, p.Rt, eta_face(p, i.T, i.u));
  } else if (t == 8) {  // outflow with a fixed pressure, src/bcOutflow.loci:206-240
    const double pref = bs[1];
    const double uit = dot(i.u, n) - 0.0;
    const double ai = sqrt(p.gamma * p.Rt * i.T);
    const double rhoi = (i.pg + p.Pamb) / (p.Rt * i.T);
    if (uit >= 0) {
      o = i;
      if (!(uit > ai)) {
        o.pg = pref;
        const double pa = o.pg + p.Pamb;
        o.T = pa / (p.Rt * rhoi);
      }
    } else {  // reversed flow
      o.u = {0.0 * n.x, 0.0 * n.y, 0.0 * n.z};
      o.pg = i.pg - .05 * (i.pg + p.Pamb);
      o.T = i.T;
    }
  } else if (t == 9) {  // supersonic outflow, src/bcOutflow.loci:127-141
    o = i;
    const double ai = sqrt(p.gamma * p.Rt * i.T);
    const double ut = dot(n, i.u) - 0.0;
    if (ut < ai) o.pg -= (i.pg + p.Pamb) * .1;
  } else if (t == 10) {  // extrapolate, src/bcOutflow.loci:84-94
    o = i;
  } else if (t == 11) {  // isentropic inflow, src/bcIsentropic.loci:146-166
    const V3 d = {-1. * n.x, -1. * n.y, -1. * n.z};
    p0t0_state(o, bs[6], bs[5], p.gamma, p.Rt, i, d, p.Pamb);
  } else {  // fixed mass, src/bcFixedMass.loci:140-153
    const V3 d = {-1. * n.x, -1. * n.y, -1. * n.z};
    fixed_mass_state(o, i.pg + p.Pamb, i.T, bs[7], bs[6], p.gamma, p.Rt, d, n, p.Pamb);
  }
  const int nb = g.nb;
  qf[b] = o.T;
  qf[nb + b] = o.u.x;
  qf[2 * nb + b] = o.u.y;
  qf[3 * nb + b] = o.u.z;
  qf[4 * nb + b] = o.pg;
  if (p.turb == 2) {
    // k_f, w_f: inflow/outflow switch on u_f.n (turbulence/KOmega/sst_viscous.loci:131-144),
    // Menter's wall values k = 0, omega = 60 mu/(rho_f beta1 d^2) with the
    // cell's mu and wall distance (:236-243), symmetry copies the cell
    const double ki = q[5 * nc + c], wi = q[6 * nc + c];
    double kb = ki, wb = wi;
    if (t == 2) {
      const double rho_f = (o.pg + p.Pamb) / (p.Rt * o.T);
      const double d = g.dist[c];
      kb = 0.0;
      wb = 10.0 * 6.0 * suth_mu(p, i.T) / (rho_f * 0.075 * pow(d, 2.0));
    } else if (t == 1) {
      kb = p.k_bc;
      wb = p.w_bc;
      if (dot(o.u, n) >= -0.0) { kb = ki; wb = wi; }
    }
    qf[5 * nb + b] = kb;
    qf[6 * nb + b] = wb;
  } else if (p.turb) {  // nuTilde_f: turbulence/SA/SA_turb.loci:181-183, src/solverScalarTransport.loci:423-463
    const double Xi = q[5 * nc + c];
    double Xb = Xi;
    if (t == 2) Xb = 0.0;
    else if (t == 1) Xb = (dot(n, o.u) - 0.0 >= -0.0) ? Xi : p.nut_bc;
    qf[5 * nb + b] = Xb;
  }
}

// neighbour across gather entry e of cell c: position and primitive values
template <int NQ>
__device__ __forceinline__ void stencil(const DevGrid &g, const double *__restrict__ q,
                                        const double *__restrict__ qf, int e, double *x, double *v) {
  const int nc = g.nc;
  const int f = e < 0 ? ~e : e;
  if (f < g.nfi) {
    const int o = e < 0 ? g.frc[f] : g.fcl[f];
    x[0] = g.ccx[o]; x[1] = g.ccy[o]; x[2] = g.ccz[o];
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = q[k * nc + o];
  } else if (f < g.bnd_begin) {
    const int o = g.frc[f];
    const int pi = f - g.nfi;
    x[0] = g.ccx[o] + g.psx[pi]; x[1] = g.ccy[o] + g.psy[pi]; x[2] = g.ccz[o] + g.psz[pi];
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = q[k * nc + o];
  } else {
    const int b = f - g.bnd_begin;
    x[0] = g.fcx[f]; x[1] = g.fcy[f]; x[2] = g.fcz[f];
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = qf[k * g.nb + b];
  }
}

// ---------------------------------------------------------------------------
// least-squares gradient and limiter per cell (Loci FVM grads/limiters;
// textbook unweighted least squares, Venkatakrishnan / Barth limiters)
// ---------------------------------------------------------------------------
template <int NQ>
__global__ void __launch_bounds__(TPB) k_grad(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ qf,
                       double *__restrict__ grad, double *__restrict__ phi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc;
  const double xc0 = g.ccx[c], xc1 = g.ccy[c], xc2 = g.ccz[c];
  double qc[NQ], b[NQ][3], qmin[NQ], qmax[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    qc[k] = q[k * nc + c];
    b[k][0] = b[k][1] = b[k][2] = 0.0;
    qmin[k] = qmax[k] = qc[k];
  }
  const int j0 = g.g_off[c], j1 = g.g_off[c + 1];
  for (int j = j0; j < j1; ++j) {
    double x[3], v[NQ];
    stencil<NQ>(g, q, qf, g.g_ent[j], x, v);
    const double dx = x[0] - xc0, dy = x[1] - xc1, dz = x[2] - xc2;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const double dq = v[k] - qc[k];
      b[k][0] += dx * dq;
      b[k][1] += dy * dq;
      b[k][2] += dz * dq;
      qmin[k] = dmin(qmin[k], v[k]);
      qmax[k] = dmax(qmax[k], v[k]);
    }
  }
  const double A0 = g.lsq[c], A1 = g.lsq[nc + c], A2 = g.lsq[2 * nc + c];
  const double A3 = g.lsq[3 * nc + c], A4 = g.lsq[4 * nc + c], A5 = g.lsq[5 * nc + c];
  double gr[NQ][3];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    gr[k][0] = A0 * b[k][0] + A1 * b[k][1] + A2 * b[k][2];
    gr[k][1] = A1 * b[k][0] + A3 * b[k][1] + A4 * b[k][2];
    gr[k][2] = A2 * b[k][0] + A4 * b[k][1] + A5 * b[k][2];
    grad[(3 * k) * nc + c] = gr[k][0];
    grad[(3 * k + 1) * nc + c] = gr[k][1];
    grad[(3 * k + 2) * nc + c] = gr[k][2];
  }
  double ph[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) ph[k] = 1.0;
  if (p.limiter == 2) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) ph[k] = 0.0;
  } else if (p.limiter == 1 || p.limiter == 3) {
    const double eps2 = p.Kl * p.Kl * p.Kl * g.vol[c];
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      const int f = e < 0 ? ~e : e;
      const double dx = g.fcx[f] - xc0, dy = g.fcy[f] - xc1, dz = g.fcz[f] - xc2;
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const double qdif = gr[k][0] * dx + gr[k][1] * dy + gr[k][2] * dz;
        double lim;
        if (p.limiter == 1) {
          const double delp = (qdif >= 0.0) ? qmax[k] - qc[k] : qmin[k] - qc[k];
          lim = venkat(delp, qdif, eps2);
        } else {
          if (qdif > 0) lim = dmin(1.0, (qmax[k] - qc[k]) / qdif);
          else if (qdif < 0) lim = dmin(1.0, (qmin[k] - qc[k]) / qdif);
          else lim = 1.0;
        }
        ph[k] = dmin(ph[k], lim);
      }
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) ph[k] = dmax(0.0, dmin(1.0, ph[k]));
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) phi[k * nc + c] = ph[k];
}

// limited reconstruction with the floors of leftsP(temperature,Zero) and
// leftsP(gagePressure,minPg)
__device__ __forceinline__ Q5 recon(const DevGrid &g, const Prm &p, const double *__restrict__ q,
                                    const double *__restrict__ grad, const double *__restrict__ phi,
                                    int c, double xc0, double xc1, double xc2, double xf0, double xf1,
                                    double xf2) {
  const int nc = g.nc;
  const double d0 = xf0 - xc0, d1 = xf1 - xc1, d2 = xf2 - xc2;
  double o[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double gd = grad[(3 * k) * nc + c] * d0 + grad[(3 * k + 1) * nc + c] * d1 +
                      grad[(3 * k + 2) * nc + c] * d2;
    o[k] = q[k * nc + c] + phi[k * nc + c] * gd;
  }
  Q5 r;
  r.T = dmax(o[0], 0.0);
  r.u = {o[1], o[2], o[3]};
  r.pg = dmax(o[4], p.minPg);
  return r;
}

// face gradient of variable k: interior/periodic faces average the two cell
// gradients and correct along the centre line; boundary faces correct the
// cell gradient along x_f - x_c (Loci gradv3d_f / grads_f)
__device__ __forceinline__ void face_grad_int(const DevGrid &g, const double *__restrict__ q,
                                              const double *__restrict__ grad, int c, int o, int k,
                                              const double *d, double dd, double *gout) {
  const int nc = g.nc;
  double ga[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) ga[j] = 0.5 * (grad[(3 * k + j) * nc + c] + grad[(3 * k + j) * nc + o]);
  const double corr = (q[k * nc + o] - q[k * nc + c] - (ga[0] * d[0] + ga[1] * d[1] + ga[2] * d[2])) / dd;
#pragma unroll
  for (int j = 0; j < 3; ++j) gout[j] = ga[j] + corr * d[j];
}
__device__ __forceinline__ void face_grad_bnd(const DevGrid &g, const double *__restrict__ q,
                                              const double *__restrict__ grad, double vf, int c, int k,
                                              const double *d, double dd, double *gout) {
  const int nc = g.nc;
  double ga[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) ga[j] = grad[(3 * k + j) * nc + c];
  const double corr = (vf - q[k * nc + c] - (ga[0] * d[0] + ga[1] * d[1] + ga[2] * d[2])) / dd;
#pragma unroll
  for (int j = 0; j < 3; ++j) gout[j] = ga[j] + corr * d[j];
}

// viscous face quantities (src/fluxViscous.loci:121-133,145-151,184-197,217-238,272-297)
// tmu_f (src/fluxViscous.loci:192-201): volume-weighted interior average,
// the interior cell on boundary faces, zero on no-slip walls
__device__ __forceinline__ double tmu_face(const DevGrid &g, const double *__restrict__ tmu, int f) {
  const int c = g.fcl[f];
  if (f < g.bnd_begin) {
    const int o = g.frc[f];
    const double vl = g.vol[c], vr = g.vol[o];
    return (tmu[o] * vl + tmu[c] * vr) / (vr + vl);
  }
  return (g.fbt[f] & 0xFF) == 2 ? 0.0 : tmu[c];
}
// muTotal_f, kconductTotal (src/fluxViscous.loci:182-189,275-278)
__device__ __forceinline__ void add_tmu(const DevGrid &g, const Prm &p, const double *__restrict__ tmu, int f,
                                        double *mu_f, double *k_f) {
  if (!p.turb) return;
  const double tmf = tmu_face(g, tmu, f);
  *k_f = *k_f + tmf * p.Cp / p.Prt;
  *mu_f = *mu_f + tmf;
}

__device__ inline void viscous_face(const DevGrid &g, const Prm &p, const double *__restrict__ q,
                                    const double *__restrict__ qf, const double *__restrict__ grad,
                                    const double *__restrict__ tmu, int f, double *mu_f, double *k_f, V3 *uf,
                                    double *vf, bool need_flux) {
  const int nc = g.nc;
  const int c = g.fcl[f];
  const V3 an = {g.fnx[f], g.fny[f], g.fnz[f]};
  const double sada = g.farea[f];
  const double xl0 = g.ccx[c], xl1 = g.ccy[c], xl2 = g.ccz[c];
  double gu[3][3], gT3[3];
  if (f < g.bnd_begin) {
    const int o = g.frc[f];
    double xr0 = g.ccx[o], xr1 = g.ccy[o], xr2 = g.ccz[o];
    if (f >= g.nfi) {
      const int pi = f - g.nfi;
      xr0 += g.psx[pi]; xr1 += g.psy[pi]; xr2 += g.psz[pi];
    }
    const double vl = g.vol[c], vr = g.vol[o];
    const double mul = suth_mu(p, q[c]), mur = suth_mu(p, q[o]);
    const double kl = suth_k(p, q[c]), kr = suth_k(p, q[o]);
    *mu_f = (vl * mur + vr * mul) / (vl + vr);
    *k_f = (vl * kr + vr * kl) / (vl + vr);
    add_tmu(g, p, tmu, f, mu_f, k_f);
    const double factor = 1. / (vl + vr);
    uf->x = factor * (vl * q[nc + o] + vr * q[nc + c]);
    uf->y = factor * (vl * q[2 * nc + o] + vr * q[2 * nc + c]);
    uf->z = factor * (vl * q[3 * nc + o] + vr * q[3 * nc + c]);
    if (!need_flux) return;
    const double d[3] = {xr0 - xl0, xr1 - xl1, xr2 - xl2};
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    for (int i = 0; i < 3; ++i) face_grad_int(g, q, grad, c, o, 1 + i, d, dd, gu[i]);
    face_grad_int(g, q, grad, c, o, 0, d, dd, gT3);
  } else {
    const int b = f - g.bnd_begin;
    const int nb = g.nb;
    const double Tf = qf[b];
    *mu_f = suth_mu(p, Tf);
    *k_f = suth_k(p, Tf);
    add_tmu(g, p, tmu, f, mu_f, k_f);
    uf->x = qf[nb + b];
    uf->y = qf[2 * nb + b];
    uf->z = qf[3 * nb + b];
    if (!need_flux) return;
    if ((g.fbt[f] & 0xFF) == 3) {
      vf[0] = vf[1] = vf[2] = vf[3] = 0.0;
      return;
    }
    const double d[3] = {g.fcx[f] - xl0, g.fcy[f] - xl1, g.fcz[f] - xl2};
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    for (int i = 0; i < 3; ++i) face_grad_bnd(g, q, grad, qf[(1 + i) * nb + b], c, 1 + i, d, dd, gu[i]);
    face_grad_bnd(g, q, grad, qf[b], c, 0, d, dd, gT3);
  }
  const V3 gT = {gT3[0], gT3[1], gT3[2]};
  viscous_flux(vf, gu, gT, *mu_f, *k_f, *uf, an, sada);
}

// ---------------------------------------------------------------------------
// low-Mach preconditioning, src/fluxInviscid.loci:35-142.  Eta_l per cell from
// Mref = |u|/a raised by the pressure-jump and laminar viscous limits of every
// interior and periodic face (a periodic face counts on its cl side, with the
// partner's state as its ghost, src/solverPeriodic.loci:39-70), then two
// passes of the maximum over the face neighbours (Eta_ll, Eta_p).  Maxima
// are exact, so the gather order is free.
// ---------------------------------------------------------------------------
__global__ void k_eta_l(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ tmu,
                        double *__restrict__ el) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc;
  const double T = q[c], pg = q[4 * nc + c];
  const V3 U = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
  const double rho = (pg + p.Pamb) / (p.Rt * T);
  const double a = sqrt(p.gamma * p.Rt * T);
  const double mu = p.viscous ? suth_mu(p, T) : 0.0;
  double mref = sqrt(dot(U, U)) / a;
  for (int j = g.g_off[c]; j < g.g_off[c + 1]; ++j) {
    const int e = g.g_ent[j];
    const int f = e < 0 ? ~e : e;
    if (f >= g.bnd_begin) continue;
    const int o = e < 0 ? g.frc[f] : g.fcl[f];
    const double To = q[o], pgo = q[4 * nc + o];
    const double dP = fabs(pg - pgo);
    mref = dmax(mref, p.PLimPC * dP / (rho * a * a));
    if (p.viscous) {
      const double area = dmax(g.farea[f], 1e-30);
      const double rhoo = (pgo + p.Pamb) / (p.Rt * To);
      const double rhom = 0.5 * (rho + rhoo);
      double mum = dmax(mu, suth_mu(p, To));
      if (p.turb) {  // src/fluxInviscid.loci:81-96: the cl side joins max(tmu_cl, tmu_cl)
        const double tl = e < 0 ? tmu[c] : tmu[o], tr = e < 0 ? tmu[o] : tmu[c];
        mum = mum + (e < 0 ? dmax(tl, tl) : dmax(tl, tr));
      }
      const double nu = mum / rhom;
      mref = dmax(mref, nu / ((g.vol[c] / area) * a));
    }
  }
  const double M2 = 2. * dmin(0.25, mref * mref);
  el[c] = dmax(M2 / (1. - M2), p.Mlim2);
}

__global__ void k_eta_max(DevGrid g, const double *__restrict__ src, double *__restrict__ dst) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  double b = src[c];
  for (int j = g.g_off[c]; j < g.g_off[c + 1]; ++j) {
    const int e = g.g_ent[j];
    const int f = e < 0 ? ~e : e;
    if (f >= g.bnd_begin) continue;
    b = dmax(b, src[e < 0 ? g.frc[f] : g.fcl[f]]);
  }
  dst[c] = b;
}

// ---------------------------------------------------------------------------
// face fluxes: iflux (src/fluxInviscid.loci:1170-1188, src/bcInflow.loci:342-355,
// src/bcViscousWall.loci:216-229, src/bcImpermeable.loci:66-73), vflux and
// maxev_f (src/solverSetup.loci:83-105)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_face_flux(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ qf,
                            const double *__restrict__ grad, const double *__restrict__ phi,
                            const double *__restrict__ etap, const double *__restrict__ tmu,
                            double *__restrict__ iflux, double *__restrict__ vflux, double *__restrict__ maxev) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nf) return;
  const int nc = g.nc, nf = g.nf, nb = g.nb;
  const V3 an = {g.fnx[f], g.fny[f], g.fnz[f]};
  const double area = g.farea[f];
  const double xf0 = g.fcx[f], xf1 = g.fcy[f], xf2 = g.fcz[f];
  const int c = g.fcl[f];
  const double us_n = 0.0;
  // interior/periodic: max over the two cells (src/fluxInviscid.loci:1187);
  // boundary: the interior cell
  double eta = 1.0;
  if (p.pre) eta = f < g.bnd_begin ? dmax(etap[c], etap[g.frc[f]]) : etap[c];
  const Q5 L = recon(g, p, q, grad, phi, c, g.ccx[c], g.ccy[c], g.ccz[c], xf0, xf1, xf2);
  double fl[5];
  double mev;
  const double Eta_P = eta + 1., Eta_M = eta - 1.;
  if (f < g.bnd_begin) {
    const int o = g.frc[f];
    double xr0 = g.ccx[o], xr1 = g.ccy[o], xr2 = g.ccz[o];
    if (f >= g.nfi) {
      const int pi = f - g.nfi;
      xr0 += g.psx[pi]; xr1 += g.psy[pi]; xr2 += g.psz[pi];
    }
    const Q5 R = recon(g, p, q, grad, phi, o, xr0, xr1, xr2, xf0, xf1, xf2);
    if (p.flux == 1) roe_flux(fl, L, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, eta);
    else hllc_flux(fl, L, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, eta);
    const V3 ul = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
    const V3 ur = {q[nc + o], q[2 * nc + o], q[3 * nc + o]};
    const double utl = dot(ul, an) - us_n, utr = dot(ur, an) - us_n;
    const double al = sqrt(p.gamma * p.Rt * q[c]), ar = sqrt(p.gamma * p.Rt * q[o]);
    const double sigmal = sqrt(utl * utl * Eta_M * Eta_M + 4. * eta * al * al);
    const double sigmar = sqrt(utr * utr * Eta_M * Eta_M + 4. * eta * ar * ar);
    const double mevl = 0.5 * dmax(fabs(utl * Eta_P - sigmal), fabs(utl * Eta_P + sigmal));
    const double mevr = 0.5 * dmax(fabs(utr * Eta_P - sigmar), fabs(utr * Eta_P + sigmar));
    mev = dmax(mevl, mevr);
  } else {
    const int b = f - g.bnd_begin;
    const int t = g.fbt[f] & 0xFF;
    Q5 R;
    R.T = qf[b];
    R.u = {qf[nb + b], qf[2 * nb + b], qf[3 * nb + b]};
    R.pg = qf[4 * nb + b];
    if (t == 1 || (t >= 6 && t <= 10)) {  // Riemann flux against the face state
      // supersonic inflow uses the face's Eta_pf (src/bcInflow.loci:280-289)
      const double eb = t == 7 ? eta_face(p, q[c], V3{q[nc + c], q[2 * nc + c], q[3 * nc + c]}) : eta;
      if (p.flux == 1) roe_flux(fl, L, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, eb);
      else hllc_flux(fl, L, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, eb);
    } else if (t == 5) {  // reflecting: against the mirrored left state, no mass flux
      const Q5 Rm = mirror(L, an, us_n);
      if (p.flux == 1) roe_flux(fl, L, Rm, an, area, p.Pamb, p.Rt, p.gamma, us_n, eta);
      else hllc_flux(fl, L, Rm, an, area, p.Pamb, p.Rt, p.gamma, us_n, eta);
      fl[0] = 0;
    } else if (t == 11 || t == 12) {  // the face state's Euler flux
      euler_flux(fl, R, an, area, p.Pamb, p.Rt, p.gamma, us_n);
    } else if (t == 2) {
      const double pg = R.pg;
      fl[0] = 0;
      fl[1] = pg * (an.x * area);
      fl[2] = pg * (an.y * area);
      fl[3] = pg * (an.z * area);
      const double pw = pg + p.Pamb;
      fl[4] = pw * area * us_n;
    } else {
      const double pg = R.pg;
      fl[0] = 0;
      const double moment = pg * area;
      fl[1] = an.x * moment;
      fl[2] = an.y * moment;
      fl[3] = an.z * moment;
      fl[4] = area * pg * dot(R.u, an);
    }
    const double utf = dot(R.u, an) - us_n;
    const double af = sqrt(p.gamma * p.Rt * R.T);
    const double sigmaf = sqrt(utf * utf * Eta_M * Eta_M + 4. * eta * af * af);
    mev = 0.5 * dmax(fabs(utf * Eta_P - sigmaf), fabs(utf * Eta_P + sigmaf));
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) iflux[k * nf + f] = fl[k];
  maxev[f] = mev;
  if (p.viscous) {
    double mu_f, k_f, vf[4];
    V3 uf;
    viscous_face(g, p, q, qf, grad, tmu, f, &mu_f, &k_f, &uf, vf, true);
#pragma unroll
    for (int k = 0; k < 4; ++k) vflux[k * nf + f] = vf[k];
  }
}

// ---------------------------------------------------------------------------
// residual gather (src/fluxInviscid.loci:1253-1268, src/fluxViscous.loci:299-310)
// and the residual norm terms (src/solverResidual.loci:74-112)
// ---------------------------------------------------------------------------
struct NormAcc {
  double rt, et, mt;
  double rmax, emax, mmax;
  int rid, eid, mid;  // caller cell index of the maxima (-1 none)
  int rcell, ecell;   // stored cell of the maxima
};

__device__ __forceinline__ bool better(double a, int ia, double b, int ib) {
  // a replaces b when |b| < |a|; on equal magnitude the lower caller index
  // wins, which is what a join in caller cell order produces
  const double fa = fabs(a), fb = fabs(b);
  if (fb < fa) return true;
  if (fa == fb) return ia >= 0 && (ib < 0 || ia < ib);
  return false;
}

__device__ __forceinline__ void norm_join(NormAcc &r, const NormAcc &s) {
  r.rt += s.rt;
  r.et += s.et;
  r.mt += s.mt;
  if (better(s.emax, s.eid, r.emax, r.eid)) { r.emax = s.emax; r.eid = s.eid; r.ecell = s.ecell; }
  if (better(s.rmax, s.rid, r.rmax, r.rid)) { r.rmax = s.rmax; r.rid = s.rid; r.rcell = s.rcell; }
  if (better(s.mmax, s.mid, r.mmax, r.mid)) { r.mmax = s.mmax; r.mid = s.mid; }
}

static constexpr int NPART = 12;  // doubles per partial record

__device__ void block_norms(NormAcc a, double *__restrict__ part) {
  __shared__ NormAcc sh[TPB];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) norm_join(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const NormAcc &r = sh[0];
    double *o = part + (size_t)blockIdx.x * NPART;
    o[0] = r.rt; o[1] = r.et; o[2] = r.mt; o[3] = r.rmax; o[4] = r.emax; o[5] = r.mmax;
    o[6] = r.rid; o[7] = r.eid; o[8] = r.mid; o[9] = r.rcell; o[10] = r.ecell; o[11] = 0;
  }
}

// norm terms of one cell from its residual s (src/solverResidual.loci:73-111)
__device__ __forceinline__ void cell_norms(const DevGrid &g, double vtot, int c, const double *s, NormAcc &a) {
  const double vol = g.vol[c];
  const double rvol = 1. / vol;
  const double weight = vol / vtot;
  const int id = g.old_id[c];
  const double r0 = rvol * s[0];
  if (fabs(0.0) < fabs(r0)) { a.rmax = r0; a.rid = id; a.rcell = c; }
  double rt = 0;
  rt += r0 * r0;
  rt *= weight;
  double mt = 0, mmax = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double v = rvol * s[1 + i];
    if (fabs(mmax) < fabs(v)) mmax = v;
    mt += v * v;
  }
  mt *= weight / 3.0;
  const double ve = rvol * s[4];
  a.rt = rt;
  a.mt = mt;
  a.et = weight * ve * ve;
  a.emax = ve; a.eid = id; a.ecell = c;
  a.mmax = mmax; a.mid = id;
}

__global__ void k_gather_src(DevGrid g, int viscous, double vtot, const double *__restrict__ iflux,
                             const double *__restrict__ vflux, double *__restrict__ src,
                             double *__restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  NormAcc a = {0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1};
  if (c < g.nc) {
    const int nc = g.nc, nf = g.nf;
    double s[5] = {0, 0, 0, 0, 0};
    const int j0 = g.g_off[c], j1 = g.g_off[c + 1];
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      if (e < 0) {
        const int f = ~e;
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] -= iflux[k * nf + f];
        if (viscous) {
#pragma unroll
          for (int k = 0; k < 4; ++k) s[k + 1] += vflux[k * nf + f];
        }
      } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] += iflux[k * nf + e];
        if (viscous) {
#pragma unroll
          for (int k = 0; k < 4; ++k) s[k + 1] -= vflux[k * nf + e];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) src[k * nc + c] = s[k];
    cell_norms(g, vtot, c, s, a);
  }
  block_norms(a, part);
}

// residual norm partials of a right-hand side held as [5][nc] (the BDF2 rhs,
// src/solverResidual.loci:113-150)
__global__ void k_rhs_norms(DevGrid g, double vtot, const double *__restrict__ rhs, double *__restrict__ part) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  NormAcc a = {0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1};
  if (c < g.nc) {
    double s[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] = rhs[(size_t)k * g.nc + c];
    cell_norms(g, vtot, c, s, a);
  }
  block_norms(a, part);
}

__global__ void k_norm_final(DevGrid g, int nparts, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ NormAcc sh[TPB];
  NormAcc a = {0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1};
  // each thread folds a contiguous slice in order, then a fixed tree
  const int per = (nparts + TPB - 1) / TPB;
  const int i0 = threadIdx.x * per, i1 = min(nparts, i0 + per);
  for (int i = i0; i < i1; ++i) {
    const double *o = part + (size_t)i * NPART;
    NormAcc s = {o[0], o[1], o[2], o[3], o[4], o[5], (int)o[6], (int)o[7], (int)o[8], (int)o[9], (int)o[10]};
    norm_join(a, s);
  }
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) norm_join(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const NormAcc &r = sh[0];
    out[0] = r.rt; out[1] = r.et; out[2] = r.mt;
    out[3] = r.rmax; out[4] = r.emax; out[5] = r.mmax;
    const int rc = r.rcell, ec = r.ecell;
    out[6] = (rc >= 0 && r.rmax != 0.0) ? g.ccx[rc] : 0.0;
    out[7] = (rc >= 0 && r.rmax != 0.0) ? g.ccy[rc] : 0.0;
    out[8] = (rc >= 0 && r.rmax != 0.0) ? g.ccz[rc] : 0.0;
    out[9] = (ec >= 0 && r.emax != 0.0) ? g.ccx[ec] : 0.0;
    out[10] = (ec >= 0 && r.emax != 0.0) ? g.ccy[ec] : 0.0;
    out[11] = (ec >= 0 && r.emax != 0.0) ? g.ccz[ec] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// face Jacobians: fjp/fjm (src/fluxInviscid.loci:1190-1221 + viscous
// src/fluxViscous.loci:315-403) written straight into the block rows;
// boundary faces: their whole srcJ contribution (src/bcInflow.loci:357-375,
// src/bcViscousWall.loci:230-240, src/bcImpermeable.loci:75-100,
// src/fluxViscous.loci:405-448)
// ---------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ void store_blk(F *__restrict__ dst, const F *blk, bool neg) {
#pragma unroll
  for (int i = 0; i < 25; ++i) dst[i] = neg ? (F)(-blk[i]) : blk[i];
}

template <typename F>
__global__ void __launch_bounds__(TPB) k_face_jac(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ qf,
                           const double *__restrict__ grad, const double *__restrict__ phi,
                           const double *__restrict__ vflux, const double *__restrict__ etap,
                           const double *__restrict__ tmu, F *__restrict__ off, F *__restrict__ pblk,
                           F *__restrict__ bblk) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nf) return;
  const int nc = g.nc, nf = g.nf, nb = g.nb;
  const V3 an = {g.fnx[f], g.fny[f], g.fnz[f]};
  const double area = g.farea[f];
  const double xf0 = g.fcx[f], xf1 = g.fcy[f], xf2 = g.fcz[f];
  const int c = g.fcl[f];
  const double xl0 = g.ccx[c], xl1 = g.ccy[c], xl2 = g.ccz[c];
  const double us_n = 0.0;
  double eta = 1.0;  // src/fluxInviscid.loci:1204,1220; src/bcInflow.loci:369
  if (p.pre) eta = f < g.bnd_begin ? dmax(etap[c], etap[g.frc[f]]) : etap[c];
  const Q5 L = recon(g, p, q, grad, phi, c, xl0, xl1, xl2, xf0, xf1, xf2);
  F blk[25];
  if (f < g.bnd_begin) {
    const int o = g.frc[f];
    double xr0 = g.ccx[o], xr1 = g.ccy[o], xr2 = g.ccz[o];
    const bool per = f >= g.nfi;
    if (per) {
      const int pi = f - g.nfi;
      xr0 += g.psx[pi]; xr1 += g.psy[pi]; xr2 += g.psz[pi];
    }
    const Q5 R = recon(g, p, q, grad, phi, o, xr0, xr1, xr2, xf0, xf1, xf2);
    const RoeAvg ra = roe_average(L, R, an, p.Pamb, p.Rt, p.gamma, us_n, eta);
    double mu_f = 0, k_f = 0, vf[4] = {0, 0, 0, 0}, dist = 1;
    V3 uf = {0, 0, 0};
    if (p.viscous) {
      viscous_face(g, p, q, qf, grad, tmu, f, &mu_f, &k_f, &uf, vf, false);
#pragma unroll
      for (int k = 0; k < 4; ++k) vf[k] = vflux[k * nf + f];
      const V3 dv = {xr0 - xl0, xr1 - xl1, xr2 - xl2};
      dist = dot(an, dv);
    }
    // fjp
#pragma unroll
    for (int i = 0; i < 25; ++i) blk[i] = (F)0;
    roe_jacobian<F>(blk, ra, an, area, p.gamma, p.Rt, 1);
    if (p.viscous) viscous_jacobian<F>(blk, 1.0, an, area, dist, mu_f, k_f, uf, vf);
    if (p.tdamp > 0) damping_jacobian<F>(blk, 1.0, L, an, area, p.Pamb, p.Rt, p.gamma, us_n, p.tdamp);
    if (!per) store_blk(off + (size_t)g.slot_cr[f] * 25, blk, true);
    else store_blk(pblk + (size_t)(f - g.nfi) * 25, blk, false);
    // fjm
#pragma unroll
    for (int i = 0; i < 25; ++i) blk[i] = (F)0;
    roe_jacobian<F>(blk, ra, an, area, p.gamma, p.Rt, -1);
    if (p.viscous) viscous_jacobian<F>(blk, -1.0, an, area, dist, mu_f, k_f, uf, vf);
    if (p.tdamp > 0) damping_jacobian<F>(blk, -1.0, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, p.tdamp);
    store_blk(off + (size_t)g.slot_cl[f] * 25, blk, false);
  } else {
    const int b = f - g.bnd_begin;
    const int t = g.fbt[f] & 0xFF;
#pragma unroll
    for (int i = 0; i < 25; ++i) blk[i] = (F)0;
    Q5 R;
    R.T = qf[b];
    R.u = {qf[nb + b], qf[2 * nb + b], qf[3 * nb + b]};
    R.pg = qf[4 * nb + b];
    if (t == 1 || (t >= 6 && t <= 10)) {  // srcJ -= fjp(left, face state)
      const double eb = t == 7 ? eta_face(p, q[c], V3{q[nc + c], q[2 * nc + c], q[3 * nc + c]}) : eta;
      F tmp[25];
#pragma unroll
      for (int i = 0; i < 25; ++i) tmp[i] = (F)0;
      const RoeAvg ra = roe_average(L, R, an, p.Pamb, p.Rt, p.gamma, us_n, eb);
      roe_jacobian<F>(tmp, ra, an, area, p.gamma, p.Rt, 1);
#pragma unroll
      for (int i = 0; i < 25; ++i) blk[i] -= tmp[i];
    } else if (t == 5) {
      reflecting_jacobian<F>(blk, L, an, area, p, us_n, eta);
    } else if (t == 11 || t == 12) {
      Q5 ci;
      ci.T = q[c];
      ci.u = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
      ci.pg = q[4 * nc + c];
      bc_state_jacobian<F>(blk, t, ci, p.bcs + 8 * (g.fbt[f] >> 8), an, area, p, us_n);
    } else if (t == 2) {
      acc(blk[1 + 5 * 4], -(an.x * area));
      acc(blk[2 + 5 * 4], -(an.y * area));
      acc(blk[3 + 5 * 4], -(an.z * area));
      acc(blk[4 + 5 * 4], -(area * dot(R.u, an)));
    } else {
      const double pp = q[4 * nc + c] + p.Pamb;
      const double T = q[c];
      const double rho = pp / (p.Rt * T);
      const double drdt = -rho / T;
      const double drdp = rho / pp;
      const double ui[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
      const double n3[3] = {an.x, an.y, an.z};
      const double U = sqrt(ui[0] * ui[0] + ui[1] * ui[1] + ui[2] * ui[2]);
      const double uit = (n3[0] * ui[0] + n3[1] * ui[1] + n3[2] * ui[2]) - us_n;
      const double coef1 = -0.5 * rho * area;
      const double coef2 = -0.5 * area;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double coefj = coef1 * (ui[j] * uit / (U + 1e-30) + n3[j] * U);
#pragma unroll
        for (int i = 0; i < 3; ++i) acc(blk[(1 + i) + 5 * (1 + j)], coefj * n3[i]);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double coefx = 0.5 * U * uit;
        acc(blk[(1 + i) + 5 * 0], coef2 * (coefx * drdt) * n3[i]);
        acc(blk[(1 + i) + 5 * 4], coef2 * (1.0 + coefx * drdp) * n3[i]);
      }
    }
    if (p.viscous) {
      double mu_f, k_f, vf[4];
      V3 uf;
      viscous_face(g, p, q, qf, grad, tmu, f, &mu_f, &k_f, &uf, vf, false);
#pragma unroll
      for (int k = 0; k < 4; ++k) vf[k] = vflux[k * nf + f];
      const V3 dv = {xf0 - xl0, xf1 - xl1, xf2 - xl2};
      viscous_jacobian<F>(blk, -1.0, an, area, dot(an, dv), mu_f, k_f, uf, vf);
    }
    store_blk(bblk + (size_t)b * 25, blk, false);
  }
}

// density, momentum and total energy per volume of a primitive state
__device__ __forceinline__ void cons5(const Prm &p, double T, double pg, const V3 &u, double *Q) {
  const double rho = (pg + p.Pamb) / (p.Rt * T);
  Q[0] = rho;
  Q[1] = rho * u.x;
  Q[2] = rho * u.y;
  Q[3] = rho * u.z;
  Q[4] = rho * (p.Rt * T / (p.gamma - 1.) + .5 * dot(u, u));
}

// KEC level n+1 from the iterate q* (src/solverTimestep.loci:366-385,
// src/solverNewton.loci:127-157): harmonic-mean temperature, biased pressure,
// density-weighted velocity; out = {T, u, v, w, pg, rho}
__device__ __forceinline__ void kec_np1(const Prm &p, double Ts, double pgs, const V3 &us, double Tn, double pgn,
                                        const V3 &un, double peps, double *out) {
  const double Tnp1 = Tn * Ts / (2. * Tn - Ts);
  const double pgnp1 = (2. * pgs - (1. - peps) * pgn) / (1 + peps);
  const double rhon = (pgn + p.Pamb) / (p.Rt * Tn);
  const double rhonp = dmax((pgnp1 + p.Pamb) / (p.Rt * Tnp1), 1e-15);
  const double srhonp = sqrt(rhonp), srhon = sqrt(rhon);
  const double w1 = (srhonp + srhon) / srhonp, w2 = srhon / srhonp;
  out[0] = Tnp1;
  out[1] = w1 * us.x - w2 * un.x;
  out[2] = w1 * us.y - w2 * un.y;
  out[3] = w1 * us.z - w2 * un.z;
  out[4] = pgnp1;
  out[5] = rhonp;
}

// KEC scalar level n+1 (and X_np1) from the iterate,
// src/solverScalarTransport.loci:367-395
__device__ __forceinline__ double kec_xnp1(double X, double Xn) {
  const double Xstar = dmax(X, 1e-15);
  double dXstar = dmax((Xstar - Xn) / Xstar, -1.95);
  if (dXstar < 0.0) dXstar += 0.25 * dXstar * dXstar;
  return Xstar * (1. + dXstar);
}

// TS: time scheme, 0 steady, 1 BDF2, 2 KEC
template <typename F, bool LS, int TS>
__global__ void __launch_bounds__(TPB) k_cell_assemble(DevGrid g, Prm p, const double *__restrict__ q,
                                const double *__restrict__ src, const double *__restrict__ maxev,
                                const F *__restrict__ off, const F *__restrict__ pblk,
                                const F *__restrict__ bblk, double *__restrict__ dtau_out,
                                F *__restrict__ lu, int *__restrict__ piv, F *__restrict__ Bout,
                                double *__restrict__ X, double *__restrict__ lu64,
                                const double *__restrict__ etap, const double *__restrict__ qn,
                                const double *__restrict__ qm1, double *__restrict__ rhs,
                                double *__restrict__ pbias) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc;
  const int j0 = g.g_off[c], j1 = g.g_off[c + 1];
  double meva = 0.0;
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    const int f = e < 0 ? ~e : e;
    meva += maxev[f] * g.farea[f];
  }
  const double vol = g.vol[c];
  const double cflpdt = 0.5 * meva / vol;
  double dtcfl = p.dtmax;
  if (p.cflmax > 0) dtcfl = dmin(dtcfl, p.cflmax / cflpdt);
  double dt = dtcfl;
  const double rvol = 1. / vol;
  const double pg = q[4 * nc + c];
  const double pp = pg + p.Pamb;
  const double T = q[c];
  const double Rt = p.Rt;
  const double rP = 1. / (Rt * T);
  const double rho = (pg + p.Pamb) * rP;
  const double rT = -rho / T;
  const double gm = p.gamma;
  const V3 U = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
  const double qq = 0.5 * dot(U, U);
  const double fT = rT * qq;
  const double fP = (1. / (gm - 1) + rP * qq);
  const double reta0 = 1. / (p.pre ? etap[c] : 1.0);  // src/solverLinearSystem.loci:195
  const double a2 = gm * Rt * T;
  const double pterm0 = (1. - reta0) * (gm - 1.) / a2;
  const double cterm = fP * rT * reta0 - fT * rP * reta0;
  const double q2 = 2. * qq;
  double s[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k] = src[k * nc + c];
  const double dsu = s[1] * U.x + s[2] * U.y + s[3] * U.z;
  const double dtemp = rvol * (s[0] * (fP * rT * reta0 + fT * pterm0) + (dsu - s[4] - q2 * s[0]) * (rT * (pterm0 + rP * reta0))) / (cterm * rT);
  dt = dmin(dt, p.urelax * T / (fabs(dtemp) + 1e-30));
  const double dp = rvol * (rT * (q2 * s[0] - dsu + s[4]) - fT * s[0]);
  dt = dmin(dt, p.urelax * pp / (fabs(dp) + 1e-30));
  dtau_out[c] = dt;

  F sJ[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) sJ[i] = (F)0;
  // upper faces (srcJ -= fjp) then lower faces (srcJ += fjm); the partner
  // slot of an upper face holds L = -fjp, of a lower face U = fjm
  const int m0 = g.m_off[c], m1 = g.m_off[c + 1];
  for (int j = m0; j < m1; ++j) {
    const int pt = g.m_part[j];
    if (pt >= 0) {
      const F *blk = off + (size_t)pt * 25;
#pragma unroll
      for (int i = 0; i < 25; ++i) sJ[i] += blk[i];
    } else {
      const F *blk = pblk + (size_t)(-1 - pt) * 25;
#pragma unroll
      for (int i = 0; i < 25; ++i) sJ[i] -= blk[i];
    }
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    const int f = e < 0 ? ~e : e;
    if (f >= g.bnd_begin) {
      const F *blk = bblk + (size_t)(f - g.bnd_begin) * 25;
#pragma unroll
      for (int i = 0; i < 25; ++i) sJ[i] += blk[i];
    }
  }
  F D[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) D[i] = -sJ[i];
  double coef = vol / dt;
  double reta = reta0, pterm = pterm0;
  if constexpr (TS == 1) {  // BDF2 fluid_D, src/solverLinearSystem.loci:384-420
    coef = 1.5 * vol / dt;
    const double pseudo_dt = p.pseudo_cflmax / cflpdt;
    const double w = 1. / (1. + pseudo_dt / dt);
    reta = 1. / (1. + w * ((p.pre ? etap[c] : 1.0) - 1.));
    pterm = (1. - reta) * (gm - 1.) / a2;
  }
  if constexpr (TS == 2) {
    // KEC: localPBias (src/solverTimestep.loci:346-351), fluid_D = -srcJ +
    // (V/dtau) dQ/dq dq/dq* (src/solverLinearSystem.loci:266-378), rhs
    // (src/solverTimestep.loci:366-396)
    const double cfl = cflpdt * p.dtmax;
    const double factor = dmin(1.0, dmax(1e-40, 0.1 * (cfl - 1.)));
    const double pb = dmax(p.pbias_ctl, pow(factor, 0.7));
    pbias[c] = pb;
    double dQ[5][5], dqs[5][5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
      for (int j = 0; j < 5; ++j) dQ[i][j] = dqs[i][j] = 0.0;
    dQ[0][0] = rT; dQ[1][0] = rT * U.x; dQ[2][0] = rT * U.y; dQ[3][0] = rT * U.z; dQ[4][0] = fT;
    dQ[0][4] = rP; dQ[1][4] = rP * U.x; dQ[2][4] = rP * U.y; dQ[3][4] = rP * U.z; dQ[4][4] = fP;
    dQ[1][1] = rho; dQ[4][1] = rho * U.x;
    dQ[2][2] = rho; dQ[4][2] = rho * U.y;
    dQ[3][3] = rho; dQ[4][3] = rho * U.z;
    const double Tn = qn[c], pgn = qn[4 * nc + c];
    const V3 un = {qn[nc + c], qn[2 * nc + c], qn[3 * nc + c]};
    double np1[6];
    kec_np1(p, T, pg, U, Tn, pgn, un, pb, np1);
    const double rhon = (pgn + p.Pamb) / (Rt * Tn);
    const double Pcoef = 2.0 / (1 + pb);
    const double srhonp1 = sqrt(np1[5]), srhon = sqrt(rhon);
    const double ucoef = (srhonp1 + srhon) / srhonp1;
    const double tfactor = 1. / (2 * Tn - T);
    const double dTdTs = 2. * Tn * Tn * tfactor * tfactor;
    const double drdt = -dTdTs * rho / T;
    const double drdp = Pcoef * rho / (pg + p.Pamb);
    dqs[0][0] = dTdTs;
    dqs[1][1] = ucoef; dqs[2][2] = ucoef; dqs[3][3] = ucoef;
    dqs[4][4] = Pcoef;
    const double ucoef2 = 0.5 * srhon / (srhonp1 * srhonp1 * srhonp1);
    const double du[3] = {un.x - U.x, un.y - U.y, un.z - U.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      dqs[1 + d][0] += ucoef2 * drdt * du[d];
      dqs[1 + d][4] += ucoef2 * drdp * du[d];
    }
    const double kcoef = vol / dt;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
      for (int j = 0; j < 5; ++j) dQ[i][j] *= kcoef;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
      for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int k = 0; k < 5; ++k) acc(D[i + 5 * j], dQ[i][k] * dqs[k][j]);
    const double e0n = p.Rt * Tn / (p.gamma - 1.) + .5 * dot(un, un);
    const V3 unp1 = {np1[1], np1[2], np1[3]};
    const double e0np = p.Rt * np1[0] / (p.gamma - 1.) + .5 * dot(unp1, unp1);
    const double rdtvol = vol / p.dtmax;
    const double rhonp = np1[5];
    s[0] = s[0] - (rhonp - rhon) * rdtvol;
    s[1] = s[1] - (rhonp * np1[1] - rhon * un.x) * rdtvol;
    s[2] = s[2] - (rhonp * np1[2] - rhon * un.y) * rdtvol;
    s[3] = s[3] - (rhonp * np1[3] - rhon * un.z) * rdtvol;
    s[4] = s[4] - (rhonp * e0np - rhon * e0n) * rdtvol;
#pragma unroll
    for (int k = 0; k < 5; ++k) rhs[(size_t)k * nc + c] = s[k];
  } else {
  acc(D[0 + 5 * 0], coef * rT);
  acc(D[1 + 5 * 0], coef * rT * U.x);
  acc(D[2 + 5 * 0], coef * rT * U.y);
  acc(D[3 + 5 * 0], coef * rT * U.z);
  acc(D[4 + 5 * 0], coef * fT);
  acc(D[0 + 5 * 4], coef * (pterm + rP * reta));
  acc(D[1 + 5 * 4], coef * (pterm + rP * reta) * U.x);
  acc(D[2 + 5 * 4], coef * (pterm + rP * reta) * U.y);
  acc(D[3 + 5 * 4], coef * (pterm + rP * reta) * U.z);
  acc(D[4 + 5 * 4], coef * (fT * pterm / rT + fP * reta));
  acc(D[1 + 5 * 1], coef * rho);
  acc(D[4 + 5 * 1], coef * rho * U.x);
  acc(D[2 + 5 * 2], coef * rho);
  acc(D[4 + 5 * 2], coef * rho * U.y);
  acc(D[3 + 5 * 3], coef * rho);
  acc(D[4 + 5 * 3], coef * rho * U.z);
  }
  if constexpr (TS == 1) {  // BDF2 rhs, src/solverTimestep.loci:400-428
    double Qp[5], Qn[5], Qm[5];
    cons5(p, T, pg, U, Qp);
    cons5(p, qn[c], qn[4 * nc + c], V3{qn[nc + c], qn[2 * nc + c], qn[3 * nc + c]}, Qn);
    cons5(p, qm1[c], qm1[4 * nc + c], V3{qm1[nc + c], qm1[2 * nc + c], qm1[3 * nc + c]}, Qm);
    const double rdtvol = vol / p.dtmax;
    const double w = 0.5;
    const double w1 = (1.0 + w) * rdtvol, w2 = w * rdtvol;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      s[k] = (s[k] - w1 * (Qp[k] - Qn[k]) + w2 * (Qn[k] - Qm[k]));
      rhs[(size_t)k * nc + c] = s[k];
    }
  }
  double b[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const F bk = (F)s[k];
    Bout[(size_t)c * 5 + k] = bk;
    b[k] = (double)bk;
  }
  double x[5];
  F *luc = lu + (size_t)c * 25;
  if constexpr (LS) {
    double Dd[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) {
      luc[i] = D[i];
      Dd[i] = (double)D[i];
    }
    piv[c] = PIV_ID;
    lu5np(Dd);
    double *l64 = lu64 + (size_t)c * 25;
#pragma unroll
    for (int i = 0; i < 25; ++i) l64[i] = Dd[i];
    lusolve5<double>(Dd, PIV_ID, b, x);
  } else {
    const int pk = lu5<F>(D);
#pragma unroll
    for (int i = 0; i < 25; ++i) luc[i] = D[i];
    piv[c] = pk;
    lusolve5<F>(D, pk, b, x);
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) X[(size_t)c * 5 + k] = x[k];
}

// ---------------------------------------------------------------------------
// block SGS (src/linearSolverSGS.loci:124-300).  One cell per lane:
//   t = -B; t += M_j X[nbr_j] over the row (upper faces, then lower faces);
//   X = LU^-1 (-t).
// Rows are read through a fixed-width neighbour list (pads point at a zero
// X entry), so every block and neighbour load of a row issues at once.
// ---------------------------------------------------------------------------
// LDS exchange inside a one-wave workgroup: a wave's LDS operations
// complete in issue order, so a lane's LDS read sees every LDS write the
// wave issued before it; the wavefront-scope fences only keep the compiler
// from moving memory operations across this point.  Unlike __syncthreads it
// does not wait for the wave's outstanding global loads and stores.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Five lanes per cell: lane i of the group owns equation i.  Each lane
// loads row i of every block of the row, component i of every neighbour's X
// and row i of the diagonal factors; the group shares the X components and
// the factors through LDS, each lane accumulates its row against the
// neighbours' whole X, the five partial sums are exchanged with cross-lane
// reads, and every lane of the group repeats the small LU solve and keeps
// its own component.  The summation order of each t_i is the one-lane order
// (blocks in row order, columns ascending), so the result is bitwise the
// same as the sequential sweep.

// one level of a forward or backward sweep; every cell of a level depends only
// on cells of other levels
template <typename F>
__global__ void __launch_bounds__(64) k_sgs_level(DevGrid g, int lo, int hi, const F *__restrict__ off,
                                                  const F *__restrict__ lu, const int *__restrict__ piv,
                                                  const F *__restrict__ B, double *__restrict__ X) {
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ F al[CPW * 25];
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int gbase = 5 * (q < CPW ? q : 0);
  const int c = lo + blockIdx.x * CPW + q;
  const bool act = q < CPW && c < hi;
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  F *aw = al + (q < CPW ? q : 0) * 25;
  constexpr bool SKIP_PADS = sizeof(F) == sizeof(double);
  double t = 0.0;
  int pk = 0, m0 = 0, len = 0;
  F mrow[MAXR][5];
  if (act) {
    t = -(double)B[(size_t)c * 5 + i];
    pk = piv[c];
    m0 = g.m_off[c];
    len = g.m_off[c + 1] - m0;
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.nbr_pad[(size_t)c * MAXR + j];
    double x1[MAXR];
    F a5[5];
    // FP64 blocks: pads (j >= len) load nothing and add nothing (C3's
    // tetrahedra: 30.2 -> 29.2 ms per step); FP32 blocks: pads load the last
    // block against the zero row, which measured faster (24.4 against 30.9 ms)
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool real = !SKIP_PADS || j < len;
      const int slot = len > 0 ? m0 + (j < len ? j : len - 1) : 0;
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = real ? M[i + 5 * jj] : (F)0;
      x1[j] = real ? X[(size_t)nb[j] * 5 + i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) a5[k] = lu[(size_t)c * 25 + 5 * i + k];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) xw[j * 5 + i] = x1[j];
#pragma unroll
    for (int k = 0; k < 5; ++k) aw[5 * i + k] = a5[k];
  }
  wave_sync();
  if (act) {
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj)
        if (!SKIP_PADS || j < len) t += mrow[j][jj] * xw[j * 5 + jj];
    for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
      const F *M = off + (size_t)(m0 + j) * 25;
      const double *xn = X + (size_t)g.m_nbr[m0 + j] * 5;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) t += M[i + 5 * jj] * xn[jj];
    }
  }
  double v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = -__shfl(t, gbase + k, 64);
  if (act) {
    double x[5];
    lusolve5<F>(aw, pk, v, x);
    const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
    X[(size_t)c * 5 + i] = xi;
  }
}

// One whole half-sweep in a single cooperative launch of W one-wave
// workgroups, all co-resident, as a dataflow over chunks.  Cells are cut
// into chunks of at most CPW cells that never straddle a level, numbered in
// sweep order; wave w takes chunks w, w+W, w+2W, ... in order.  A chunk
// waits only for the chunks holding its cells' dependencies (lower
// neighbours forward, upper neighbours backward; periodic ghosts are lagged
// copies made before the launch), each flagged with this launch's epoch.
// The wave holding the smallest unfinished chunk has finished its earlier
// chunks and all of that chunk's dependencies are smaller chunks, so it
// always proceeds: every wait ends.  A wait that exceeds `spin_limit` polls
// sets *err and returns (the host reports it), so a broken schedule cannot
// hang the device.
//
// Loads that cannot change before the chunk's dependencies are met are
// issued before the wait: indices, blocks, factors, B, and X of the
// neighbours on the side the sweep has not reached (upper forward, lower
// backward).  Forward, the upper terms come first in flowPsi's summation
// order, so they are also accumulated before the wait; the order of every
// sum is that of k_sgs_level, so the result is bitwise the same.  X is read
// and written with device-coherent accesses; a wave's stores complete
// (vmcnt = 0) before lane 0 publishes the chunk's flag.
//
// With ns > 0 the same launch also sweeps the ns turbulence scalar systems
// (segregated nuTilde, or k and omega): lane sc < ns of each cell's group
// updates scalar sc of that cell with the arithmetic of k_sa_sgs_level.  The
// scalar rows share the fluid's row structure, so the chunk dependencies
// cover them; this is the 6x6 / 7x7 block sweep with every fluid-scalar
// coupling zero, equal bit for bit to the separate sweeps.
struct ScalarRows {
  int ns;
  long nnz;
  const double *off, *dinv, *B;
  double *X;  // [ns][nc + nper + 1]
};

// turbulence scalar rows of a line-SGS half-sweep: scalar sc is swept when
// bit sc of mask is set; xt keeps each cell's value before the forward sweep
struct LineScalars {
  int ns, mask;
  long nnz;
  const double *off, *dinv, *B;
  double *X, *xt;
};

template <typename F, bool BACK>
__global__ void __launch_bounds__(64) k_sgs_sweep(DevGrid g, int nchunks, int nwaves, const int *__restrict__ chunk_lo,
                                                  const int *__restrict__ dep_off, const int *__restrict__ dep,
                                                  int *__restrict__ flag, int epoch, int spin_limit,
                                                  int *__restrict__ err, const F *__restrict__ off,
                                                  const F *__restrict__ lu, const int *__restrict__ piv,
                                                  const F *__restrict__ B, double *__restrict__ X, ScalarRows sr) {
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int gbase = 5 * (q < CPW ? q : 0);
  // component i of each neighbour's X and row i of the diagonal factors,
  // shared by the cell's five lanes after the wait
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ F al[CPW * 25];
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  F *aw = al + (q < CPW ? q : 0) * 25;
  for (int t = blockIdx.x; t < nchunks; t += nwaves) {
    const int ch = BACK ? nchunks - 1 - t : t;
    const int lo = chunk_lo[ch], hi = chunk_lo[ch + 1];
    const int c = lo + q;
    const bool act = q < CPW && c < hi;
    const int cc = act ? c : lo;
    double tsum = -(double)B[(size_t)cc * 5 + i];
    const int pk = piv[cc];
    const int m0 = g.m_off[cc];
    const int len = g.m_off[cc + 1] - m0;
    const int nup = g.m_nup[cc];
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.nbr_pad[(size_t)cc * MAXR + j];
    F mrow[MAXR][5];
    double x1[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const int slot = len > 0 ? m0 + (j < len ? j : len - 1) : 0;
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = M[i + 5 * jj];
    }
    // row i of the diagonal factors goes to LDS before the wait: no registers held across it
    if (q < CPW) {
#pragma unroll
      for (int k = 0; k < 5; ++k) aw[5 * i + k] = lu[(size_t)cc * 25 + 5 * i + k];
    }
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool pre = BACK ? (j >= nup) : (j < nup);
      x1[j] = pre ? ldx<false>(X + (size_t)nb[j] * 5 + i) : 0.0;
    }
    // turbulence scalar sc = i of this cell (lanes i < ns)
    const bool sact = act && i < sr.ns;
    const size_t xs = (size_t)g.nc + g.nper + 1;
    const int sc = sact ? i : 0;
    const double *soff = sr.ns ? sr.off + (size_t)sc * sr.nnz + m0 : nullptr;
    double *sx = sr.ns ? sr.X + sc * xs : nullptr;
    double ssum = 0.0, sdv = 0.0, sco[MAXR], sxv[MAXR];
    if (sact) {
      ssum = -sr.B[(size_t)sc * g.nc + cc];
      sdv = sr.dinv[(size_t)sc * g.nc + cc];  // before the wait: off the chain
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        sco[j] = j < len ? soff[j] : 0.0;
        const bool pre = BACK ? (j >= nup) : (j < nup);
        sxv[j] = (pre && j < len) ? ldx<false>(sx + nb[j]) : 0.0;
      }
      if (!BACK) {
#pragma unroll
        for (int j = 0; j < MAXR; ++j)
          if (j < nup && j < len) ssum += sxv[j] * sco[j];
      }
    }
    if (!BACK) {
      if (q < CPW) {
#pragma unroll
        for (int j = 0; j < MAXR; ++j)
          if (j < nup) xw[j * 5 + i] = x1[j];
      }
      wave_sync();
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
#pragma unroll
        for (int jj = 0; jj < 5; ++jj)
          if (j < nup) tsum += mrow[j][jj] * xw[j * 5 + jj];
    }
    // wait for the chunks this one depends on
    const int d0 = dep_off[ch], nd = dep_off[ch + 1] - d0;
    const int dl = lane < nd ? dep[d0 + lane] : -1;
    int spins = 0;
    while (true) {
      const int v = dl >= 0 ? __hip_atomic_load(flag + dl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : epoch;
      if (__all(v >= epoch)) break;
      if (++spins > spin_limit) {
        if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      __builtin_amdgcn_s_sleep(4);
    }
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool post = BACK ? (j < nup) : (j >= nup);
      if (post) x1[j] = ldx<true>(X + (size_t)nb[j] * 5 + i);
    }
    // the scalars' reached neighbours in the same round of loads
    if (sact) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool post = BACK ? (j < nup) : (j >= nup);
        if (post && j < len) sxv[j] = ldx<true>(sx + nb[j]);
      }
    }
    if (q < CPW) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (BACK || j >= nup) xw[j * 5 + i] = x1[j];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj)
        if (BACK || j >= nup) tsum += mrow[j][jj] * xw[j * 5 + jj];
    if (act)
      for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
        const F *M = off + (size_t)(m0 + j) * 25;
        const double *xn = X + (size_t)g.m_nbr[m0 + j] * 5;
#pragma unroll
        for (int jj = 0; jj < 5; ++jj) tsum += M[i + 5 * jj] * ldx<true>(xn + jj);
      }
    if (sact) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if ((BACK || j >= nup) && j < len) ssum += sxv[j] * sco[j];
      for (int j = MAXR; j < len; ++j) ssum += ldx<true>(sx + g.m_nbr[m0 + j]) * soff[j];
      const double sv = -ssum;
      stx<true>(sx + c, sdv * sv);
    }
    double v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = -__shfl(tsum, gbase + k, 64);
    if (act) {
      double x[5];
      lusolve5<F>(aw, pk, v, x);
      const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
      stx<true>(X + (size_t)c * 5 + i, xi);
    }
    // every store of the wave has completed (vmcnt = 0) before the publish
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (lane == 0) __hip_atomic_store(flag + ch, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The same half-sweep without flags, for the fluid system alone.  X is
// double-buffered: the source buffer Xs holds the values of the previous
// half-sweep and the target Xt starts as the sentinel X_SENT (a signalling
// NaN, which no arithmetic produces) on every interior row, with the lagged
// periodic ghosts and the zero pad row in place.  A lane reads component i
// of each neighbour the sweep has not reached (upper forward, lower
// backward) from Xs, and polls component i of each neighbour it has reached
// in Xt until no lane of the wave sees a sentinel: an aligned 8-byte store
// is seen whole, so a non-sentinel word is that neighbour's final value.
// The wait then costs one poll round trip on the neighbours' own words, and
// the producer neither waits for its stores nor publishes a flag; after its
// store it writes the sentinel over its own word of Xs, which no unfinished
// cell still reads, so that buffer is ready as the next half-sweep's
// target.  Chunks, order and summation are those of k_sgs_sweep, so the
// result is bitwise the same; the wave holding the smallest unfinished chunk
// always proceeds, and a wait longer than spin_limit polls (or an error
// raised by another wave) ends the launch with *err set.
constexpr unsigned long long X_SENT = 0x7FF45EE5DEAD0001ull;

__device__ __forceinline__ bool is_sent(double v) {
  return (unsigned long long)__double_as_longlong(v) == X_SENT;
}

// The value v, opaque to the optimiser: a poll loop recomputes the
// addresses it derives from v on every iteration (a few integer ops)
// instead of hoisting them and holding 64-bit addresses across the loop.
__device__ __forceinline__ unsigned opaque(unsigned v) {
  asm volatile("" : "+v"(v));
  return v;
}

// The reached neighbour a wave is likely to wait for longest: the largest
// (forward) or smallest (backward) candidate row over its lanes; a lane
// with no candidate passes -1 (forward) or INT_MAX (backward).  Rows are
// numbered in sweep order, so that row's producer comes last in the order
// the waves take their chunks.
template <bool BACK>
__device__ __forceinline__ int wave_latest(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int w = __shfl_xor(v, o, 64);
    v = BACK ? (w < v ? w : v) : (w > v ? w : v);
  }
  return v;
}

// The first stage of a flag-free wait: poll one word (p, the same for every
// lane, or nothing when p is null), with a pause between polls, until it is
// no longer the sentinel.  The word is the first component of the wave's
// latest reached neighbour (wave_latest): the full poll that follows finds
// the others written, or nearly so, and while the producers are still
// levels away the waiting wave issues one short request per poll instead
// of full gathers.  spins counts polls towards the caller's spin limit;
// false: the limit was passed here (*err set) or another wave raised *err.
__device__ __forceinline__ bool wait_witness(const double *p, int &spins, int spin_limit, int *err) {
  while (true) {
    const bool rdy = p ? !is_sent(ldx<true>(p)) : true;
    if (__all(rdy)) return true;
    if (++spins > spin_limit) {
      if (threadIdx.x == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __builtin_amdgcn_s_sleep(2);
  }
}

template <typename F, bool BACK>
__global__ void __launch_bounds__(64) k_sgs_flow(DevGrid g, int nchunks, int nwaves, const int *__restrict__ chunk_lo,
                                                 int spin_limit, int *__restrict__ err, const F *__restrict__ off,
                                                 const F *__restrict__ lu, const int *__restrict__ piv,
                                                 const F *__restrict__ B, double *__restrict__ Xs,
                                                 double *__restrict__ Xt) {
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int gbase = 5 * (q < CPW ? q : 0);
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ F al[CPW * 25];
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  F *aw = al + (q < CPW ? q : 0) * 25;
  for (int t = blockIdx.x; t < nchunks; t += nwaves) {
    const int ch = BACK ? nchunks - 1 - t : t;
    const int lo = chunk_lo[ch], hi = chunk_lo[ch + 1];
    const int c = lo + q;
    const bool act = q < CPW && c < hi;
    const int cc = act ? c : lo;
    double tsum = -(double)B[(size_t)cc * 5 + i];
    const int pk = piv[cc];
    const int m0 = g.m_off[cc];
    const int len = g.m_off[cc + 1] - m0;
    const int nup = g.m_nup[cc];
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.nbr_pad[(size_t)cc * MAXR + j];
    F mrow[MAXR][5];
    double x1[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const int slot = len > 0 ? m0 + (j < len ? j : len - 1) : 0;
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = M[i + 5 * jj];
    }
    // row i of the diagonal factors goes to LDS before the wait: no registers held across it
    if (q < CPW) {
#pragma unroll
      for (int k = 0; k < 5; ++k) aw[5 * i + k] = lu[(size_t)cc * 25 + 5 * i + k];
    }
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool pre = BACK ? (j >= nup) : (j < nup);
      x1[j] = pre ? ldx<false>(Xs + (size_t)nb[j] * 5 + i) : 0.0;
    }
    // the values the wait does not cover go to LDS now, so only the polled
    // words are held in registers across the wait
    if (q < CPW) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool pre = BACK ? (j >= nup) : (j < nup);
        if (pre) xw[j * 5 + i] = x1[j];
      }
    }
    if (!BACK) {  // forward, the upper terms come first
      wave_sync();
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
#pragma unroll
        for (int jj = 0; jj < 5; ++jj)
          if (j < nup) tsum += mrow[j][jj] * xw[j * 5 + jj];
    }
    int spins = 0;
    {
      int latest = BACK ? 0x7fffffff : -1;  // over the reached neighbours, pads excluded
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (BACK ? (j < nup) : (j >= nup && j < len))
          latest = BACK ? (nb[j] < latest ? nb[j] : latest) : (nb[j] > latest ? nb[j] : latest);
      latest = wave_latest<BACK>(latest);
      const bool any = BACK ? latest != 0x7fffffff : latest >= 0;
      if (!wait_witness(any ? Xt + (size_t)latest * 5 : nullptr, spins, spin_limit, err)) return;
    }
    // every poll issues all its loads before it checks any (no load waits
    // for the one before it) and loads only the words still holding the
    // sentinel (a word seen written is final); entries the wait does not
    // cover load nothing and stay zero.  Offsets are 32-bit and recomputed
    // on every poll, so no 64-bit addresses are held across it.
    const double sent = __longlong_as_double((long long)X_SENT);
    double xv[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) xv[j] = (BACK ? (j < nup) : (j >= nup)) ? sent : 0.0;
    while (true) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (is_sent(xv[j])) xv[j] = ldx<true>(Xt + opaque((unsigned)nb[j] * 5u + (unsigned)i));
      bool ready = true;
#pragma unroll
      for (int j = 0; j < MAXR; ++j) ready &= !is_sent(xv[j]);
      if (__all(ready)) break;
      if (++spins > spin_limit) {
        if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
      __builtin_amdgcn_s_sleep(1);
    }
    if (q < CPW) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool post = BACK ? (j < nup) : (j >= nup);
        if (post) xw[j * 5 + i] = xv[j];
      }
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj)
        if (BACK || j >= nup) tsum += mrow[j][jj] * xw[j * 5 + jj];
    if (act)
      for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
        const F *M = off + (size_t)(m0 + j) * 25;
        const bool post = BACK ? (j < nup) : (j >= nup);
        const double *xn = (post ? (const double *)Xt : Xs) + (size_t)g.m_nbr[m0 + j] * 5;
#pragma unroll
        for (int jj = 0; jj < 5; ++jj) {
          double xnv = ldx<true>(xn + jj);
          for (int sp = 0; post && is_sent(xnv); ++sp) {
            if (sp > spin_limit) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              xnv = 0.0;
              break;
            }
            xnv = ldx<true>(xn + jj);
          }
          tsum += M[i + 5 * jj] * xnv;
        }
      }
    double v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = -__shfl(tsum, gbase + k, 64);
    if (act) {
      double x[5];
      lusolve5<F>(aw, pk, v, x);
      const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
      stx<true>(Xt + (size_t)c * 5 + i, xi);
      // every cell that reads this word of Xs (a neighbour the sweep had not
      // reached when it ran) finished before this cell's poll succeeded, so
      // the word becomes the next half-sweep's sentinel
      stx<true>(Xs + (size_t)c * 5 + i, __longlong_as_double((long long)X_SENT));
    }
  }
}

// the target buffer of a flag-free half-sweep: sentinels on the interior
// rows (fill, needed before the first half-sweep of a solve: each later one
// finds them left by the previous half-sweep); the lagged periodic ghosts
// of the latest values in both buffers
__global__ void k_sweep_prep(DevGrid g, int fill, double *__restrict__ Xs, double *__restrict__ Xt) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = 5L * g.nc;
  if (fill && e < n) Xt[e] = __longlong_as_double((long long)X_SENT);
  if (e < 5L * g.nper) {
    const int pi = (int)(e / 5), k = (int)(e - 5L * pi);
    const double v = Xs[(size_t)g.frc[g.nfi + pi] * 5 + k];
    Xs[(size_t)(g.nc + pi) * 5 + k] = v;
    Xt[(size_t)(g.nc + pi) * 5 + k] = v;
  }
}

// Line-SGS / FSGS half-sweep as one cooperative dataflow launch: the
// chunks, dependency lists and flags of k_sgs_sweep (valid when the
// breadth-first levels are the SGS levels), the per-cell update of
// k_lsgs_level (rows rotated lower-first, FP64 unpivoted diagonal, X kept in
// xt before the backward sweep overwrites it) or of k_fsgs_level (DG = F:
// the unpivoted factors in the Jacobian precision, no xt).  The summation
// order per component is the level kernels', so the result is bitwise the
// same: forward, the lower terms come first and are summed after the wait;
// backward, the lower terms (final since the forward sweep) are summed
// before it.
// With line-scalar rows (ls.mask != 0) lane i < ns of each cell also sweeps
// turbulence scalar i as k_sa_lsgs_level does (X = (B - sum off X_nb) / D,
// rows rotated lower-first), in the same launch; fluid = 0 leaves the block
// rows alone.  Every system keeps its own per-cell order, so the result is
// bitwise that of separate sweeps.
template <typename F, typename DG, bool BACK>
__global__ void __launch_bounds__(64) k_lsgs_sweep(DevGrid g, int nchunks, int nwaves, const int *__restrict__ chunk_lo,
                                                   const int *__restrict__ dep_off, const int *__restrict__ dep,
                                                   int *__restrict__ flag, int epoch, int spin_limit,
                                                   int *__restrict__ err, const F *__restrict__ off,
                                                   const DG *__restrict__ diag, const F *__restrict__ B,
                                                   double *__restrict__ X, double *__restrict__ xt, int fluid,
                                                   LineScalars ls) {
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int gbase = 5 * (q < CPW ? q : 0);
  // component i of each neighbour's X and row i of the diagonal factors,
  // shared by the cell's five lanes after the wait
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ DG al[CPW * 25];
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  DG *aw = al + (q < CPW ? q : 0) * 25;
  for (int t = blockIdx.x; t < nchunks; t += nwaves) {
    const int ch = BACK ? nchunks - 1 - t : t;
    const int lo = chunk_lo[ch], hi = chunk_lo[ch + 1];
    const int c = lo + q;
    const bool act = q < CPW && c < hi;
    const bool fact = act && fluid;
    const int cc = act ? c : lo;
    double tsum = -(double)B[(size_t)cc * 5 + i];
    const int m0 = g.m_off[cc];
    const int len = g.m_off[cc + 1] - m0;
    const int nup = g.m_nup[cc];
    const int nlow = len - nup;  // rotated rows: lower entries j < nlow, then upper
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.lnbr_pad[(size_t)cc * MAXR + j];
    F mrow[MAXR][5];
    double x1[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      int jr = j + nup;
      if (jr >= len) jr -= len;
      const int slot = j < len ? m0 + jr : (len > 0 ? m0 : 0);
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = M[i + 5 * jj];
    }
    DG a5[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) a5[k] = diag[(size_t)cc * 25 + 5 * i + k];
    // values this half-sweep does not write: upper (forward) / lower (backward), and pads
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool pre = BACK ? (j < nlow || j >= len) : (j >= nlow);
      x1[j] = pre ? ldx<false>(X + (size_t)nb[j] * 5 + i) : 0.0;
    }
    double xold = 0.0;
    if constexpr (BACK) {
      if (xt) xold = ldx<false>(X + (size_t)cc * 5 + i);
    }
    // turbulence scalar sc = i of this cell
    const bool sact = act && i < ls.ns && ((ls.mask >> i) & 1);
    const int sc = sact ? i : 0;
    const size_t xs = (size_t)g.nc + g.nper + 1;
    const double *soff = ls.mask ? ls.off + (size_t)sc * ls.nnz + m0 : nullptr;
    double *sx = ls.mask ? ls.X + sc * xs : nullptr;
    double sres = 0.0, sold = 0.0, sdv = 0.0, sco[MAXR], sxv[MAXR];
    if (sact) {
      sres = ls.B[(size_t)sc * g.nc + cc];
      sdv = ls.dinv[(size_t)sc * g.nc + cc];  // before the wait: off the chain
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        int jr = j + nup;
        if (jr >= len) jr -= len;
        sco[j] = j < len ? soff[jr] : 0.0;
        const bool pre = j < len && (BACK ? (j < nlow) : (j >= nlow));
        sxv[j] = pre ? ldx<false>(sx + nb[j]) : 0.0;
      }
      if (!BACK) sold = ldx<false>(sx + cc);
      if (BACK) {
#pragma unroll
        for (int j = 0; j < MAXR; ++j)
          if (j < nlow) sres -= sxv[j] * sco[j];
      }
    }
    if constexpr (BACK) {  // the lower terms are final since the forward sweep
      if (q < CPW) {
#pragma unroll
        for (int j = 0; j < MAXR; ++j)
          if (j < nlow) xw[j * 5 + i] = x1[j];
      }
      wave_sync();
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
#pragma unroll
        for (int jj = 0; jj < 5; ++jj)
          if (j < nlow) tsum += mrow[j][jj] * xw[j * 5 + jj];
    }
    const int d0 = dep_off[ch], nd = dep_off[ch + 1] - d0;
    const int dl = lane < nd ? dep[d0 + lane] : -1;
    int spins = 0;
    while (true) {
      const int v = dl >= 0 ? __hip_atomic_load(flag + dl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : epoch;
      if (__all(v >= epoch)) break;
      if (++spins > spin_limit) {
        if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    // from the wait to the publish this chunk is on the dependency chain:
    // it goes ahead of the SIMD's waves still loading their own chunks
    __builtin_amdgcn_s_setprio(3);
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool post = BACK ? (j >= nlow && j < len) : (j < nlow);
      if (post) x1[j] = ldx<true>(X + (size_t)nb[j] * 5 + i);
    }
    if (q < CPW) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (!BACK || j >= nlow) xw[j * 5 + i] = x1[j];
#pragma unroll
      for (int k = 0; k < 5; ++k) aw[5 * i + k] = a5[k];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj)
        if (!BACK || j >= nlow) tsum += mrow[j][jj] * xw[j * 5 + jj];
    if (fact)
      for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
        int jr = j + nup;
        if (jr >= len) jr -= len;
        const F *M = off + (size_t)(m0 + jr) * 25;
        const double *xn = X + (size_t)g.m_nbr[m0 + jr] * 5;
#pragma unroll
        for (int jj = 0; jj < 5; ++jj) tsum += M[i + 5 * jj] * ldx<true>(xn + jj);
      }
    if (sact) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool post = j < len && (BACK ? (j >= nlow) : (j < nlow));
        if (post) sxv[j] = ldx<true>(sx + nb[j]);
      }
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (j < len && (!BACK || j >= nlow)) sres -= sxv[j] * sco[j];
      for (int j = MAXR; j < len; ++j) {
        int jr = j + nup;
        if (jr >= len) jr -= len;
        sres -= ldx<true>(sx + g.m_nbr[m0 + jr]) * soff[jr];
      }
      if (!BACK) ls.xt[(size_t)sc * g.nc + c] = sold;
      stx<true>(sx + c, sdv * sres);
    }
    double v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = -__shfl(tsum, gbase + k, 64);
    if (fact) {
      if constexpr (BACK) {
        if (xt) xt[(size_t)c * 5 + i] = xold;
      }
      double x[5];
      lusolve5<DG>(aw, PIV_ID, v, x);
      const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
      stx<true>(X + (size_t)c * 5 + i, xi);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (lane == 0) __hip_atomic_store(flag + ch, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_setprio(0);
  }
}

// Line-SGS half-sweep as one cooperative dataflow launch without flags, for
// the fluid system and the turbulence scalars of its fused loops.  X is
// double-buffered.  The forward half-sweep reads the previous iterate of
// the neighbours it has not reached (upper) from Xa and polls the ones it
// has reached (lower) in Xb, whose interior rows hold the sentinel X_SENT
// when it starts; it writes Xb and then the sentinel over its own word of
// Xa: every cell that reads that word (a lower neighbour) finished before
// this cell's poll succeeded.  The backward half-sweep reads its lower
// neighbours' forward values from Xb (final since the forward sweep) and
// polls its upper neighbours in Xa; it leaves Xb alone, so Xb keeps X after
// the forward sweep, the relaxation's xt (k_lsgs_relax_flow forms
// X = w Xa + (1 - w) Xb in Xa and refills Xb with sentinels).  A turbulence
// scalar does the same with its own pair (ls.X, sXb) and keeps its own xt
// (the iterate before the forward sweep) as k_lsgs_sweep does.  Periodic
// ghost rows and the zero pad row are never sentinels: ghosts are upper
// entries, read from Xa in both directions; pads are read from the source
// buffer of either direction, so Xb's pad row must hold zero.  Chunks,
// order and summation are those of k_lsgs_sweep, so the result is bitwise
// the same; the wave holding the smallest unfinished chunk always proceeds,
// and a wait longer than spin_limit polls (or an error raised by another
// wave) ends the launch with *err set.  A system that is not in the launch
// (fluid = 0, or its scalar's mask bit clear) polls and stores nothing.
template <typename F, typename DG, bool BACK>
__global__ void __launch_bounds__(64, 3) k_lsgs_flow(DevGrid g, int nchunks, int nwaves, const int *__restrict__ chunk_lo,
                                                  int spin_limit, int *__restrict__ err, const F *__restrict__ off,
                                                  const DG *__restrict__ diag, const F *__restrict__ B,
                                                  double *__restrict__ Xa, double *__restrict__ Xb, int fluid,
                                                  LineScalars ls, double *__restrict__ sXb) {
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int gbase = 5 * (q < CPW ? q : 0);
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ DG al[CPW * 25];
  // With FP64 blocks the diagonal row and the scalar row coefficients wait
  // for the dependencies in LDS rather than registers (no spills at 3
  // waves per SIMD); with FP32 blocks registers are free and the LDS
  // stores would wait for the loads before the polls are issued.
  constexpr bool STAGE = sizeof(F) == 8;
  // scalar i's row coefficients of the cell (lanes i < ns) when staged
  __shared__ double sl[CPW * 5 * MAXR];
  // scalar i's neighbour values of the cell (lanes i < ns), held in LDS across the wait
  __shared__ double svl[CPW * 5 * MAXR];
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  DG *aw = al + (q < CPW ? q : 0) * 25;
  double *sw = sl + ((q < CPW ? q : 0) * 5 + i) * MAXR;
  double *svw = svl + ((q < CPW ? q : 0) * 5 + i) * MAXR;
  const double *Xs = BACK ? Xb : Xa;  // values this half-sweep does not write
  double *Xt = BACK ? Xa : Xb;        // the half-sweep's results, polled
  const double sent = __longlong_as_double((long long)X_SENT);
  for (int t = blockIdx.x; t < nchunks; t += nwaves) {
    const int ch = BACK ? nchunks - 1 - t : t;
    const int lo = chunk_lo[ch], hi = chunk_lo[ch + 1];
    const int c = lo + q;
    const bool act = q < CPW && c < hi;
    const bool fact = act && fluid;
    const int cc = act ? c : lo;
    double tsum = -(double)B[(size_t)cc * 5 + i];
    const int m0 = g.m_off[cc];
    const int len = g.m_off[cc + 1] - m0;
    const int nup = g.m_nup[cc];
    const int nlow = len - nup;  // rotated rows: lower entries j < nlow, then upper
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.lnbr_pad[(size_t)cc * MAXR + j];
    F mrow[MAXR][5];
    double x1[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      int jr = j + nup;
      if (jr >= len) jr -= len;
      const int slot = j < len ? m0 + jr : (len > 0 ? m0 : 0);
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = M[i + 5 * jj];
    }
    DG a5[5];
    if constexpr (STAGE) {  // row i of the diagonal factors into LDS before the wait
      if (q < CPW) {
#pragma unroll
        for (int k = 0; k < 5; ++k) aw[5 * i + k] = diag[(size_t)cc * 25 + 5 * i + k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 5; ++k) a5[k] = diag[(size_t)cc * 25 + 5 * i + k];
    }
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      const bool pre = BACK ? (j < nlow || j >= len) : (j >= nlow);
      x1[j] = pre ? ldx<false>(Xs + (size_t)nb[j] * 5 + i) : 0.0;
    }
    // turbulence scalar sc = i of this cell
    const bool sact = act && i < ls.ns && ((ls.mask >> i) & 1);
    const int sc = sact ? i : 0;
    const size_t xs = (size_t)g.nc + g.nper + 1;
    const double *soff = ls.mask ? ls.off + (size_t)sc * ls.nnz + m0 : nullptr;
    double *sxa = ls.mask ? ls.X + sc * xs : nullptr;
    double *sxb = ls.mask ? sXb + sc * xs : nullptr;
    const double *sxs = BACK ? sxb : sxa;
    double *sxt = BACK ? sxa : sxb;
    double sres = 0.0, sold = 0.0, sdv = 0.0, sco[MAXR], sxv[MAXR];
    if (sact) {
      sres = ls.B[(size_t)sc * g.nc + cc];
      sdv = ls.dinv[(size_t)sc * g.nc + cc];  // before the wait: off the chain
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        int jr = j + nup;
        if (jr >= len) jr -= len;
        const double co = j < len ? soff[jr] : 0.0;
        if constexpr (STAGE) sw[j] = co;
        else sco[j] = co;
        const bool pre = j < len && (BACK ? (j < nlow) : (j >= nlow));
        sxv[j] = pre ? ldx<false>(sxs + nb[j]) : 0.0;
      }
      if (!BACK) sold = ldx<false>(sxa + cc);
      if (BACK) {
#pragma unroll
        for (int j = 0; j < MAXR; ++j)
          if (j < nlow) sres -= sxv[j] * (STAGE ? sw[j] : sco[j]);
      }
#pragma unroll
      for (int j = 0; j < MAXR; ++j) svw[j] = sxv[j];
    }
    // the values the wait does not cover go to LDS now, so only the polled
    // words are held in registers across the wait
    if (q < CPW) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool pre = BACK ? (j < nlow || j >= len) : (j >= nlow);
        if (pre) xw[j * 5 + i] = x1[j];
      }
    }
    if constexpr (BACK) {  // the lower terms are final since the forward sweep
      wave_sync();
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
#pragma unroll
        for (int jj = 0; jj < 5; ++jj)
          if (j < nlow) tsum += mrow[j][jj] * xw[j * 5 + jj];
    }
    // the wait: poll the words of the reached neighbours in the target buffers
    int spins = 0;
    {
      int latest = BACK ? 0x7fffffff : -1;  // the fluid's reached neighbours
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (BACK ? (j >= nlow && j < len) : (j < nlow))
          latest = BACK ? (nb[j] < latest ? nb[j] : latest) : (nb[j] > latest ? nb[j] : latest);
      latest = wave_latest<BACK>(latest);
      const bool any = fluid && (BACK ? latest != 0x7fffffff : latest >= 0);
      if (!wait_witness(any ? Xt + (size_t)latest * 5 : nullptr, spins, spin_limit, err)) return;
    }
    // every poll issues all its loads before it checks any (no load waits
    // for the one before it) and loads only the words still holding the
    // sentinel (a word seen written is final); entries the wait does not
    // cover load nothing and stay zero.  Offsets are 32-bit and recomputed
    // on every poll, so no 64-bit addresses are held across it.
    const double *sbase = ls.mask ? (BACK ? ls.X : sXb) : Xt;  // the scalars' target buffers
    const unsigned sofs = (unsigned)sc * (unsigned)xs;
    double xv[MAXR], sv[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      xv[j] = fluid && (BACK ? (j >= nlow && j < len) : (j < nlow)) ? sent : 0.0;
      sv[j] = sact && j < len && (BACK ? (j >= nlow) : (j < nlow)) ? sent : 0.0;
    }
    while (true) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (is_sent(xv[j])) xv[j] = ldx<true>(Xt + opaque((unsigned)nb[j] * 5u + (unsigned)i));
#pragma unroll
      for (int j = 0; j < MAXR; ++j)
        if (is_sent(sv[j])) sv[j] = ldx<true>(sbase + opaque(sofs + (unsigned)nb[j]));
      bool ready = true;
#pragma unroll
      for (int j = 0; j < MAXR; ++j) ready &= !is_sent(xv[j]) & !is_sent(sv[j]);
      if (__all(ready)) break;
      if (++spins > spin_limit) {
        if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
      }
      if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
      __builtin_amdgcn_s_sleep(1);
    }
    if (q < CPW && fluid) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool post = BACK ? (j >= nlow && j < len) : (j < nlow);
        if (post) xw[j * 5 + i] = xv[j];
      }
      if constexpr (!STAGE) {
#pragma unroll
        for (int k = 0; k < 5; ++k) aw[5 * i + k] = a5[k];
      }
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj)
        if (!BACK || j >= nlow) tsum += mrow[j][jj] * xw[j * 5 + jj];
    if (fact)
      for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
        int jr = j + nup;
        if (jr >= len) jr -= len;
        const F *M = off + (size_t)(m0 + jr) * 25;
        const bool post = BACK ? (j >= nlow) : (j < nlow);
        const double *xn = (post ? (const double *)Xt : Xs) + (size_t)g.m_nbr[m0 + jr] * 5;
#pragma unroll
        for (int jj = 0; jj < 5; ++jj) {
          double xnv = ldx<true>(xn + jj);
          for (int sp = 0; post && is_sent(xnv); ++sp) {
            if (sp > spin_limit) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              xnv = 0.0;
              break;
            }
            xnv = ldx<true>(xn + jj);
          }
          tsum += M[i + 5 * jj] * xnv;
        }
      }
    if (sact) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) {
        const bool post = j < len && (BACK ? (j >= nlow) : (j < nlow));
        const double xj = post ? sv[j] : svw[j];
        if (j < len && (!BACK || j >= nlow)) sres -= xj * (STAGE ? sw[j] : sco[j]);
      }
      for (int j = MAXR; j < len; ++j) {
        int jr = j + nup;
        if (jr >= len) jr -= len;
        const bool post = BACK ? (j >= nlow) : (j < nlow);
        const double *xp = (post ? (const double *)sxt : sxs) + g.m_nbr[m0 + jr];
        double xv = ldx<true>(xp);
        for (int sp = 0; post && is_sent(xv); ++sp) {
          if (sp > spin_limit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            xv = 0.0;
            break;
          }
          xv = ldx<true>(xp);
        }
        sres -= xv * soff[jr];
      }
      if (!BACK) ls.xt[(size_t)sc * g.nc + c] = sold;
      stx<true>(sxt + c, sdv * sres);
      if (!BACK) stx<true>(sxa + c, sent);
    }
    double v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = -__shfl(tsum, gbase + k, 64);
    if (fact) {
      double x[5];
      lusolve5<DG>(aw, PIV_ID, v, x);
      const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
      stx<true>(Xt + (size_t)c * 5 + i, xi);
      if (!BACK) stx<true>(Xa + (size_t)c * 5 + i, sent);
    }
  }
}

// Relaxation after a flag-free line-SGS sweep (src/linearSolverLSGS.loci:
// 1868-1874 for the fluid, :1330-1340 for a scalar): X = w Xa + (1 - w) xt
// into Xa with the block sums of (X - xt)^2, as k_lsgs_relax; then the
// sentinel into `fill`, the forward half-sweep's target of the next sweep
// (the fluid's xt is that buffer itself, read before it is overwritten).
__global__ void k_lsgs_relax_flow(int n, const double *__restrict__ st, double *__restrict__ Xa,
                                  const double *xt, double *fill, double *__restrict__ part) {
  __shared__ double sh[TPB];
  const int i = blockIdx.x * TPB + threadIdx.x;
  const double w = st[0];
  double r2 = 0.0;
  if (i < n) {
    const double xo = xt[i];
    const double x = w * Xa[i] + (1. - w) * xo;
    Xa[i] = x;
    fill[i] = __longlong_as_double((long long)X_SENT);
    const double r = x - xo;
    r2 = r * r;
  }
  sh[threadIdx.x] = r2;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// the forward target of a flag-free line-SGS solve: sentinels on the n
// interior words, zero on the pad row (npad words at `pad`)
__global__ void k_flow_fill(long n, double *__restrict__ X, long npad, double *__restrict__ pad) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) X[e] = __longlong_as_double((long long)X_SENT);
  if (e < npad) pad[e] = 0.0;
}

// periodic ghost copies, src/linearSolverSGS.loci:124-127 and :215-219
__global__ void k_ghost(DevGrid g, double *__restrict__ X) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= g.nper) return;
  const int src = g.frc[g.nfi + pi];
#pragma unroll
  for (int k = 0; k < 5; ++k) X[(size_t)(g.nc + pi) * 5 + k] = X[(size_t)src * 5 + k];
}

__global__ void k_ghost1(DevGrid g, int ns, double *__restrict__ X) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= g.nper) return;
  const size_t xs = (size_t)g.nc + g.nper + 1;
  for (int sc = 0; sc < ns; ++sc) X[sc * xs + g.nc + pi] = X[sc * xs + g.frc[g.nfi + pi]];
}

// ---------------------------------------------------------------------------
// line SGS, src/linearSolverLSGS.loci:1692-1954.  The line search of that
// file links cells only through duplicate couples, which it removes first
// (:625-632 then :642-664), so every line is one cell: a line solve is
// X_c = D_c^-1 (B_c - sum_lower L X[cl] - sum_upper U X[cr]) with D_c's FP64
// unpivoted factors, in the breadth-first line order (:902-937).  Levels of
// that order run as one launch each; 5 lanes per cell as in k_sgs_level, with
// the row read lower faces first (linerhs :1498-1516 over the outer entries
// built at :1036-1066).  The backward sweep also keeps X after the forward
// sweep in xt (:1827-1829) for the relaxation.
// ---------------------------------------------------------------------------
template <typename F, bool BACK>
__global__ void __launch_bounds__(64) k_lsgs_level(DevGrid g, const int *__restrict__ cells, int lo, int hi,
                                                   const F *__restrict__ off, const double *__restrict__ lu64,
                                                   const F *__restrict__ B, double *__restrict__ X,
                                                   double *__restrict__ xt) {
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int k = lo + blockIdx.x * CPW + q;
  const bool act = q < CPW && k < hi;
  const int gbase = 5 * (q < CPW ? q : 0);
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ double al[CPW * 25];
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  double *aw = al + (q < CPW ? q : 0) * 25;
  double t = 0.0;
  int c = 0, m0 = 0, len = 0, nup = 0;
  F mrow[MAXR][5];
  if (act) {
    c = cells ? cells[k] : k;
    t = -(double)B[(size_t)c * 5 + i];
    m0 = g.m_off[c];
    len = g.m_off[c + 1] - m0;
    nup = g.m_nup[c];
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.lnbr_pad[(size_t)c * MAXR + j];
    double x1[MAXR];
    double a5[5];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      int jr = j + nup;
      if (jr >= len) jr -= len;
      const int slot = j < len ? m0 + jr : (len > 0 ? m0 : 0);
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = M[i + 5 * jj];
      x1[j] = X[(size_t)nb[j] * 5 + i];
    }
#pragma unroll
    for (int kk = 0; kk < 5; ++kk) a5[kk] = lu64[(size_t)c * 25 + 5 * i + kk];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) xw[j * 5 + i] = x1[j];
#pragma unroll
    for (int kk = 0; kk < 5; ++kk) aw[5 * i + kk] = a5[kk];
  }
  wave_sync();
  if (act) {
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) t += mrow[j][jj] * xw[j * 5 + jj];
    for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
      int jr = j + nup;
      if (jr >= len) jr -= len;
      const F *M = off + (size_t)(m0 + jr) * 25;
      const double *xn = X + (size_t)g.m_nbr[m0 + jr] * 5;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) t += M[i + 5 * jj] * xn[jj];
    }
  }
  double v[5];
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) v[kk] = -__shfl(t, gbase + kk, 64);
  if (act) {
    if constexpr (BACK) xt[(size_t)c * 5 + i] = X[(size_t)c * 5 + i];
    double x[5];
    lusolve5<double>(aw, PIV_ID, v, x);
    const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
    X[(size_t)c * 5 + i] = xi;
  }
}

// FSGS (src/linearSolverFSGS.loci:1297-1356): the diagonal factored without
// pivots in the Jacobian precision F (from the unfactored fluid_D the LSGS
// assembly keeps in lu), and one level of a sweep in the line order with
// those factors; rows read lower faces first, as in line SGS
template <typename F>
__global__ void k_fsgs_factor(int nc, const F *__restrict__ D, F *__restrict__ luf) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  F a[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = D[(size_t)c * 25 + i];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      const F l = a[i + 5 * k] / a[k + 5 * k];
      a[i + 5 * k] = l;
#pragma unroll
      for (int j = k + 1; j < 5; ++j) a[i + 5 * j] = a[i + 5 * j] - l * a[k + 5 * j];
    }
#pragma unroll
  for (int i = 0; i < 25; ++i) luf[(size_t)c * 25 + i] = a[i];
}

template <typename F>
__global__ void __launch_bounds__(64) k_fsgs_level(DevGrid g, const int *__restrict__ cells, int lo, int hi,
                                                   const F *__restrict__ off, const F *__restrict__ luf,
                                                   const F *__restrict__ B, double *__restrict__ X) {
  const int lane = threadIdx.x;
  const int q = lane / 5, i = lane - 5 * q;
  const int k = lo + blockIdx.x * CPW + q;
  const bool act = q < CPW && k < hi;
  const int gbase = 5 * (q < CPW ? q : 0);
  __shared__ double xl[CPW * MAXR * 5];
  __shared__ F al[CPW * 25];
  double *xw = xl + (q < CPW ? q : 0) * (MAXR * 5);
  F *aw = al + (q < CPW ? q : 0) * 25;
  double t = 0.0;
  int c = 0, m0 = 0, len = 0, nup = 0;
  F mrow[MAXR][5];
  if (act) {
    c = cells ? cells[k] : k;
    t = -(double)B[(size_t)c * 5 + i];
    m0 = g.m_off[c];
    len = g.m_off[c + 1] - m0;
    nup = g.m_nup[c];
    int nb[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) nb[j] = g.lnbr_pad[(size_t)c * MAXR + j];
    double x1[MAXR];
    F a5[5];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      int jr = j + nup;
      if (jr >= len) jr -= len;
      const int slot = j < len ? m0 + jr : (len > 0 ? m0 : 0);
      const F *M = off + (size_t)slot * 25;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = M[i + 5 * jj];
      x1[j] = X[(size_t)nb[j] * 5 + i];
    }
#pragma unroll
    for (int kk = 0; kk < 5; ++kk) a5[kk] = luf[(size_t)c * 25 + 5 * i + kk];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) xw[j * 5 + i] = x1[j];
#pragma unroll
    for (int kk = 0; kk < 5; ++kk) aw[5 * i + kk] = a5[kk];
  }
  wave_sync();
  if (act) {
#pragma unroll
    for (int j = 0; j < MAXR; ++j)
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) t += mrow[j][jj] * xw[j * 5 + jj];
    for (int j = MAXR; j < len; ++j) {  // rows wider than MAXR
      int jr = j + nup;
      if (jr >= len) jr -= len;
      const F *M = off + (size_t)(m0 + jr) * 25;
      const double *xn = X + (size_t)g.m_nbr[m0 + jr] * 5;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) t += M[i + 5 * jj] * xn[jj];
    }
  }
  double v[5];
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) v[kk] = -__shfl(t, gbase + kk, 64);
  if (act) {
    double x[5];
    lusolve5<F>(aw, PIV_ID, v, x);
    const double xi = i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : i == 3 ? x[3] : x[4];
    X[(size_t)c * 5 + i] = xi;
  }
}

__global__ void k_zero_d(long n, double *__restrict__ x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// st: w, zero_resid, resid, iterations run, done
__global__ void k_lsgs_init(double *__restrict__ st, double w0) {
  if (threadIdx.x == 0) {
    st[0] = w0; st[1] = 1.0; st[2] = 0.0; st[3] = 0.0; st[4] = 0.0;
  }
}

// X = w X + (1-w) xt and the block sums of (X - xt)^2 (:1869-1874)
__global__ void k_lsgs_relax(int n, const double *__restrict__ st, double *__restrict__ X,
                             const double *__restrict__ xt, double *__restrict__ part) {
  __shared__ double sh[TPB];
  const int i = blockIdx.x * TPB + threadIdx.x;
  const double w = st[0];
  double r2 = 0.0;
  if (i < n) {
    const double xo = xt[i];
    const double x = w * X[i] + (1. - w) * xo;
    X[i] = x;
    const double r = x - xo;
    r2 = r * r;
  }
  sh[threadIdx.x] = r2;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// update norm, relaxation control and stopping test (:1927-1948)
__global__ void k_lsgs_check(int nparts, const double *__restrict__ part, double *__restrict__ st, int iter,
                             double abstol, double reltol) {
  __shared__ double sh[TPB];
  double a = 0.0;
  {
    // the same in-order sum per thread, with eight loads in flight
    int j = threadIdx.x;
    for (; j + 7 * TPB < nparts; j += 8 * TPB) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part[j + u * TPB];
#pragma unroll
      for (int u = 0; u < 8; ++u) a += v[u];
    }
    for (; j < nparts; j += TPB) a += part[j];
  }
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double resid = sqrt(sh[0]);
    const double zero = iter == 0 ? (resid > 1e-16 ? resid : 1e-16) : st[1];
    const double rel = resid / zero;
    double w = st[0];
    if (rel > 1.0) w *= 1. / rel;
    st[0] = w;
    st[1] = zero;
    st[2] = resid;
    st[3] = (double)(iter + 1);
    st[4] = (resid < abstol || rel < reltol) ? 1.0 : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Spalart-Allmaras nuTilde equation (turbulence/SA/SA_turb.loci, baseline
// model, standard production; src/solverScalarTransport.loci), segregated
// as in the reference: its own scalar system and scalar SGS per Newton step.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sa_fv1(const Prm &p, double x) {
  return pow(x, 3.0) / (pow(x, 3.0) + pow(p.sa_cv1, 3.0));
}

// ---------------------------------------------------------------------------
// Menter SST k-omega (turbulence/KOmega/sst.loci) with the default options:
// turbulence_model SST, kprod_func SOmega, compressibilityCorrection wilcox,
// dissipationCorrection none, smooth walls, Kprod_clip 1e100, tmu_clip 1e5.
// sst1 (k-omega set) and sst2 (k-epsilon set), sst.loci:319-347; sigma_k of
// the SST model 0.85 (:682-684).
// ---------------------------------------------------------------------------
constexpr double SST_SIGK1 = 0.85, SST_SIGW1 = 0.5, SST_BETA1 = 0.075, SST_BETAS = 0.09, SST_KAPPA = 0.41,
                 SST_A1 = 0.31, SST_SIGK2 = 1.0, SST_SIGW2 = 0.856, SST_BETA2 = 0.0828;

// strainRate_t, divu_t (sst.loci:65-104), vortMag_t (src/solverTurbulence.loci:45-49)
// of gu[3i+j] = d u_i / d x_j
__device__ __forceinline__ void sst_kinematics(const double *gu, double &S, double &divu, double &Om) {
  const double dudx = gu[0], dudy = gu[1], dudz = gu[2];
  const double dvdx = gu[3], dvdy = gu[4], dvdz = gu[5];
  const double dwdx = gu[6], dwdy = gu[7], dwdz = gu[8];
  divu = (dudx + dvdy + dwdz);
  const double S2 = 2. * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + (dudy + dvdx) * (dudy + dvdx) +
                    (dudz + dwdx) * (dudz + dwdx) + (dvdz + dwdy) * (dvdz + dwdy);
  S = sqrt(S2);
  const double wx = dwdy - dvdz, wy = dudz - dwdx, wz = dvdx - dudy;
  Om = sqrt(wx * wx + wy * wy + wz * wz);
}
// bF2 (:514-521) and kw_tmu with bF3 = 1 (:768-771)
__device__ __forceinline__ double sst_F2(double k, double w, double y, double nu) {
  const double arg2 = dmin(dmax(2.0 * sqrt(k) / (0.09 * w * y), 500.0 * nu / (w * y * y)), 1000.0);
  return tanh(arg2 * arg2);
}
__device__ __forceinline__ double sst_tmu(double rho, double k, double w, double S, double F2) {
  return dmin(rho * SST_A1 * k / dmax(1e-30, dmax(SST_A1 * w, S * F2 * 1.0)), 1e5);
}
__device__ __forceinline__ void cell_gu(const double *__restrict__ grad, int nc, int c, double *gu) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gu[3 * i + j] = grad[(3 * (1 + i) + j) * nc + c];
}

struct SstCell {
  double F1, sigk, sigw, ksrc, wsrc, kJ, wJ;
};
// One cell: crossDiffusionTerm (:464-468), bF1 (:487-499), bF2, kw_tmu, the
// blended coefficients (:702-711), the SOmega stress (:790-796), P_k
// (:851-854), k and omega sources and destruction Jacobians (:870-904),
// cross diffusion (:910-913), Wilcox compressibility (:283-287, 960-1006);
// each source keeps the file's join order.
__device__ SstCell sst_core(const Prm &p, double T, double pg, double k, double w, const double *gu, double y,
                            double gkgw, double vol) {
  SstCell o;
  const double rho = (pg + p.Pamb) / (p.Rt * T);
  const double mu = suth_mu(p, T);
  const double nu = mu / rho;
  double S, divu, Om;
  sst_kinematics(gu, S, divu, Om);
  const double crossterm = dmax(SST_SIGW2 * gkgw, 0.0);
  const double CD = 2. * rho * crossterm / dmax(w, 1e-30);
  const double CDkw = dmax(1.0e-10, CD);
  const double ls = dmax(sqrt(k) / (0.09 * w * y), 500.0 * nu / (w * y * y));
  const double arg1 = dmin(dmin(ls, 4. * rho * SST_SIGW2 * k / (CDkw * y * y)), 32.0);
  const double arg1_2 = arg1 * arg1, arg1_4 = arg1_2 * arg1_2;
  const double F1 = tanh(arg1_4);
  const double F2 = sst_F2(k, w, y, nu);
  const double tmu = sst_tmu(rho, k, w, S, F2);
  const double beta = (F1 * SST_BETA1 + (1.0 - F1) * SST_BETA2);
  const double gama1 = SST_BETA1 / SST_BETAS - SST_SIGW1 * pow(SST_KAPPA, 2.0) / sqrt(SST_BETAS);
  const double gama2 = SST_BETA2 / SST_BETAS - SST_SIGW2 * pow(SST_KAPPA, 2.0) / sqrt(SST_BETAS);
  const double gama = F1 * gama1 + (1.0 - F1) * gama2;
  const double stress = dmax(tmu * (S * Om - 2. / 3. * divu * divu) - 2. / 3. * rho * k * divu, 0.0);
  const double BetaS = SST_BETAS * 1.0;
  const double Pk = dmin(stress, (1e100) * rho * BetaS * k * w);
  const double a = sqrt(p.gamma * p.Rt * T);
  const double Mt2 = 2.0 * k / (a * a);
  const double mt0 = 0.25;
  const double fmt = dmax(Mt2 - mt0 * mt0, 0.0);
  const double BetaF = SST_BETAS * 1.0 * 1.5 * fmt;
  const double nut = tmu / rho;
  const double Dw = (gama / (nut + 1e-30)) * stress;
  double ks = 0.0, ws = 0.0;
  ks += Pk * vol;
  ws += Dw * vol;
  ks -= BetaS * rho * w * k * vol;
  ws -= beta * rho * w * w * vol;
  ws += (1.0 - F1) * CD * vol;
  ks -= BetaF * rho * w * k * vol;
  ws -= -BetaF * rho * w * w * vol * 1.0;
  o.F1 = F1;
  o.sigk = F1 * SST_SIGK1 + (1.0 - F1) * SST_SIGK2;
  o.sigw = F1 * SST_SIGW1 + (1.0 - F1) * SST_SIGW2;
  o.ksrc = ks;
  o.wsrc = ws;
  o.kJ = -SST_BETAS * 1.0 * rho * w * vol;
  o.wJ = -2.0 * beta * rho * w * vol;
  return o;
}

// eddy viscosity: SA tmu = nuTilde fv1 rho (:93-107); SST kw_tmu
__global__ void k_tmu(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ grad,
                      double *__restrict__ tmu) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc;
  if (p.turb == 2) {
    const double T = q[c], pg = q[4 * nc + c], k = q[5 * nc + c], w = q[6 * nc + c];
    const double rho = (pg + p.Pamb) / (p.Rt * T);
    const double nu = suth_mu(p, T) / rho;
    double gu[9], S, divu, Om;
    cell_gu(grad, nc, c, gu);
    sst_kinematics(gu, S, divu, Om);
    tmu[c] = sst_tmu(rho, k, w, S, sst_F2(k, w, g.dist[c], nu));
    return;
  }
  const double T = q[c], pg = q[4 * nc + c], nut = q[5 * nc + c];
  const double rho = (pg + p.Pamb) / (p.Rt * T);
  const double x = rho * nut / suth_mu(p, T);
  tmu[c] = nut * sa_fv1(p, x) * rho;
}

// production - destruction and the destruction Jacobian of one cell
// (:109-165, 258-315, 347-351)
__device__ __forceinline__ void sa_core(const Prm &p, double T, double pg, double nut, double vortMag, double y,
                                        double vol, double *srcv, double *srcJ) {
  const double rho = (pg + p.Pamb) / (p.Rt * T);
  const double mu = suth_mu(p, T);
  const double xcoeff = rho * nut / mu;
  const double fv1 = sa_fv1(p, xcoeff);
  const double fv2 = 1.0 - (xcoeff / (1.0 + xcoeff * fv1));
  const double k = p.sa_k;
  const double denom = k * k * y * y;
  double S_tilda = 0.0;
  S_tilda += vortMag;
  S_tilda += nut * fv2 / denom;
  const double S_t = dmax(S_tilda, 0.3 * vortMag);
  const double rclip = 10;
  const double r = (S_t < nut / (rclip * denom)) ? rclip : nut / (S_t * denom);
  const double r2 = r * r, r6 = r2 * r2 * r2;
  const double gc = r + p.sa_cw2 * (r6 - r);
  const double cw36 = pow(p.sa_cw3, 6.0);
  const double g2 = gc * gc, g6 = g2 * g2 * g2;
  const double fw = gc * pow((1.0 + cw36) / (g6 + cw36), 1.0 / 6.0);
  const double cw1 = (p.sa_cb1) / (k * k) + (1.0 + (p.sa_cb2)) / (p.sa_sigma);
  double sv = 0.0;
  sv += rho * nut * S_t * p.sa_cb1 * vol;
  const double t = nut / y;
  sv -= rho * fw * cw1 * t * t * vol;
  double rp;
  const double rclip2 = 9e6;
  if (S_t < nut / (rclip2 * denom)) {
    rp = 0;
  } else {
    const double chi = xcoeff, chiP = 1. / mu;
    const double chi3 = chi * chi * chi, chi4 = chi3 * chi, chi6 = chi3 * chi3;
    const double cv1 = p.sa_cv1, cv13 = cv1 * cv1 * cv1;
    const double fv2p = chi * (4 * (chi3 * chiP) / (chi3 + cv13) - 3 * (chi6 * chiP) / pow(chi3 + cv13, 2.0)) /
                            pow(chi4 / (chi3 + cv13) + 1, 2.0) -
                        chiP / (chi4 / (chi3 + cv13) + 1);
    const double Stp = (fv2 / rho + nut * fv2p) / (denom);
    rp = 1. / (rho * denom * S_t) - nut * Stp / (denom * S_t * S_t);
  }
  const double gp = rp + p.sa_cw2 * (6. * pow(r, 5.0) * rp - rp);
  const double term = pow((cw36 + 1.) / (cw36 + g6), 1. / 6.);
  const double term2 = term * term, term5 = term2 * term2 * term;
  const double rterm5 = 1. / term5, rsum = 1. / (g6 + cw36);
  const double fwp = gp * term - gp * ((cw36 + 1) * g6) * rsum * rsum * rterm5;
  *srcv = sv;
  *srcJ = -(vol * cw1 * rho * nut * (rho * nut * fwp + 2. * fw) / (y * y));
}

// Per-face terms of the nuTilde system, saf[k][nf]:
//   0 convective flux mdot max(X_upwind, 0), 1 diff1_f, 2 div_nuTilde_f,
//   3 fjp, 4 fjm (interior/periodic), 5 srcJ_L, 6 srcJ_U, 7/8 cb2 diagonal
//   parts of cl / cr, 9 boundary srcJ of the interior cell;
// the scalar rows get U = fjm in row cl and L = -fjp in row cr
// (src/solverScalarTransport.loci:50-135,520-535; SA_turb.loci:354-482,737-745).
__global__ void k_sa_face(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ qf,
                          const double *__restrict__ grad, const double *__restrict__ phi,
                          const double *__restrict__ iflux, double *__restrict__ saf, double *__restrict__ soff) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nf) return;
  const int nc = g.nc, nf = g.nf, nb = g.nb;
  const int c = g.fcl[f];
  const double an[3] = {g.fnx[f], g.fny[f], g.fnz[f]};
  const double xf[3] = {g.fcx[f], g.fcy[f], g.fcz[f]};
  const double xl[3] = {g.ccx[c], g.ccy[c], g.ccz[c]};
  const double area = g.farea[f];
  const double mdot = iflux[f];
  const double sig = p.sa_sigma, cb2 = p.sa_cb2;
  const double factorx = 2, factory = .5;
  const double rhol = (q[4 * nc + c] + p.Pamb) / (p.Rt * q[c]);
  const double nul = q[5 * nc + c];
  auto recon6 = [&](int cc, const double *xc) {
    const double d0 = xf[0] - xc[0], d1 = xf[1] - xc[1], d2 = xf[2] - xc[2];
    const double gd = grad[15 * nc + cc] * d0 + grad[16 * nc + cc] * d1 + grad[17 * nc + cc] * d2;
    return q[5 * nc + cc] + phi[5 * nc + cc] * gd;
  };
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (f < g.bnd_begin) {
    const int o = g.frc[f];
    double xr[3] = {g.ccx[o], g.ccy[o], g.ccz[o]};
    if (f >= g.nfi) {
      const int pi = f - g.nfi;
      xr[0] += g.psx[pi]; xr[1] += g.psy[pi]; xr[2] += g.psz[pi];
    }
    const double XL = recon6(c, xl), XR = recon6(o, xr);
    const double Xval = (mdot < 0.0) ? XR : XL;
    v[0] = mdot * dmax(Xval, 0.0);
    const double vl = g.vol[c], vr = g.vol[o];
    const double nur = q[5 * nc + o];
    const double vrvl = 1.0 / (vr + vl);
    const double nuf = (vr * nul + vl * nur) * vrvl;
    const double dvec[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
    const double ddsq = dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2];
    double gfx[3];
    face_grad_int(g, q, grad, c, o, 5, dvec, ddsq, gfx);
    const double rhor = (q[4 * nc + o] + p.Pamb) / (p.Rt * q[o]);
    const double factor = 1. / (vl + vr);
    const double rho_f = factor * (vl * rhor + vr * rhol);
    const double mu_f = (vl * suth_mu(p, q[o]) + vr * suth_mu(p, q[c])) / (vl + vr);
    const double gdn = gfx[0] * an[0] + gfx[1] * an[1] + gfx[2] * an[2];
    const double d1 = gdn * area;
    v[1] = (d1 * (mu_f / rho_f + (1. + cb2) * nuf)) / (sig);
    v[2] = gdn * area;
    double fjp = 0.0, fjm = 0.0;
    fjp += dmax(mdot, 0.0);
    fjm += dmin(mdot, 0.0);
    const double rhom = .5 * (rhol + rhor);
    const double C = (mu_f + (1. + cb2) * rhom * nuf) * area;
    const double dd = dvec[0] * an[0] + dvec[1] * an[1] + dvec[2] * an[2];
    fjp -= -C / (dd * sig);
    fjm += -C / (dd * sig);
    const double C2 = cb2 * area;
    const double d2 = sig * dd;
    v[6] = -factorx * (C2 / d2) * rhol * nul;
    v[5] = -factorx * (C2 / d2) * rhor * nur;
    v[7] = factorx * factory * (C2 / d2) * rhol * nul;
    v[8] = factorx * factory * (C2 / d2) * rhor * nur;
    fjp += v[5];
    fjm -= v[6];
    v[3] = fjp;
    v[4] = fjm;
    soff[g.slot_cl[f]] = fjm;
    if (f < g.nfi) soff[g.slot_cr[f]] = -fjp;
  } else {
    const int b = f - g.bnd_begin;
    const double XL = recon6(c, xl);
    const double Xf = qf[5 * nb + b];
    const double Xval = (mdot < 0.0) ? Xf : XL;
    v[0] = mdot * dmax(Xval, 0.0);
    const double Tf = qf[b];
    const double rho_f = (qf[4 * nb + b] + p.Pamb) / (p.Rt * Tf);
    const double mu_f = suth_mu(p, Tf);
    const double dvec[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
    const double ddsq = dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2];
    double gfx[3];
    face_grad_bnd(g, q, grad, Xf, c, 5, dvec, ddsq, gfx);
    const double gdn = gfx[0] * an[0] + gfx[1] * an[1] + gfx[2] * an[2];
    const double d1 = gdn * area;
    v[1] = (d1 * (mu_f / rho_f + (1. + cb2) * Xf)) / (sig);
    v[2] = gdn * area;
    double bj = 0.0;
    bj -= dmax(mdot, 0.0);
    if ((g.fbt[f] & 0xFF) != 3) {
      const double C = (mu_f + (1. + cb2) * rho_f * Xf) * area;
      const double dd = dvec[0] * an[0] + dvec[1] * an[1] + dvec[2] * an[2];
      bj += -C / (dd * sig);
      const double C2 = cb2 * area;
      const double d2 = sig * dd;
      bj += factorx * factory * (C2 / d2) * rhol * nul;
    }
    v[9] = bj;
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) saf[(size_t)k * nf + f] = v[k];
}

// Per-cell gather of the nuTilde system in a fixed order (convection and
// continuity over lower, upper, boundary faces; stabilisation; SA sources;
// diffusion and its divergence; the cb2 term), srcJ, D = -srcJ + rho V/dtau,
// B = src, 1/D and the Jacobi start X = B/D (src/linearSolverSGS.loci:343-352).
// D and B of a scalar row from its source and Jacobian sums: steady
// (src/solverScalarTransport.loci:584-590, 680-690) or BDF2 with the density
// of time levels n and n - 1 (:602-616, :702-716)
__device__ __forceinline__ void scalar_time(const Prm &p, double vol, double dtau, double rho, double X,
                                            const double *__restrict__ qn, const double *__restrict__ qm1, int nc,
                                            int c, int kv, double sv, double sJ, double &D, double &b) {
  if (p.unsteady == 2) {  // KEC, :592-600 and :691-700
    const double rn = (qn[4 * nc + c] + p.Pamb) / (p.Rt * qn[c]);
    const double factor = vol / p.dtmax;
    const double Xn = qn[(size_t)kv * nc + c];
    const double Xnp1 = kec_xnp1(X, Xn);
    const double rhs = factor * ((Xnp1 - Xn) * rn) - sv;
    b = -rhs;
    const double sd = (2. * rn) * (vol / dtau);
    D = -sJ + sd;
    return;
  }
  if (!p.unsteady) {
    const double sd = rho * (vol / dtau);
    D = -sJ + sd;
    const double rhs = 0.0 - sv;
    b = -rhs;
    return;
  }
  const double rn = (qn[4 * nc + c] + p.Pamb) / (p.Rt * qn[c]);
  const double rnm1 = (qm1[4 * nc + c] + p.Pamb) / (p.Rt * qm1[c]);
  const double Xn = qn[(size_t)kv * nc + c], Xm = qm1[(size_t)kv * nc + c];
  const double factor = vol / p.dtmax;
  const double w = 0.5;
  const double w1 = (1. + w) * factor, w2 = -w * factor;
  const double rhs = w1 * (X - Xn) * rn + w2 * (Xn * rn - Xm * rnm1) - w2 * X * (rn - rnm1) - sv;
  b = -rhs;
  const double fd = vol / dtau;
  const double sd = (1. + w) * fd * rn + w * fd * (rn - rnm1);
  D = -sJ + sd;
}

__global__ void k_sa_cell(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ grad,
                          const double *__restrict__ dtau, const double *__restrict__ iflux,
                          const double *__restrict__ saf, double *__restrict__ ssrc, double *__restrict__ sdinv,
                          double *__restrict__ sB, double *__restrict__ sX, const double *__restrict__ qn,
                          const double *__restrict__ qm1) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc, nf = g.nf;
  const double rho = (q[4 * nc + c] + p.Pamb) / (p.Rt * q[c]);
  const double X = q[5 * nc + c];
  const int j0 = g.g_off[c], j1 = g.g_off[c + 1];
  double sv = 0.0, sJ = 0.0, rhodot = 0.0;
  // g_ent order is lower faces (e >= 0), upper faces, boundary faces
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e >= 0) {
      sv += saf[e];
      rhodot += iflux[e];
    } else {
      const int f = ~e;
      sv -= saf[f];
      rhodot -= iflux[f];
    }
  }
  // stabilisation, src/solverScalarTransport.loci:508-546
  sv -= (p.unsteady == 2 ? kec_xnp1(X, qn[5 * nc + c]) : X) * rhodot;
  sJ += p.unsteady == 2 ? -2. * rhodot : p.unsteady ? -1.5 * rhodot : -rhodot;
#define GU(i, jj) grad[(3 * (1 + (i)) + (jj)) * nc + c]
  const double wx = GU(2, 1) - GU(1, 2), wy = GU(0, 2) - GU(2, 0), wz = GU(1, 0) - GU(0, 1);
#undef GU
  const double vortMag = sqrt(wx * wx + wy * wy + wz * wz);
  double ssa, sdJ;
  sa_core(p, q[c], q[4 * nc + c], X, vortMag, g.dist[c], g.vol[c], &ssa, &sdJ);
  sv += ssa;
  sJ += sdJ;
  double div = 0.0;
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e >= 0) {
      sv -= rho * saf[(size_t)nf + e];
      div -= saf[(size_t)2 * nf + e];
    } else {
      const int f = ~e;
      sv += rho * saf[(size_t)nf + f];
      div += saf[(size_t)2 * nf + f];
    }
  }
  sv += -rho * X * div * p.sa_cb2 / p.sa_sigma;
  // srcJ: cb2 diagonal parts (lower: cr part, upper: cl part), boundary terms,
  // then -fjp over upper and +fjm over lower faces, then the L/U cb2 parts
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e >= 0) sJ += saf[(size_t)8 * nf + e];
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e < 0 && ~e < g.bnd_begin) sJ += saf[(size_t)7 * nf + ~e];
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e < 0 && ~e >= g.bnd_begin) sJ += saf[(size_t)9 * nf + ~e];
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e < 0 && ~e < g.bnd_begin) sJ -= saf[(size_t)3 * nf + ~e];
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e >= 0) sJ += saf[(size_t)4 * nf + e];
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e < 0 && ~e < g.bnd_begin) sJ += saf[(size_t)5 * nf + ~e];
  }
  for (int j = j0; j < j1; ++j) {
    const int e = g.g_ent[j];
    if (e >= 0) sJ += saf[(size_t)6 * nf + e];
  }
  ssrc[c] = sv;
  double D, b;
  scalar_time(p, g.vol[c], dtau[c], rho, X, qn, qm1, nc, c, 5, sv, sJ, D, b);
  const double dinv = 1. / D;
  sB[c] = b;
  sdinv[c] = dinv;
  sX[c] = dinv * b;
}

// SST per-face pass 1, saf rows 0-1: area grad_f(k).n and its omega-weighted
// form (w_f = max(0, mean of the reconstructions) on interior faces, the
// face value on boundary faces), sst.loci:427-462
__device__ __forceinline__ double recon_k(const double *__restrict__ q, const double *__restrict__ grad,
                                          const double *__restrict__ phi, int nc, int k, int cc, const double *xc,
                                          const double *xf) {
  const double d0 = xf[0] - xc[0], d1 = xf[1] - xc[1], d2 = xf[2] - xc[2];
  const double gd = grad[(3 * k) * nc + cc] * d0 + grad[(3 * k + 1) * nc + cc] * d1 + grad[(3 * k + 2) * nc + cc] * d2;
  return q[k * nc + cc] + phi[k * nc + cc] * gd;
}

__global__ void k_sst_face_a(DevGrid g, const double *__restrict__ q, const double *__restrict__ qf,
                             const double *__restrict__ grad, const double *__restrict__ phi,
                             double *__restrict__ saf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nf) return;
  const int nc = g.nc, nf = g.nf, nb = g.nb;
  const int c = g.fcl[f];
  const double an[3] = {g.fnx[f], g.fny[f], g.fnz[f]};
  const double xf[3] = {g.fcx[f], g.fcy[f], g.fcz[f]};
  const double xl[3] = {g.ccx[c], g.ccy[c], g.ccz[c]};
  double gk[3];
  if (f < g.bnd_begin) {
    const int o = g.frc[f];
    double xr[3] = {g.ccx[o], g.ccy[o], g.ccz[o]};
    if (f >= g.nfi) {
      const int pi = f - g.nfi;
      xr[0] += g.psx[pi]; xr[1] += g.psy[pi]; xr[2] += g.psz[pi];
    }
    const double dvec[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
    const double ddsq = dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2];
    face_grad_int(g, q, grad, c, o, 5, dvec, ddsq, gk);
    const double gkn = g.farea[f] * (gk[0] * an[0] + gk[1] * an[1] + gk[2] * an[2]);
    const double wL = recon_k(q, grad, phi, nc, 6, c, xl, xf), wR = recon_k(q, grad, phi, nc, 6, o, xr, xf);
    const double w_f = dmax(0.0, 0.5 * (wL + wR));
    saf[f] = gkn;
    saf[(size_t)nf + f] = w_f * gkn;
  } else {
    const int b = f - g.bnd_begin;
    const double dvec[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
    const double ddsq = dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2];
    face_grad_bnd(g, q, grad, qf[5 * nb + b], c, 5, dvec, ddsq, gk);
    const double gkn = g.farea[f] * (gk[0] * an[0] + gk[1] * an[1] + gk[2] * an[2]);
    saf[f] = gkn;
    saf[(size_t)nf + f] = qf[6 * nb + b] * gkn;
  }
}

// SST per-cell pass 1: grad k . grad omega = (div(w grad k) - w lap k) / vol
// gathered in the lower/upper/boundary order, then sst_core into
// scell[8][nc]: F1, grad k.grad w, sigma_k, sigma_w, k and omega sources,
// k and omega Jacobians
__global__ void k_sst_cell_a(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ grad,
                             const double *__restrict__ saf, double *__restrict__ scell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc, nf = g.nf;
  double lap = 0.0, dwg = 0.0;
  for (int j = g.g_off[c]; j < g.g_off[c + 1]; ++j) {
    const int e = g.g_ent[j];
    if (e >= 0) {
      lap += -saf[e];
      dwg += -saf[(size_t)nf + e];
    } else {
      lap += saf[~e];
      dwg += saf[(size_t)nf + ~e];
    }
  }
  const double w = q[6 * nc + c];
  const double gkgw = (dwg - w * lap) / g.vol[c];
  double gu[9];
  cell_gu(grad, nc, c, gu);
  const SstCell o = sst_core(p, q[c], q[4 * nc + c], q[5 * nc + c], w, gu, g.dist[c], gkgw, g.vol[c]);
  scell[c] = o.F1;
  scell[(size_t)nc + c] = gkgw;
  scell[2 * (size_t)nc + c] = o.sigk;
  scell[3 * (size_t)nc + c] = o.sigw;
  scell[4 * (size_t)nc + c] = o.ksrc;
  scell[5 * (size_t)nc + c] = o.wsrc;
  scell[6 * (size_t)nc + c] = o.kJ;
  scell[7 * (size_t)nc + c] = o.wJ;
}

// SST per-face pass 2, for k (saf rows 2-6) and omega (rows 7-11): upwind
// convection of max(X, 0), diffusion X_nu_f grad_f(X).n area with X_nu_f =
// mu_f + tmu_f sigma_f, fjp, fjm (also into the scalar rows soff), and the
// boundary Jacobian (src/solverScalarTransport.loci:121-135, 548-567, 779-826)
__global__ void __launch_bounds__(TPB) k_sst_face_b(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ qf,
                             const double *__restrict__ grad, const double *__restrict__ phi,
                             const double *__restrict__ iflux, const double *__restrict__ tmu,
                             const double *__restrict__ scell, double *__restrict__ saf, double *__restrict__ soff,
                             long nnz) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nf) return;
  const int nc = g.nc, nf = g.nf, nb = g.nb;
  const int c = g.fcl[f];
  const double an[3] = {g.fnx[f], g.fny[f], g.fnz[f]};
  const double xf[3] = {g.fcx[f], g.fcy[f], g.fcz[f]};
  const double xl[3] = {g.ccx[c], g.ccy[c], g.ccz[c]};
  const double area = g.farea[f];
  const double mdot = iflux[f];
  const double tmf = tmu_face(g, tmu, f);
  for (int sc = 0; sc < 2; ++sc) {
    const int kv = 5 + sc;
    const double sl = scell[(2 + sc) * (size_t)nc + c];
    double v[5] = {0, 0, 0, 0, 0};
    if (f < g.bnd_begin) {
      const int o = g.frc[f];
      double xr[3] = {g.ccx[o], g.ccy[o], g.ccz[o]};
      if (f >= g.nfi) {
        const int pi = f - g.nfi;
        xr[0] += g.psx[pi]; xr[1] += g.psy[pi]; xr[2] += g.psz[pi];
      }
      const double XL = recon_k(q, grad, phi, nc, kv, c, xl, xf), XR = recon_k(q, grad, phi, nc, kv, o, xr, xf);
      const double Xval = (mdot < 0.0) ? XR : XL;
      v[0] = mdot * dmax(Xval, 0.0);
      const double vl = g.vol[c], vr = g.vol[o];
      const double sr = scell[(2 + sc) * (size_t)nc + o];
      const double rvlvr = 1. / (vl + vr);
      const double sig_f = (vr * sl + vl * sr) * rvlvr;
      const double mu_f = (vl * suth_mu(p, q[o]) + vr * suth_mu(p, q[c])) / (vl + vr);
      const double nu_f = (mu_f + tmf * sig_f);
      const double dvec[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
      const double ddsq = dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2];
      double gx[3];
      face_grad_int(g, q, grad, c, o, kv, dvec, ddsq, gx);
      v[1] = nu_f * (gx[0] * an[0] + gx[1] * an[1] + gx[2] * an[2]) * area;
      const double C = nu_f * area;
      const double d = dvec[0] * an[0] + dvec[1] * an[1] + dvec[2] * an[2];
      double fjp = 0.0, fjm = 0.0;
      fjp += dmax(mdot, 0.0);
      fjm += dmin(mdot, 0.0);
      fjp += C / d;
      fjm += -C / d;
      v[2] = fjp;
      v[3] = fjm;
      soff[(size_t)sc * nnz + g.slot_cl[f]] = fjm;
      if (f < g.nfi) soff[(size_t)sc * nnz + g.slot_cr[f]] = -fjp;
    } else {
      const int b = f - g.bnd_begin;
      const double XL = recon_k(q, grad, phi, nc, kv, c, xl, xf);
      const double Xf = qf[kv * nb + b];
      const double Xval = (mdot < 0.0) ? Xf : XL;
      v[0] = mdot * dmax(Xval, 0.0);
      const double mu_f = suth_mu(p, qf[b]);
      const double nu_f = (mu_f + tmf * sl);
      const double dvec[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
      const double ddsq = dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2];
      double gx[3];
      face_grad_bnd(g, q, grad, Xf, c, kv, dvec, ddsq, gx);
      v[1] = nu_f * (gx[0] * an[0] + gx[1] * an[1] + gx[2] * an[2]) * area;
      const double C = nu_f * area;
      const double d = dvec[0] * an[0] + dvec[1] * an[1] + dvec[2] * an[2];
      double bj = 0.0;
      bj -= dmax(mdot, 0.0);
      bj += -C / d;
      v[4] = bj;
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) saf[(size_t)(2 + 5 * sc + r) * nf + f] = v[r];
  }
}

// SST per-cell pass 2: the k and omega rows in a fixed order (convection
// lower/upper/boundary, stabilisation -X rhodot, cell sources, diffusion
// lower/upper/boundary; Jacobian: stabilisation, cell, boundary faces,
// upper fjp, lower fjm), D = -srcJ + rho V/dtau, B = src, Jacobi start
__global__ void k_sst_cell_b(DevGrid g, Prm p, const double *__restrict__ q, const double *__restrict__ dtau,
                             const double *__restrict__ iflux, const double *__restrict__ saf,
                             const double *__restrict__ scell, double *__restrict__ ssrc, double *__restrict__ sdinv,
                             double *__restrict__ sB, double *__restrict__ sX, const double *__restrict__ qn,
                             const double *__restrict__ qm1) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc, nf = g.nf;
  const size_t xs = (size_t)nc + g.nper + 1;
  const double rho = (q[4 * nc + c] + p.Pamb) / (p.Rt * q[c]);
  const int j0 = g.g_off[c], j1 = g.g_off[c + 1];
  for (int sc = 0; sc < 2; ++sc) {
    const double *F = saf + (size_t)(2 + 5 * sc) * nf;
    const double X = q[(5 + sc) * nc + c];
    double sv = 0.0, sJ = 0.0, rhodot = 0.0;
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      if (e >= 0) {
        sv += F[e];
        rhodot += iflux[e];
      } else {
        sv -= F[~e];
        rhodot -= iflux[~e];
      }
    }
    sv -= (p.unsteady == 2 ? kec_xnp1(X, qn[(5 + sc) * nc + c]) : X) * rhodot;
    sJ += p.unsteady == 2 ? -2. * rhodot : p.unsteady ? -1.5 * rhodot : -rhodot;
    sv += scell[(4 + sc) * (size_t)nc + c];
    sJ += scell[(6 + sc) * (size_t)nc + c];
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      if (e >= 0) sv -= F[(size_t)nf + e];
      else sv += F[(size_t)nf + ~e];
    }
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      if (e < 0 && ~e >= g.bnd_begin) sJ += F[(size_t)4 * nf + ~e];
    }
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      if (e < 0 && ~e < g.bnd_begin) sJ -= F[(size_t)2 * nf + ~e];
    }
    for (int j = j0; j < j1; ++j) {
      const int e = g.g_ent[j];
      if (e >= 0) sJ += F[(size_t)3 * nf + e];
    }
    ssrc[(size_t)sc * nc + c] = sv;
    double D, b;
    scalar_time(p, g.vol[c], dtau[c], rho, X, qn, qm1, nc, c, 5 + sc, sv, sJ, D, b);
    const double dinv = 1. / D;
    sB[(size_t)sc * nc + c] = b;
    sdinv[(size_t)sc * nc + c] = dinv;
    sX[sc * xs + c] = dinv * b;
  }
}

// one level of a scalar SGS half-sweep, one lane per cell (the row holds
// U = fjm for upper faces, then L = -fjp for lower faces)
__global__ void k_sa_sgs_level(DevGrid g, const int *__restrict__ cells, int lo, int hi, int sc0, int ns, long nnz,
                               const double *__restrict__ soff, const double *__restrict__ sdinv,
                               const double *__restrict__ sB, double *__restrict__ sX) {
  // lanes [0, w) scalar sc0, [w, 2w) scalar sc0 + 1: the systems are independent
  const int w = hi - lo;
  const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= w * ns) return;
  const int sc = sc0 + t0 / w;
  const int k = lo + (t0 - (sc - sc0) * w);
  const int c = cells ? cells[k] : k;
  const int nc = g.nc;
  const size_t xs = (size_t)nc + g.nper + 1;
  const double *off = soff + (size_t)sc * nnz;
  double *X = sX + (size_t)sc * xs;
  double t = -sB[(size_t)sc * nc + c];
  for (int j = g.m_off[c]; j < g.m_off[c + 1]; ++j) t += X[g.m_nbr[j]] * off[j];
  const double v = -t;
  X[c] = sdinv[(size_t)sc * nc + c] * v;
}

// one level of a scalar line-SGS sweep (src/linearSolverLSGS.loci:1240-1317):
// X_c = (B_c - lower couplings - upper couplings) / D_c, the row rotated to
// lower-first; the forward sweep keeps the value it replaces in xt
template <bool BACK>
__global__ void k_sa_lsgs_level(DevGrid g, const int *__restrict__ cells, int lo, int hi,
                                const double *__restrict__ soff, const double *__restrict__ sdinv,
                                const double *__restrict__ sB, double *__restrict__ sX, double *__restrict__ sxt) {
  const int k = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= hi) return;
  const int c = cells ? cells[k] : k;
  double r = sB[c];
  const int m0 = g.m_off[c], len = g.m_off[c + 1] - m0, nup = g.m_nup[c];
  for (int j = 0; j < len; ++j) {
    int jr = j + nup;
    if (jr >= len) jr -= len;
    r -= sX[g.m_nbr[m0 + jr]] * soff[m0 + jr];
  }
  if constexpr (!BACK) sxt[c] = sX[c];
  sX[c] = sdinv[c] * r;
}

// X_next = X + min(1, urelax X/(|dX| + 1e-15)) dX, floored at 1e-15
// (src/solverScalarTransport.loci:400-408)
__global__ void k_sa_update(DevGrid g, Prm p, int ns, const double *__restrict__ sX, double *__restrict__ q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const size_t xs = (size_t)g.nc + g.nper + 1;
  for (int sc = 0; sc < ns; ++sc) {
    const double Xc = q[(5 + sc) * (size_t)g.nc + c], dX = sX[sc * xs + c];
    double tscale = 1.0;
    tscale = dmin(tscale, p.urelax * Xc / (fabs(dX) + 1e-15));
    const double Xn = Xc + tscale * dX;
    q[(5 + sc) * (size_t)g.nc + c] = dmax(Xn, 1e-15);
  }
}

// damped Newton update, src/solverNewton.loci:62-78 (steady), :160-180
// (BDF2: at most 75 % of the current pressure and temperature) and :93-125
// (KEC: limits from the level n+1 reconstruction)
__global__ void k_update(DevGrid g, Prm p, const double *__restrict__ X, double *__restrict__ q,
                         const double *__restrict__ qn, const double *__restrict__ pbias) {
  const double Pamb = p.Pamb;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc;
  const double *dq = X + (size_t)c * 5;
  const double dT = dq[0], dp = dq[4];
  double alpha = 1.0;
  const double T = q[c], pg = q[4 * nc + c];
  if (p.unsteady == 1) {
    const double relax = 0.75;
    const double pnit = pg + Pamb;
    alpha = dmin(alpha, relax * pnit / dmax(fabs(dp), 1e-30));
    alpha = dmin(alpha, relax * T / dmax(fabs(dT), 1e-30));
  } else if (p.unsteady == 2) {
    const double relax = 0.8;
    const double Tn = qn[c];
    const double dTlim = (relax * (2 * Tn - T)) / 2.0;
    alpha = dmin(alpha, dTlim / dmax(fabs(dT), 1e-30));
    const double peps = pbias[c];
    const double pgnp1 = (2. * pg - (1. - peps) * qn[4 * nc + c]) / (1 + peps);
    const double dplim = (1 + peps) * relax * (pgnp1 + Pamb) / 2.0;
    alpha = dmin(alpha, dplim / dmax(fabs(dp), 1e-30));
  } else {
    alpha = dmin(alpha, 0.9 * T / (fabs(dT) + 1e-30));
    alpha = dmin(alpha, 0.9 * (pg + Pamb) / (fabs(dp) + 1e-30));
  }
  q[c] = T + alpha * dT;
  q[4 * nc + c] = pg + alpha * dp;
  q[nc + c] = q[nc + c] + alpha * dq[1];
  q[2 * nc + c] = q[2 * nc + c] + alpha * dq[2];
  q[3 * nc + c] = q[3 * nc + c] + alpha * dq[3];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_bface(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  if (g.nb) hipLaunchKernelGGL(k_bface, dim3(nblk(g.nb)), dim3(TPB), 0, st, g, p, s.q, s.qf);
}
void launch_grad(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  if (p.turb == 2) hipLaunchKernelGGL(k_grad<7>, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, s.q, s.qf, s.grad, s.phi);
  else if (p.turb) hipLaunchKernelGGL(k_grad<6>, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, s.q, s.qf, s.grad, s.phi);
  else hipLaunchKernelGGL(k_grad<5>, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, s.q, s.qf, s.grad, s.phi);
}
void launch_eta(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  if (!p.pre) return;
  hipLaunchKernelGGL(k_eta_l, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, s.q, (const double *)s.tmu, s.eta);
  hipLaunchKernelGGL(k_eta_max, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, (const double *)s.eta, s.eta_t);
  hipLaunchKernelGGL(k_eta_max, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, (const double *)s.eta_t, s.eta);
}
void launch_face_flux(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  hipLaunchKernelGGL(k_face_flux, dim3(nblk(g.nf)), dim3(TPB), 0, st, g, p, s.q, s.qf, s.grad, s.phi,
                     (const double *)s.eta, (const double *)s.tmu, s.iflux, s.vflux, s.maxev);
}
int norm_parts(const DevGrid &g) { return nblk(g.nc); }
void launch_rhs_norms(const DevGrid &g, double vtot, DevState &s, hipStream_t st) {
  hipLaunchKernelGGL(k_rhs_norms, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, vtot, (const double *)s.rhs, s.npart);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(TPB), 0, st, g, nblk(g.nc), s.npart, s.nout);
}
void launch_gather(const DevGrid &g, const Prm &p, double vtot, DevState &s, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_src, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p.viscous, vtot, s.iflux,
                     s.vflux, s.src, s.npart);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(TPB), 0, st, g, nblk(g.nc), s.npart, s.nout);
}
void launch_face_jac(const DevGrid &g, const Prm &p, int fp32, DevState &s, hipStream_t st) {
  if (fp32)
    hipLaunchKernelGGL(k_face_jac<float>, dim3(nblk(g.nf)), dim3(TPB), 0, st, g, p, s.q, s.qf, s.grad,
                       s.phi, s.vflux, (const double *)s.eta, (const double *)s.tmu, (float *)s.off,
                       (float *)s.pblk, (float *)s.bblk);
  else
    hipLaunchKernelGGL(k_face_jac<double>, dim3(nblk(g.nf)), dim3(TPB), 0, st, g, p, s.q, s.qf, s.grad,
                       s.phi, s.vflux, (const double *)s.eta, (const double *)s.tmu, (double *)s.off,
                       (double *)s.pblk, (double *)s.bblk);
}
template <typename F, bool LS, int TS>
static void cell_assemble_ts(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  hipLaunchKernelGGL((k_cell_assemble<F, LS, TS>), dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, s.q, s.src, s.maxev,
                     (const F *)s.off, (const F *)s.pblk, (const F *)s.bblk, s.dtau, (F *)s.lu, s.piv,
                     (F *)s.B, s.X, s.lu64, (const double *)s.eta, (const double *)s.qn, (const double *)s.qm1,
                     s.rhs, s.pbias);
}
template <typename F, bool LS>
static void cell_assemble_t(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  if (p.unsteady == 2) cell_assemble_ts<F, LS, 2>(g, p, s, st);
  else if (p.unsteady == 1) cell_assemble_ts<F, LS, 1>(g, p, s, st);
  else cell_assemble_ts<F, LS, 0>(g, p, s, st);
}
void launch_cell_assemble(const DevGrid &g, const Prm &p, int fp32, int lsgs, DevState &s, hipStream_t st) {
  if (fp32) {
    if (lsgs) cell_assemble_t<float, true>(g, p, s, st);
    else cell_assemble_t<float, false>(g, p, s, st);
  } else {
    if (lsgs) cell_assemble_t<double, true>(g, p, s, st);
    else cell_assemble_t<double, false>(g, p, s, st);
  }
}
void launch_sgs_level(const DevGrid &g, int fp32, int lo, int hi, DevState &s, hipStream_t st) {
  const int n = hi - lo;
  if (n <= 0) return;
  const int nb = (n + CPW - 1) / CPW;
  if (fp32)
    hipLaunchKernelGGL(k_sgs_level<float>, dim3(nb), dim3(64), 0, st, g, lo, hi,
                       (const float *)s.off, (const float *)s.lu, s.piv, (const float *)s.B, s.X);
  else
    hipLaunchKernelGGL(k_sgs_level<double>, dim3(nb), dim3(64), 0, st, g, lo, hi,
                       (const double *)s.off, (const double *)s.lu, s.piv, (const double *)s.B, s.X);
}
int sgs_sweep_waves(int fp32, int ncu) {
  int per_cu = 0, a = 0, b = 0;
  hipError_t e = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_sgs_sweep<float, false>, 64, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_sgs_sweep<double, false>, 64, 0);
  if (e != hipSuccess) return 0;
  e = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_sgs_sweep<float, true>, 64, 0)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_sgs_sweep<double, true>, 64, 0);
  if (e != hipSuccess) return 0;
  per_cu = std::min(a, b);
  // the flag-free sweep runs with the same wave count
  e = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_sgs_flow<float, false>, 64, 0)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_sgs_flow<double, false>, 64, 0);
  if (e != hipSuccess) return 0;
  e = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_sgs_flow<float, true>, 64, 0)
           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_sgs_flow<double, true>, 64, 0);
  if (e != hipSuccess) return 0;
  per_cu = std::min(per_cu, std::min(a, b));
  if (per_cu <= 0) return 0;
  return per_cu * ncu;
}
template <typename F, typename DG>
static int lsgs_waves_t(int ncu) {
  int a = 0, b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_lsgs_sweep<F, DG, false>, 64, 0) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_lsgs_sweep<F, DG, true>, 64, 0) != hipSuccess) return 0;
  const int per_cu = std::min(a, b);
  return per_cu > 0 ? per_cu * ncu : 0;
}
int lsgs_sweep_waves(int fp32, int ncu) {
  const int a = fp32 ? lsgs_waves_t<float, double>(ncu) : lsgs_waves_t<double, double>(ncu);
  const int b = fp32 ? lsgs_waves_t<float, float>(ncu) : a;
  return std::min(a, b);
}
template <typename F>
static int lsgs_flow_waves_t(int ncu) {
  int a = 0, b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_lsgs_flow<F, double, false>, 64, 0) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_lsgs_flow<F, double, true>, 64, 0) != hipSuccess) return 0;
  const int per_cu = std::min(a, b);
  return per_cu > 0 ? per_cu * ncu : 0;
}
int lsgs_flow_waves(int fp32, int ncu) { return fp32 ? lsgs_flow_waves_t<float>(ncu) : lsgs_flow_waves_t<double>(ncu); }
hipError_t launch_lsgs_flow(const DevGrid &g, int fp32, int backward, int nchunks, const int *chunk_lo, int spin_limit,
                            int *err, int nwaves, int fluid, int ns, int smask, long nnz, DevState &s,
                            hipStream_t st) {
  int nw = std::max(1, std::min(nwaves, nchunks));
  LineScalars ls{ns, smask, nnz, s.soff, s.sdinv, s.sB, s.sX, s.sxt};
  double *sXb = s.sX2;
  if (!smask) {
    ls = LineScalars{0, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
    sXb = nullptr;
  }
  void *args[] = {(void *)&g, &nchunks, &nw, &chunk_lo, &spin_limit, &err, (void *)&s.off, (void *)&s.lu64,
                  (void *)&s.B, (void *)&s.X, (void *)&s.X2, &fluid, (void *)&ls, (void *)&sXb};
  const void *fn;
  if (fp32) fn = backward ? (const void *)k_lsgs_flow<float, double, true> : (const void *)k_lsgs_flow<float, double, false>;
  else fn = backward ? (const void *)k_lsgs_flow<double, double, true> : (const void *)k_lsgs_flow<double, double, false>;
  return hipLaunchCooperativeKernel(fn, dim3(nw), dim3(64), args, 0, st);
}
void launch_lsgs_flow_fill(const DevGrid &g, DevState &s, int fluid, int ns, hipStream_t st) {
  const size_t xs = (size_t)g.nc + g.nper + 1;
  if (fluid) {
    const long n = 5L * g.nc;
    hipLaunchKernelGGL(k_flow_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, s.X2, 5L,
                       s.X2 + 5 * (xs - 1));
  }
  for (int sc = 0; sc < ns; ++sc) {
    const long n = g.nc;
    hipLaunchKernelGGL(k_flow_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, s.sX2 + sc * xs, 1L,
                       s.sX2 + sc * xs + (xs - 1));
  }
}
void launch_lsgs_relax_flow(const DevGrid &g, int iter, double abstol, double reltol, DevState &s, hipStream_t st) {
  const int n = g.nc * 5, np = lsgs_parts(g);
  hipLaunchKernelGGL(k_lsgs_relax_flow, dim3(np), dim3(TPB), 0, st, n, (const double *)s.lst, s.X,
                     (const double *)s.X2, s.X2, s.lparts);
  hipLaunchKernelGGL(k_lsgs_check, dim3(1), dim3(TPB), 0, st, np, (const double *)s.lparts, s.lst, iter, abstol,
                     reltol);
}
void launch_sa_lsgs_relax_flow(const DevGrid &g, int sc, int iter, double abstol, double reltol, DevState &s,
                               hipStream_t st, int slot) {
  const int np = nblk(g.nc);
  const size_t nc = g.nc, xs = nc + g.nper + 1;
  double *lst = s.lst + 8 * slot, *parts = s.lparts + (size_t)slot * lsgs_parts(g);
  hipLaunchKernelGGL(k_lsgs_relax_flow, dim3(np), dim3(TPB), 0, st, g.nc, (const double *)lst, s.sX + sc * xs,
                     (const double *)(s.sxt + sc * nc), s.sX2 + sc * xs, parts);
  hipLaunchKernelGGL(k_lsgs_check, dim3(1), dim3(TPB), 0, st, np, (const double *)parts, lst, iter, abstol, reltol);
}
hipError_t launch_lsgs_sweep(const DevGrid &g, int fp32, int fsgs, int backward, int nchunks, const int *chunk_lo,
                             const int *dep_off, const int *dep, int *flag, int epoch, int spin_limit, int *err,
                             int nwaves, int fluid, int ns, int smask, long nnz, DevState &s, hipStream_t st) {
  int nw = std::max(1, std::min(nwaves, nchunks));
  const void *diag = fsgs ? (const void *)s.luf : (const void *)s.lu64;
  double *xt = fsgs ? nullptr : s.xt;
  LineScalars ls{ns, smask, nnz, s.soff, s.sdinv, s.sB, s.sX, s.sxt};
  if (!smask) ls = LineScalars{0, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
  void *args[] = {(void *)&g, &nchunks, &nw, &chunk_lo, &dep_off, &dep, &flag, &epoch, &spin_limit, &err,
                  (void *)&s.off, (void *)&diag, (void *)&s.B, (void *)&s.X, (void *)&xt, &fluid, (void *)&ls};
  const void *fn;
  if (fp32) {
    if (fsgs) fn = backward ? (const void *)k_lsgs_sweep<float, float, true> : (const void *)k_lsgs_sweep<float, float, false>;
    else fn = backward ? (const void *)k_lsgs_sweep<float, double, true> : (const void *)k_lsgs_sweep<float, double, false>;
  } else {
    fn = backward ? (const void *)k_lsgs_sweep<double, double, true> : (const void *)k_lsgs_sweep<double, double, false>;
  }
  return hipLaunchCooperativeKernel(fn, dim3(nw), dim3(64), args, 0, st);
}
hipError_t launch_sgs_sweep(const DevGrid &g, int fp32, int backward, int nchunks, const int *chunk_lo,
                            const int *dep_off, const int *dep, int *flag, int epoch, int spin_limit, int *err,
                            int nwaves, int ns, long nnz, DevState &s, hipStream_t st) {
  int nw = std::max(1, std::min(nwaves, nchunks));
  ScalarRows sr{ns, nnz, s.soff, s.sdinv, s.sB, s.sX};
  if (!ns) sr = ScalarRows{0, 0, nullptr, nullptr, nullptr, nullptr};
  void *args[] = {(void *)&g, &nchunks, &nw, &chunk_lo, &dep_off, &dep, &flag, &epoch, &spin_limit, &err,
                  (void *)&s.off, (void *)&s.lu, (void *)&s.piv, (void *)&s.B, (void *)&s.X, (void *)&sr};
  const void *fn = fp32 ? (backward ? (const void *)k_sgs_sweep<float, true> : (const void *)k_sgs_sweep<float, false>)
                        : (backward ? (const void *)k_sgs_sweep<double, true> : (const void *)k_sgs_sweep<double, false>);
  return hipLaunchCooperativeKernel(fn, dim3(nw), dim3(64), args, 0, st);
}
hipError_t launch_sgs_flow(const DevGrid &g, int fp32, int backward, int nchunks, const int *chunk_lo, int spin_limit,
                           int *err, int nwaves, DevState &s, double *Xs, double *Xt, hipStream_t st) {
  int nw = std::max(1, std::min(nwaves, nchunks));
  void *args[] = {(void *)&g, &nchunks, &nw, &chunk_lo, &spin_limit, &err, (void *)&s.off, (void *)&s.lu,
                  (void *)&s.piv, (void *)&s.B, (void *)&Xs, (void *)&Xt};
  const void *fn = fp32 ? (backward ? (const void *)k_sgs_flow<float, true> : (const void *)k_sgs_flow<float, false>)
                        : (backward ? (const void *)k_sgs_flow<double, true> : (const void *)k_sgs_flow<double, false>);
  return hipLaunchCooperativeKernel(fn, dim3(nw), dim3(64), args, 0, st);
}
void launch_sweep_prep(const DevGrid &g, int fill, double *Xs, double *Xt, hipStream_t st) {
  const long n = 5L * (fill ? std::max(g.nc, g.nper) : g.nper);
  if (n > 0) hipLaunchKernelGGL(k_sweep_prep, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, g, fill, Xs, Xt);
}
void launch_ghost(const DevGrid &g, DevState &s, hipStream_t st) {
  if (g.nper) hipLaunchKernelGGL(k_ghost, dim3(nblk(g.nper)), dim3(TPB), 0, st, g, s.X);
}
void launch_update(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  hipLaunchKernelGGL(k_update, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, s.X, s.q, (const double *)s.qn,
                     (const double *)s.pbias);
}

// KEC: q{n+1} from the last iterate with its localPBias
// (src/solverNewton.loci:127-157), scalars (src/solverScalarTransport.loci:367-377)
__global__ void k_kec_final(DevGrid g, Prm p, int ns, const double *__restrict__ qn,
                            const double *__restrict__ pbias, double *__restrict__ q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.nc) return;
  const int nc = g.nc;
  const V3 us = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
  const V3 un = {qn[nc + c], qn[2 * nc + c], qn[3 * nc + c]};
  double np1[6];
  kec_np1(p, q[c], q[4 * nc + c], us, qn[c], qn[4 * nc + c], un, pbias[c], np1);
#pragma unroll
  for (int k = 0; k < 5; ++k) q[(size_t)k * nc + c] = np1[k];
  for (int j = 0; j < ns; ++j)
    q[(size_t)(5 + j) * nc + c] = kec_xnp1(q[(size_t)(5 + j) * nc + c], qn[(size_t)(5 + j) * nc + c]);
}
void launch_kec_final(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  const int ns = p.turb == 2 ? 2 : p.turb ? 1 : 0;
  hipLaunchKernelGGL(k_kec_final, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, ns, (const double *)s.qn,
                     (const double *)s.pbias, s.q);
}


// ---------------------------------------------------------------------------
// linear residual of the assembled system, r = B - A X, A = D + L + U, in the
// norm sum_c sum_k (r_ck / sB_k)^2 with sB_k the RMS of B_k over the cells.
// D X is formed from the packed LU factors: P D = L U.
// ---------------------------------------------------------------------------
template <typename F>
__global__ void k_b_stats(int nc, const F *__restrict__ B, double *__restrict__ part) {
  __shared__ double sh[TPB][5];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double b = c < nc ? (double)B[(size_t)c * 5 + k] : 0.0;
    sh[threadIdx.x][k] = b * b;
  }
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 5; ++k) sh[threadIdx.x][k] += sh[threadIdx.x + w][k];
    __syncthreads();
  }
  if (threadIdx.x < 5) part[(size_t)blockIdx.x * 5 + threadIdx.x] = sh[0][threadIdx.x];
}

// out[i] = sum_p part[p][i] in block order (one thread per i)
// out[i] = sum_p part[p][i] for i < len, one workgroup per component:
// thread t sums parts t, t + TPB, ... in order, then a fixed-order tree
// over the workgroup, so the result does not depend on timing
__global__ void __launch_bounds__(TPB) k_sum_parts_d(int nparts, int len, const double *__restrict__ part,
                                                     double *__restrict__ out) {
  __shared__ double sh[TPB];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int p = threadIdx.x; p < nparts; p += TPB) s += part[(size_t)p * len + i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = sh[0];
}

// D x of one cell: from the packed pivoted LU factors (P D = L U) under
// block SGS, or from the unfactored block (RAWD: line SGS and FSGS keep
// fluid_D itself in this array)
template <typename F, bool RAWD>
__device__ __forceinline__ void diag_times(const F *__restrict__ lu, int pk, int c, const double x[5], double out[5]) {
  F a[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = lu[(size_t)c * 25 + i];
  if constexpr (RAWD) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 5; ++j) s += a[i + 5 * j] * x[j];
      out[i] = s;
    }
  } else {
    double y[5], z[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = i; j < 5; ++j) s += a[i + 5 * j] * x[j];
      y[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double s = y[i];
#pragma unroll
      for (int j = 0; j < i; ++j) s += a[i + 5 * j] * y[j];
      z[i] = s;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      const int p = (pk >> (3 * i)) & 7;
#pragma unroll
      for (int k = 0; k < 5; ++k)
        if (k == p) out[k] += z[i];
    }
  }
}

// (A V)_c = D_c V_c + sum over the row of M_j V_nbr(j)
template <typename F, bool RAWD>
__device__ __forceinline__ void apply_a(const DevGrid &g, const F *__restrict__ off, const F *__restrict__ lu,
                                        int pk, int c, const double *__restrict__ V, double out[5]) {
  double x[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) x[k] = V[(size_t)c * 5 + k];
  diag_times<F, RAWD>(lu, pk, c, x, out);
  const int m0 = g.m_off[c], m1 = g.m_off[c + 1];
  for (int j = m0; j < m1; ++j) {
    const F *M = off + (size_t)j * 25;
    const double *xn = V + (size_t)g.m_nbr[j] * 5;
#pragma unroll
    for (int jj = 0; jj < 5; ++jj) {
      const double xv = xn[jj];
#pragma unroll
      for (int i = 0; i < 5; ++i) out[i] += M[i + 5 * jj] * xv;
    }
  }
}

// block sums of the weighted norm of r = B - A X (per-component weights
// 1 / rms(B_k)^2); r itself into rout[nc][5] when rout is not null
template <typename F, bool RAWD>
__global__ void k_linres(DevGrid g, const double *__restrict__ bsum, const F *__restrict__ off,
                         const F *__restrict__ lu, const int *__restrict__ piv, const F *__restrict__ B,
                         const double *__restrict__ X, double *__restrict__ part, double *__restrict__ rout) {
  __shared__ double sh[TPB];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (c < g.nc) {
    double ax[5];
    apply_a<F, RAWD>(g, off, lu, RAWD ? 0 : piv[c], c, X, ax);
    const double ncd = (double)g.nc;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double r = (double)B[(size_t)c * 5 + k] - ax[k];
      if (rout) rout[(size_t)c * 5 + k] = r;
      const double sB = sqrt(bsum[k] / ncd) + 1e-300;
      const double e = r / sB;
      acc += e * e;
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Weighted inner products of r = B - A X0 and q = A D (same operator and
// per-component weights 1/rms(B_k)^2 as k_linres): part[3 * block + {0,1,2}]
// = block sums of r.r, r.q, q.q.  Both vectors need current ghost copies.
template <typename F, bool RAWD>
__global__ void k_linres_pair(DevGrid g, const double *__restrict__ bsum, const F *__restrict__ off,
                              const F *__restrict__ lu, const int *__restrict__ piv, const F *__restrict__ B,
                              const double *__restrict__ X0, const double *__restrict__ D,
                              double *__restrict__ part) {
  __shared__ double sh[TPB][3];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0.0, rq = 0.0, qq = 0.0;
  if (c < g.nc) {
    const int pk = RAWD ? 0 : piv[c];
    double ax[5], ad[5];
    apply_a<F, RAWD>(g, off, lu, pk, c, X0, ax);
    apply_a<F, RAWD>(g, off, lu, pk, c, D, ad);
    const double ncd = (double)g.nc;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double w = 1.0 / (bsum[k] / ncd + 1e-300);
      const double r = (double)B[(size_t)c * 5 + k] - ax[k];
      rr += w * r * r;
      rq += w * r * ad[k];
      qq += w * ad[k] * ad[k];
    }
  }
  sh[threadIdx.x][0] = rr; sh[threadIdx.x][1] = rq; sh[threadIdx.x][2] = qq;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 3; ++k) sh[threadIdx.x][k] += sh[threadIdx.x + w][k];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[(size_t)blockIdx.x * 3 + threadIdx.x] = sh[0][threadIdx.x];
}

// d = a - b
__global__ void k_sub_d(long n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ d) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = a[i] - b[i];
}
// x = a + alpha d
__global__ void k_axpy_d(long n, const double *__restrict__ a, double alpha, const double *__restrict__ d,
                         double *__restrict__ x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = a[i] + alpha * d[i];
}

__global__ void k_copy_d(long n, const double *__restrict__ a, double *__restrict__ b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// linear residual norm squared into out[0]; needs the ghost copies current
template <typename F, bool RAWD>
static void linres_t(const DevGrid &g, DevState &s, double *parts, double *bsum, double *rout, hipStream_t st) {
  const int nb = nblk(g.nc);
  hipLaunchKernelGGL(k_b_stats<F>, dim3(nb), dim3(TPB), 0, st, g.nc, (const F *)s.B, parts);
  hipLaunchKernelGGL(k_sum_parts_d, dim3(5), dim3(TPB), 0, st, nb, 5, parts, bsum);
  hipLaunchKernelGGL((k_linres<F, RAWD>), dim3(nb), dim3(TPB), 0, st, g, (const double *)bsum, (const F *)s.off,
                     (const F *)s.lu, (const int *)s.piv, (const F *)s.B, (const double *)s.X, parts, rout);
}
void launch_linres(const DevGrid &g, int fp32, int rawd, DevState &s, double *parts, double *bsum, double *out,
                   hipStream_t st, double *rout) {
  launch_ghost(g, s, st);
  if (fp32) {
    if (rawd) linres_t<float, true>(g, s, parts, bsum, rout, st);
    else linres_t<float, false>(g, s, parts, bsum, rout, st);
  } else {
    if (rawd) linres_t<double, true>(g, s, parts, bsum, rout, st);
    else linres_t<double, false>(g, s, parts, bsum, rout, st);
  }
  hipLaunchKernelGGL(k_sum_parts_d, dim3(1), dim3(TPB), 0, st, nblk(g.nc), 1, parts, out);
}

int lsgs_parts(const DevGrid &g) { return (int)(((long)g.nc * 5 + TPB - 1) / TPB); }
void launch_lsgs_init(DevState &s, double w0, hipStream_t st, int slot) {
  hipLaunchKernelGGL(k_lsgs_init, dim3(1), dim3(64), 0, st, s.lst + 8 * slot, w0);
}
template <typename F, bool BACK>
static void lsgs_level_t(const DevGrid &g, const int *cells, int lo, int hi, DevState &s, hipStream_t st) {
  const int nb = (hi - lo + CPW - 1) / CPW;
  hipLaunchKernelGGL((k_lsgs_level<F, BACK>), dim3(nb), dim3(64), 0, st, g, cells, lo, hi, (const F *)s.off,
                     (const double *)s.lu64, (const F *)s.B, s.X, s.xt);
}
void launch_lsgs_level(const DevGrid &g, int fp32, int backward, const int *cells, int lo, int hi, DevState &s,
                       hipStream_t st) {
  if (hi <= lo) return;
  if (fp32) {
    if (backward) lsgs_level_t<float, true>(g, cells, lo, hi, s, st);
    else lsgs_level_t<float, false>(g, cells, lo, hi, s, st);
  } else {
    if (backward) lsgs_level_t<double, true>(g, cells, lo, hi, s, st);
    else lsgs_level_t<double, false>(g, cells, lo, hi, s, st);
  }
}
void launch_fsgs_factor(const DevGrid &g, int fp32, DevState &s, hipStream_t st) {
  if (fp32)
    hipLaunchKernelGGL(k_fsgs_factor<float>, dim3(nblk(g.nc)), dim3(TPB), 0, st, g.nc, (const float *)s.lu,
                       (float *)s.luf);
  else
    hipLaunchKernelGGL(k_fsgs_factor<double>, dim3(nblk(g.nc)), dim3(TPB), 0, st, g.nc, (const double *)s.lu,
                       (double *)s.luf);
}
void launch_fsgs_level(const DevGrid &g, int fp32, const int *cells, int lo, int hi, DevState &s, hipStream_t st) {
  if (hi <= lo) return;
  const int nb = (hi - lo + CPW - 1) / CPW;
  if (fp32)
    hipLaunchKernelGGL(k_fsgs_level<float>, dim3(nb), dim3(64), 0, st, g, cells, lo, hi, (const float *)s.off,
                       (const float *)s.luf, (const float *)s.B, s.X);
  else
    hipLaunchKernelGGL(k_fsgs_level<double>, dim3(nb), dim3(64), 0, st, g, cells, lo, hi, (const double *)s.off,
                       (const double *)s.luf, (const double *)s.B, s.X);
}
void launch_zero_d(long n, double *x, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_zero_d, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, x);
}
void launch_lsgs_relax(const DevGrid &g, int iter, double abstol, double reltol, DevState &s, hipStream_t st) {
  const int n = g.nc * 5, np = lsgs_parts(g);
  hipLaunchKernelGGL(k_lsgs_relax, dim3(np), dim3(TPB), 0, st, n, (const double *)s.lst, s.X,
                     (const double *)s.xt, s.lparts);
  hipLaunchKernelGGL(k_lsgs_check, dim3(1), dim3(TPB), 0, st, np, (const double *)s.lparts, s.lst, iter, abstol,
                     reltol);
}
// slot > 0: a turbulence scalar's own relaxation state and partial sums
// (lst + 8 slot, lparts + slot x lsgs_parts), so its loop can run beside the fluid's

template <typename F, bool RAWD>
static void linres_pair_t(const DevGrid &g, DevState &s, const double *X0, const double *D, double *parts,
                          double *bsum, hipStream_t st) {
  const int nb = nblk(g.nc);
  hipLaunchKernelGGL(k_b_stats<F>, dim3(nb), dim3(TPB), 0, st, g.nc, (const F *)s.B, parts);
  hipLaunchKernelGGL(k_sum_parts_d, dim3(5), dim3(TPB), 0, st, nb, 5, parts, bsum);
  hipLaunchKernelGGL((k_linres_pair<F, RAWD>), dim3(nb), dim3(TPB), 0, st, g, (const double *)bsum,
                     (const F *)s.off, (const F *)s.lu, (const int *)s.piv, (const F *)s.B, X0, D, parts);
}
void launch_linres_pair(const DevGrid &g, int fp32, int rawd, DevState &s, const double *X0, const double *D,
                        double *parts, double *bsum, double *out, hipStream_t st) {
  if (fp32) {
    if (rawd) linres_pair_t<float, true>(g, s, X0, D, parts, bsum, st);
    else linres_pair_t<float, false>(g, s, X0, D, parts, bsum, st);
  } else {
    if (rawd) linres_pair_t<double, true>(g, s, X0, D, parts, bsum, st);
    else linres_pair_t<double, false>(g, s, X0, D, parts, bsum, st);
  }
  hipLaunchKernelGGL(k_sum_parts_d, dim3(3), dim3(TPB), 0, st, nblk(g.nc), 3, parts, out);
}
void launch_sub_d(long n, const double *a, const double *b, double *d, hipStream_t st) {
  hipLaunchKernelGGL(k_sub_d, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, a, b, d);
}
void launch_axpy_d(long n, const double *a, double alpha, const double *d, double *x, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_d, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, a, alpha, d, x);
}
void launch_ghost_of(const DevGrid &g, double *V, hipStream_t st) {
  if (g.nper) hipLaunchKernelGGL(k_ghost, dim3(nblk(g.nper)), dim3(TPB), 0, st, g, V);
}
static inline int nscal(const Prm &p) { return p.turb == 2 ? 2 : p.turb ? 1 : 0; }
void launch_ghost1(const DevGrid &g, int ns, double *V, hipStream_t st) {
  if (g.nper && ns) hipLaunchKernelGGL(k_ghost1, dim3(nblk(g.nper)), dim3(TPB), 0, st, g, ns, V);
}

void launch_tmu(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  if (p.turb)
    hipLaunchKernelGGL(k_tmu, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, (const double *)s.q, (const double *)s.grad,
                       s.tmu);
}
void launch_sa_assemble(const DevGrid &g, const Prm &p, long nnz, DevState &s, hipStream_t st) {
  if (p.turb == 2) {
    hipLaunchKernelGGL(k_sst_face_a, dim3(nblk(g.nf)), dim3(TPB), 0, st, g, (const double *)s.q,
                       (const double *)s.qf, (const double *)s.grad, (const double *)s.phi, s.saf);
    hipLaunchKernelGGL(k_sst_cell_a, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, (const double *)s.q,
                       (const double *)s.grad, (const double *)s.saf, s.scell);
    hipLaunchKernelGGL(k_sst_face_b, dim3(nblk(g.nf)), dim3(TPB), 0, st, g, p, (const double *)s.q,
                       (const double *)s.qf, (const double *)s.grad, (const double *)s.phi, (const double *)s.iflux,
                       (const double *)s.tmu, (const double *)s.scell, s.saf, s.soff, nnz);
    hipLaunchKernelGGL(k_sst_cell_b, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, (const double *)s.q,
                       (const double *)s.dtau, (const double *)s.iflux, (const double *)s.saf,
                       (const double *)s.scell, s.ssrc, s.sdinv, s.sB, s.sX, (const double *)s.qn,
                       (const double *)s.qm1);
    return;
  }
  hipLaunchKernelGGL(k_sa_face, dim3(nblk(g.nf)), dim3(TPB), 0, st, g, p, (const double *)s.q,
                     (const double *)s.qf, (const double *)s.grad, (const double *)s.phi, (const double *)s.iflux,
                     s.saf, s.soff);
  hipLaunchKernelGGL(k_sa_cell, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, (const double *)s.q,
                     (const double *)s.grad, (const double *)s.dtau, (const double *)s.iflux,
                     (const double *)s.saf, s.ssrc, s.sdinv, s.sB, s.sX, (const double *)s.qn,
                     (const double *)s.qm1);
}
void launch_sa_level(const DevGrid &g, int sc0, int ns, long nnz, const int *cells, int lo, int hi, DevState &s,
                     hipStream_t st) {
  if (hi <= lo || ns <= 0) return;
  const long lanes = (long)(hi - lo) * ns;
  hipLaunchKernelGGL(k_sa_sgs_level, dim3((unsigned)((lanes + TPB - 1) / TPB)), dim3(TPB), 0, st, g, cells, lo, hi,
                     sc0, ns, nnz, (const double *)s.soff, (const double *)s.sdinv, (const double *)s.sB, s.sX);
}
void launch_sa_lsgs_level(const DevGrid &g, int sc, long nnz, int backward, const int *cells, int lo, int hi,
                          DevState &s, hipStream_t st) {
  if (hi <= lo) return;
  const dim3 grid((hi - lo + TPB - 1) / TPB);
  const size_t nc = g.nc, xs = nc + g.nper + 1;
  const double *off = s.soff + sc * nnz, *dinv = s.sdinv + sc * nc, *B = s.sB + sc * nc;
  double *X = s.sX + sc * xs, *xt = s.sxt + sc * nc;
  if (backward)
    hipLaunchKernelGGL(k_sa_lsgs_level<true>, grid, dim3(TPB), 0, st, g, cells, lo, hi, off, dinv, B, X, xt);
  else
    hipLaunchKernelGGL(k_sa_lsgs_level<false>, grid, dim3(TPB), 0, st, g, cells, lo, hi, off, dinv, B, X, xt);
}
void launch_sa_lsgs_relax(const DevGrid &g, int sc, int iter, double abstol, double reltol, DevState &s,
                          hipStream_t st, int slot) {
  const int np = nblk(g.nc);
  const size_t nc = g.nc, xs = nc + g.nper + 1;
  double *lst = s.lst + 8 * slot, *parts = s.lparts + (size_t)slot * lsgs_parts(g);
  hipLaunchKernelGGL(k_lsgs_relax, dim3(np), dim3(TPB), 0, st, g.nc, (const double *)lst, s.sX + sc * xs,
                     (const double *)(s.sxt + sc * nc), parts);
  hipLaunchKernelGGL(k_lsgs_check, dim3(1), dim3(TPB), 0, st, np, (const double *)parts, lst, iter, abstol, reltol);
}
void launch_sa_update(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st) {
  hipLaunchKernelGGL(k_sa_update, dim3(nblk(g.nc)), dim3(TPB), 0, st, g, p, nscal(p), (const double *)s.sX, s.q);
}

void launch_copy_d(long n, const double *a, double *b, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_d, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, st, n, a, b);
}

}  // namespace fpsi
