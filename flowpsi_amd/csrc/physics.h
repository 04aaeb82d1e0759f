// Device physics of the flowPsi Newton step, one face (or cell) per lane.
//
// Every function restates a flowPsi rule body (cited per function); the
// arithmetic keeps the rule's operation order and is compiled without FMA
// contraction so that results agree with flowPsi's FP64 evaluation to
// rounding.  Blocks are 5x5 column-major (element (i,j) at i + 5*j).
#pragma once
#include <hip/hip_runtime.h>

namespace fpsi {

struct Prm {
  double gamma, Rt, Pamb;
  int flux, limiter;
  double Kl;
  int viscous;
  double a1, a2, a3, k1, k2, k3;
  double Twall, pg0, T0, u0x, u0y, u0z;
  double cflmax, dtmax, urelax, minPg;
  // low-Mach preconditioning: pre = 0 when Mlim2 >= 1 (eta = 1 everywhere)
  int pre;
  double Mlim2, PLimPC;
  // turbulence: 0 none, 1 Spalart-Allmaras (constants, nuTilde of inflow
  // faces), 2 Menter SST (k and omega of inflow faces); Cp and the
  // turbulent Prandtl number of the eddy conductivity
  int turb;
  double sa_cb1, sa_k, sa_cv1, sa_cb2, sa_cw2, sa_cw3, sa_sigma, Cp, Prt, nut_bc;
  double k_bc, w_bc;
  // time-accurate BDF2 (timeStepMode unsteady): time step dtmax, pseudo-time
  // CFL of the preconditioning blend
  int unsteady;           // 0 steady, 1 BDF2, 2 KEC
  double pseudo_cflmax;
  double pbias_ctl;       // timePressureBiasControl (KEC)
  double tdamp;           // temporalDamping beta (<= 0: not set)
  // boundary parameter sets [set][8]: T, pg, ux, uy, uz, p0, T0, massFlux
  const double *bcs;
};

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// primitive state of one side of a face
struct Q5 {
  double T;
  V3 u;
  double pg;
};

// include/inviscidFlux.h:25-45 is the consistent flux; not needed on the path.

// src/fluxInviscid.loci:157-256
__device__ inline void hllc_flux(double *f, const Q5 &L, const Q5 &R, const V3 &an, double area,
                                 double pamb, double Rt, double gamma, double Us_n, double Eta) {
  const double gm1 = gamma - 1;
  const double rl = (L.pg + pamb) / (Rt * L.T);
  const double rr = (R.pg + pamb) / (Rt * R.T);
  const double h0l = Rt * L.T * gamma / gm1 + 0.5 * dot(L.u, L.u);
  const double h0r = Rt * R.T * gamma / gm1 + 0.5 * dot(R.u, R.u);
  const double e0l = h0l - Rt * L.T;
  const double e0r = h0r - Rt * R.T;
  const double srl = sqrt(rl), srr = sqrt(rr);
  const double rf = 1. / (srl + srr);
  V3 U_ = {(L.u.x * srl + R.u.x * srr) * rf, (L.u.y * srl + R.u.y * srr) * rf,
           (L.u.z * srl + R.u.z * srr) * rf};
  const double ut_ = dot(U_, an) - Us_n;
  const double h0_ = (h0l * srl + h0r * srr) * rf;
  const double a_ = sqrt(dmax(gm1 * (h0_ - 0.5 * dot(U_, U_)), 0.0));
  const double EtaP = Eta + 1., EtaM = Eta - 1.;
  const double sigma = sqrt(dmax(ut_ * ut_ * EtaM * EtaM + 4. * Eta * a_ * a_, 1e-30));
  const double evl_ = 0.5 * (ut_ * EtaP - sigma);
  const double evr_ = 0.5 * (ut_ * EtaP + sigma);
  const double utla = dot(L.u, an), utra = dot(R.u, an);
  const double utl = utla - Us_n, utr = utra - Us_n;
  const double al2 = gamma * Rt * L.T, ar2 = gamma * Rt * R.T;
  const double sigmar = sqrt(dmax(utr * utr * EtaM * EtaM + 4. * Eta * ar2, 0.0));
  const double sigmal = sqrt(dmax(utl * utl * EtaM * EtaM + 4. * Eta * al2, 0.0));
  const double SL = dmin(evl_, .5 * (utl * EtaP - sigmal));
  const double SR = dmax(evr_, .5 * (utr * EtaP + sigmar));
  const double lmdot = rl * utl, rmdot = rr * utr;
  const double pl = L.pg + pamb, pr = R.pg + pamb;
  const double coef = 0.5 * area;
  const double AL = coef * ((SL > 0) ? (SR < 0 ? 1. : 2.) : (SR < 0 ? 0. : 1.));
  const double AR = coef * ((SR < 0) ? (SL > 0 ? 1. : 2.) : (SL > 0 ? 0. : 1.));
  const double AM = coef * ((SL > 0 || SR < 0) ? 0. : 1.);
  f[0] = AL * lmdot + AR * rmdot;
  f[1] = AL * lmdot * L.u.x + AR * rmdot * R.u.x + an.x * (AL * L.pg + AR * R.pg);
  f[2] = AL * lmdot * L.u.y + AR * rmdot * R.u.y + an.y * (AL * L.pg + AR * R.pg);
  f[3] = AL * lmdot * L.u.z + AR * rmdot * R.u.z + an.z * (AL * L.pg + AR * R.pg);
  f[4] = AL * (lmdot * e0l + pl * utla) + AR * (rmdot * e0r + pr * utra);
  f[0] -= AM * (SL * rl + SR * rr);
  f[1] -= AM * (SL * rl * L.u.x + SR * rr * R.u.x);
  f[2] -= AM * (SL * rl * L.u.y + SR * rr * R.u.y);
  f[3] -= AM * (SL * rl * L.u.z + SR * rr * R.u.z);
  f[4] -= AM * (SL * rl * e0l + SR * rr * e0r);
  const double SM = ((rr * utr * (SR - utr) - rl * utl * (SL - utl) + L.pg - R.pg) /
                     (rr * (SR - utr) - rl * (SL - utl)));
  const double rl_star = rl * (SL - utl) / (SL - SM);
  const double rr_star = rr * (SR - utr) / (SR - SM);
  const double pg_star = rl * (utl - SL) * (utl - SM) + L.pg;
  const double p_star = pg_star + pamb;
  const double rul_star = ((SL - utl) * rl * L.u.x + (pg_star - L.pg) * an.x) / (SL - SM);
  const double rvl_star = ((SL - utl) * rl * L.u.y + (pg_star - L.pg) * an.y) / (SL - SM);
  const double rwl_star = ((SL - utl) * rl * L.u.z + (pg_star - L.pg) * an.z) / (SL - SM);
  const double rur_star = ((SR - utr) * rr * R.u.x + (pg_star - R.pg) * an.x) / (SR - SM);
  const double rvr_star = ((SR - utr) * rr * R.u.y + (pg_star - R.pg) * an.y) / (SR - SM);
  const double rwr_star = ((SR - utr) * rr * R.u.z + (pg_star - R.pg) * an.z) / (SR - SM);
  const double re0l_star = ((SL - utl) * rl * e0l - pl * utl + p_star * SM) / (SL - SM);
  const double re0r_star = ((SR - utr) * rr * e0r - pr * utr + p_star * SM) / (SR - SM);
  const double CR = SR - fabs(SM);
  const double CL = SL + fabs(SM);
  f[0] += AM * (CR * rr_star + CL * rl_star);
  f[1] += AM * (CR * rur_star + CL * rul_star);
  f[2] += AM * (CR * rvr_star + CL * rvl_star);
  f[3] += AM * (CR * rwr_star + CL * rwl_star);
  f[4] += AM * (CR * re0r_star + CL * re0l_star);
}

// Coefficients shared by the Roe flux dissipation and RoeJacobian
// (src/fluxInviscid.loci:985-1000 and :1042-1066).
struct RoeT {
  double T4, T7, T8, T9, T10, T11;
};

__device__ __forceinline__ RoeT roe_terms(double ev1, double ev2, double ev3, double un, double una,
                                          double a, double beta, double sigma, double gamma,
                                          double Rt, double rho, double h0, double rT) {
  const double BM = beta - 1.;
  const double T1 = ev2 - ev3;
  const double T2 = (ev2 + ev3 - 2 * ev1);
  const double T3 = rho * (BM * BM * un * un - sigma * sigma) / (4. * sigma);
  RoeT r;
  r.T4 = T2 * sigma + BM * T1 * un;
  const double T5 = (T2 * sigma - BM * T1 * un) / (2. * sigma);
  const double T6 = ((ev2 + ev3) * sigma - BM * T1 * un) / (2. * sigma);
  const double gm1 = gamma - 1.0;
  const double Cp = gamma * Rt / gm1;
  const double rbeta = 1. / beta;
  const double ra2 = 1. / (a * a);
  r.T7 = rbeta * T1 * T3 * ra2;
  r.T8 = (((rbeta - 1) + gamma) * T6 - gm1 * T5) * ra2;
  r.T9 = Cp * rho + h0 * rT;
  const double x1 = gamma * h0 * T1 * T3 * ra2 * rbeta;
  const double x2 = -rbeta * gm1 * T1 * T3 * r.T9 * ra2 / rT;
  const double x3 = -T1 * T3 * rbeta;
  const double x4 = .5 * rho * r.T4 * una / sigma;
  r.T10 = (x1 + x2 + x3 + x4);
  r.T11 = (rbeta * ra2 * h0 + (r.T9 - rho * rbeta * Cp) * ra2 * gm1 / rT - rbeta) * T6 -
          gm1 * ra2 * r.T9 * T5 / rT - una * T1 / sigma;
  return r;
}

// src/fluxInviscid.loci:869-1014 (roeCT_flux with the Harten-Hyman entropy fix)
__device__ inline void roe_flux(double *f, const Q5 &L, const Q5 &R, const V3 &an, double area,
                                double pamb, double Rt, double gamma, double Us_n, double Eta) {
  const double pl = L.pg + pamb, pr = R.pg + pamb;
  const double rl = pl / (Rt * L.T), rr = pr / (Rt * R.T);
  const double gm1 = gamma - 1.;
  const double re0l = rl * (Rt * L.T / gm1 + 0.5 * dot(L.u, L.u));
  const double re0r = rr * (Rt * R.T / gm1 + 0.5 * dot(R.u, R.u));
  const double h0l = (re0l + pl) / rl, h0r = (re0r + pr) / rr;
  const double r_ = sqrt(rl * rr);
  const double srl = sqrt(rl), srr = sqrt(rr);
  const double rf = 1. / (srl + srr);
  V3 U_ = {(L.u.x * srl + R.u.x * srr) * rf, (L.u.y * srl + R.u.y * srr) * rf,
           (L.u.z * srl + R.u.z * srr) * rf};
  const double ut_ = dot(U_, an) - Us_n;
  const double uta_ = dot(U_, an);
  const double q2_ = dot(U_, U_);
  const double h0_ = (h0l * srl + h0r * srr) * rf;
  const double h_ = h0_ - 0.5 * q2_;
  const double al2 = gamma * Rt * L.T, ar2 = gamma * Rt * R.T;
  const double utl = dot(L.u, an) - Us_n, utr = dot(R.u, an) - Us_n;
  const double utla = dot(L.u, an), utra = dot(R.u, an);
  if (fabs(pl - pr) / (pl + pr + 1e-33) > .1) Eta = 1.;
  const double af2 = dmax(gm1 * h_, 1e-30);
  const double EtaP = Eta + 1., EtaM = Eta - 1.;
  const double a_ = sqrt(af2);
  const double sigma = sqrt(ut_ * ut_ * EtaM * EtaM + 4. * Eta * a_ * a_);
  double ev1 = fabs(ut_);
  double ev2 = fabs(ut_ * EtaP - sigma) / 2.;
  double ev3 = fabs(ut_ * EtaP + sigma) / 2.;
  const double sigmal = sqrt(utl * utl * EtaM * EtaM + 4. * Eta * al2);
  const double sigmar = sqrt(utr * utr * EtaM * EtaM + 4. * Eta * ar2);
  const double e2l = (utl * EtaP - sigmal) / 2., e2r = (utr * EtaP - sigmar) / 2.;
  const double e3l = (utl * EtaP + sigmal) / 2., e3r = (utr * EtaP + sigmar) / 2.;
  const double ev2s = (ut_ * EtaP - sigma) / 2., ev3s = (ut_ * EtaP + sigma) / 2.;
  const double eps2 = dmax(0.0, dmax(ev2s - e2l, e2r - ev2s));
  const double eps3 = dmax(0.0, dmax(ev3s - e3l, e3r - ev3s));
  ev2 = ev2 < eps2 ? 0.5 * ((ev2s * ev2s) / eps2 + eps2) : ev2;
  ev3 = ev3 < eps3 ? 0.5 * ((ev3s * ev3s) / eps3 + eps3) : ev3;
  const V3 Uj = {R.u.x - L.u.x, R.u.y - L.u.y, R.u.z - L.u.z};
  const double utj = dot(Uj, an);
  const double pj = R.pg - L.pg;
  const double Tj = R.T - L.T;
  const double rT_ = -gamma * Rt * r_ / (h_ * gm1);
  const RoeT t = roe_terms(ev1, ev2, ev3, ut_, uta_, a_, Eta, sigma, gamma, Rt, r_, h0_, rT_);
  const double T1 = ev2 - ev3;
  double d0 = pj * t.T8 + utj * t.T7 + ev1 * Tj * rT_;
  double d1 = U_.x * t.T7 * utj + ev1 * (U_.x * Tj * rT_ + r_ * Uj.x) +
              pj * (t.T8 * U_.x - an.x * T1 / sigma) + an.x * r_ * t.T4 * utj / (2. * sigma);
  double d2 = U_.y * t.T7 * utj + ev1 * (U_.y * Tj * rT_ + r_ * Uj.y) +
              pj * (t.T8 * U_.y - an.y * T1 / sigma) + an.y * r_ * t.T4 * utj / (2. * sigma);
  double d3 = U_.z * t.T7 * utj + ev1 * (U_.z * Tj * rT_ + r_ * Uj.z) +
              pj * (t.T8 * U_.z - an.z * T1 / sigma) + an.z * r_ * t.T4 * utj / (2. * sigma);
  double d4 = pj * t.T11 + t.T10 * utj + ev1 * (t.T9 * Tj + r_ * dot(U_, Uj));
  const double Ah = 0.5 * area;
  f[0] = Ah * (rl * utl + rr * utr - d0);
  f[1] = Ah * (L.u.x * rl * utl + R.u.x * rr * utr + an.x * (L.pg + R.pg) - d1);
  f[2] = Ah * (L.u.y * rl * utl + R.u.y * rr * utr + an.y * (L.pg + R.pg) - d2);
  f[3] = Ah * (L.u.z * rl * utl + R.u.z * rr * utr + an.z * (L.pg + R.pg) - d3);
  f[4] = Ah * ((re0l * utl + pl * utla) + (re0r * utr + pr * utra) - d4);
}

// Accumulate v into a real_fj element the way `real_fj += real` rounds.
template <typename F>
__device__ __forceinline__ void acc(F &a, double v) { a = (F)((double)a + v); }

// src/fluxInviscid.loci:1018-1090 (RoeJacobian), accumulated into Fj.
template <typename F>
__device__ inline void roe_jacobian_ev(F *Fj, double ev1, double ev2, double ev3, const V3 &U,
                                       double un, double a, double beta, double sigma, double gamma,
                                       double Rt, double rho, double h0, const V3 &n, double area) {
  const double una = dot(U, n);
  const double gm1 = gamma - 1.0;
  const double rT = -gamma * Rt * rho / ((h0 - 0.5 * dot(U, U)) * gm1);
  const RoeT t = roe_terms(ev1, ev2, ev3, un, una, a, beta, sigma, gamma, Rt, rho, h0, rT);
  const double T1 = ev2 - ev3;
#define FJ(i, j) Fj[(i) + 5 * (j)]
  acc(FJ(0, 0), area * ev1 * rT);
  acc(FJ(0, 1), area * n.x * t.T7);
  acc(FJ(0, 2), area * n.y * t.T7);
  acc(FJ(0, 3), area * n.z * t.T7);
  acc(FJ(0, 4), area * t.T8);
  acc(FJ(1, 0), area * ev1 * rT * U.x);
  acc(FJ(2, 0), area * ev1 * rT * U.y);
  acc(FJ(3, 0), area * ev1 * rT * U.z);
  acc(FJ(1, 1), area * (n.x * (t.T7 * U.x + n.x * rho * t.T4 / (2. * sigma)) + ev1 * rho));
  acc(FJ(1, 2), area * n.y * (t.T7 * U.x + n.x * rho * t.T4 / (2. * sigma)));
  acc(FJ(1, 3), area * n.z * (t.T7 * U.x + n.x * rho * t.T4 / (2. * sigma)));
  acc(FJ(2, 1), area * n.x * (t.T7 * U.y + n.y * rho * t.T4 / (2. * sigma)));
  acc(FJ(2, 2), area * (n.y * (t.T7 * U.y + n.y * rho * t.T4 / (2. * sigma)) + ev1 * rho));
  acc(FJ(2, 3), area * n.z * (t.T7 * U.y + n.y * rho * t.T4 / (2. * sigma)));
  acc(FJ(3, 1), area * n.x * (t.T7 * U.z + n.z * rho * t.T4 / (2. * sigma)));
  acc(FJ(3, 2), area * n.y * (t.T7 * U.z + n.z * rho * t.T4 / (2. * sigma)));
  acc(FJ(3, 3), area * (n.z * (t.T7 * U.z + n.z * rho * t.T4 / (2. * sigma)) + ev1 * rho));
  acc(FJ(1, 4), area * (t.T8 * U.x - n.x * T1 / sigma));
  acc(FJ(2, 4), area * (t.T8 * U.y - n.y * T1 / sigma));
  acc(FJ(3, 4), area * (t.T8 * U.z - n.z * T1 / sigma));
  acc(FJ(4, 0), area * ev1 * t.T9);
  acc(FJ(4, 1), area * (n.x * t.T10 + ev1 * rho * U.x));
  acc(FJ(4, 2), area * (n.y * t.T10 + ev1 * rho * U.y));
  acc(FJ(4, 3), area * (n.z * t.T10 + ev1 * rho * U.z));
  acc(FJ(4, 4), area * t.T11);
#undef FJ
}

// Roe-averaged state and eigenvalues of roeCT_jacobian (src/fluxInviscid.loci:1092-1160)
struct RoeAvg {
  V3 U;
  double ut, a, sigma, rho, h0, eta;
  double ev1, ev2, ev3, eps2;
};

__device__ inline RoeAvg roe_average(const Q5 &L, const Q5 &R, const V3 &an, double pamb, double Rt,
                                     double gamma, double Us_n, double Eta) {
  RoeAvg r;
  const double gm1 = gamma - 1;
  const double pl = L.pg + pamb, pr = R.pg + pamb;
  const double rl = pl / (Rt * L.T), rr = pr / (Rt * R.T);
  const double re0l = rl * (Rt * L.T / gm1 + 0.5 * dot(L.u, L.u));
  const double re0r = rr * (Rt * R.T / gm1 + 0.5 * dot(R.u, R.u));
  const double h0l = (re0l + pl) / rl, h0r = (re0r + pr) / rr;
  r.rho = sqrt(rl * rr);
  const double srl = sqrt(rl), srr = sqrt(rr);
  const double rf = 1. / (srl + srr);
  r.U = {(L.u.x * srl + R.u.x * srr) * rf, (L.u.y * srl + R.u.y * srr) * rf,
         (L.u.z * srl + R.u.z * srr) * rf};
  r.ut = dot(r.U, an) - Us_n;
  const double q2_ = dot(r.U, r.U);
  r.h0 = (h0l * srl + h0r * srr) * rf;
  const double h_ = r.h0 - 0.5 * q2_;
  if (fabs(pl - pr) / (pl + pr + 1e-33) > .1) Eta = 1.;
  r.eta = Eta;
  const double af2 = dmax(gm1 * h_, 1e-30);
  const double EtaP = Eta + 1., EtaM = Eta - 1.;
  r.a = sqrt(af2);
  r.sigma = sqrt(r.ut * r.ut * EtaM * EtaM + 4. * Eta * r.a * r.a);
  r.ev1 = r.ut;
  r.ev2 = (r.ut * EtaP - r.sigma) / 2.;
  r.ev3 = (r.ut * EtaP + r.sigma) / 2.;
  const double mxv = dmax(r.ev1, dmax(r.ev2, r.ev3));
  const double mnv = dmin(r.ev1, dmin(r.ev2, r.ev3));
  const double eps = 0.02 * (mxv - mnv);
  r.eps2 = eps * eps;
  return r;
}

// roeCT_jacobian(dir) from a Roe average (src/fluxInviscid.loci:1161-1167)
template <typename F>
__device__ inline void roe_jacobian(F *Fj, const RoeAvg &r, const V3 &an, double area, double gamma,
                                    double Rt, int dir) {
  const double d = (dir < 0) ? -1 : 1;
  const double e1 = .5 * (r.ev1 + d * sqrt(r.ev1 * r.ev1 + r.eps2));
  const double e2 = .5 * (r.ev2 + d * sqrt(r.ev2 * r.ev2 + r.eps2));
  const double e3 = .5 * (r.ev3 + d * sqrt(r.ev3 * r.ev3 + r.eps2));
  roe_jacobian_ev<F>(Fj, e1, e2, e3, r.U, r.ut, r.a, r.eta, r.sigma, gamma, Rt, r.rho, r.h0, an, area);
}

// include/bcTools.h:25-51
__device__ inline void cv_inflow(Q5 &b, const Q5 &i, const Q5 &r0s, const V3 &n, double us_n,
                                 double Pambient, double gamma, double Rtilde, double Beta) {
  const double uit = dot(i.u, n) - us_n;
  const double u0t = dot(r0s.u, n) - us_n;
  const double a02 = gamma * Rtilde * r0s.T;
  const double a0 = sqrt(a02);
  const double r0 = (r0s.pg + Pambient) / (Rtilde * r0s.T);
  const double MinPressure = 0.1 * (r0s.pg + Pambient) - Pambient;
  const double BetaM = 0.5 * (1.0 - Beta);
  const double uitBeta = uit * (1. - Beta);
  const double sigma2 = uitBeta * uitBeta + 4. * Beta * a0 * a0;
  const double sigma = sqrt(sigma2);
  double p = 0.5 * (r0s.pg + i.pg - 0.5 * r0 * sigma * (u0t - uit)) -
             ((i.pg - r0s.pg) + r0 * (uit - u0t) * uit * BetaM) * uit * BetaM / (2. * sigma);
  p = dmax(p, MinPressure);
  b.u.x = r0s.u.x + 2. * n.x * (p - r0s.pg) / (r0 * (sigma + uit * BetaM));
  b.u.y = r0s.u.y + 2. * n.y * (p - r0s.pg) / (r0 * (sigma + uit * BetaM));
  b.u.z = r0s.u.z + 2. * n.z * (p - r0s.pg) / (r0 * (sigma + uit * BetaM));
  const double rhob = r0 + (p - r0s.pg) * (sigma - uit * BetaM) / (Beta * a0 * a0 * (sigma + uit * BetaM));
  b.T = (p + Pambient) / (Rtilde * rhob);
  b.pg = p;
}

// include/bcTools.h:52-80
__device__ inline void cv_outflow(Q5 &b, const Q5 &i, const Q5 &r0s, const V3 &n, double us_n,
                                  double Pambient, double gamma, double Rtilde, double Beta) {
  const double uit = dot(i.u, n) - us_n;
  const double u0t = dot(r0s.u, n) - us_n;
  const double a02 = gamma * Rtilde * r0s.T;
  const double a0 = sqrt(a02);
  const double r0 = (r0s.pg + Pambient) / (Rtilde * r0s.T);
  const double MinPressure = 0.1 * (i.pg + Pambient) - Pambient;
  const double BetaM = 0.5 * (1.0 - Beta);
  const double uitBeta = uit * (1. - Beta);
  const double sigma2 = uitBeta * uitBeta + 4. * Beta * a0 * a0;
  const double sigma = sqrt(sigma2);
  double p = .5 * (i.pg + r0s.pg - 0.5 * r0 * sigma * (u0t - uit)) +
             ((i.pg - r0s.pg) + r0 * (uit - u0t) * uit * BetaM) * uit * BetaM / (2. * sigma);
  p = dmax(p, MinPressure);
  b.u.x = i.u.x + 2. * n.x * (i.pg - p) / (r0 * (sigma + uit * BetaM));
  b.u.y = i.u.y + 2. * n.y * (i.pg - p) / (r0 * (sigma + uit * BetaM));
  b.u.z = i.u.z + 2. * n.z * (i.pg - p) / (r0 * (sigma + uit * BetaM));
  const double rhob = r0 + (p - i.pg) * (sigma - uit * BetaM) / (Beta * a0 * a0 * (sigma + uit * BetaM));
  b.T = (p + Pambient) / (Rtilde * rhob);
  b.pg = p;
}

// ---------------------------------------------------------------------------
// boundary conditions with a parameter set of their own (include/fpsi.h
// FPSI_BC_REFLECTING .. FIXED_MASS)
// ---------------------------------------------------------------------------
// inviscidFlux (include/inviscidFlux.h:25-45): the Euler flux of one state
__device__ inline void euler_flux(double *f, const Q5 &s, const V3 &an, double area, double Pamb, double Rt,
                                  double gamma, double us_n) {
  const double gm1 = gamma - 1;
  const double rho = (s.pg + Pamb) / (Rt * s.T);
  const double h0 = Rt * s.T * gamma / gm1 + 0.5 * dot(s.u, s.u);
  const double e0 = h0 - Rt * s.T;
  const double uta = dot(s.u, an);
  const double ut = uta - us_n;
  const double mdot = area * rho * ut;
  const double pa = s.pg + Pamb;
  f[0] = mdot;
  f[1] = mdot * s.u.x + area * s.pg * an.x;
  f[2] = mdot * s.u.y + area * s.pg * an.y;
  f[3] = mdot * s.u.z + area * s.pg * an.z;
  f[4] = mdot * e0 + area * pa * uta;
}

// reflecting face values, src/bcImpermeable.loci:102-117 (rho_i the cell's density)
__device__ inline void reflecting_state(Q5 &o, const Q5 &i, double rho_i, const V3 &n, double us_n, double Pamb) {
  const double uit = dot(i.u, n) - us_n;
  o.u = {i.u.x - uit * n.x, i.u.y - uit * n.y, i.u.z - uit * n.z};
  const double Pmin = 0.5 * (i.pg + Pamb) - Pamb;
  o.pg = dmax(Pmin, i.pg + 0.5 * rho_i * uit * fabs(uit));
  o.T = i.T;
}

// the mirrored right state of the reflecting flux, src/bcImpermeable.loci:119-130
__device__ inline Q5 mirror(const Q5 &l, const V3 &n, double us_n) {
  const double uit = dot(l.u, n) - us_n;
  Q5 r;
  r.T = l.T;
  r.pg = l.pg;
  r.u = {l.u.x - 2. * uit * n.x, l.u.y - 2. * uit * n.y, l.u.z - 2. * uit * n.z};
  return r;
}

// fixed_p0t0_bc and total_pressure_bc_imp, src/bcIsentropic.loci:66-144:
// the boundary state from p0, T0 and the outgoing Riemann invariant along
// the flow direction d (the inward normal), with no rigid rotation
__device__ inline void p0t0_state(Q5 &o, double T0, double p0, double gamma, double Rt, const Q5 &i, const V3 &d,
                                  double Pamb) {
  const double gm1 = gamma - 1.;
  const double a0 = sqrt(gamma * Rt * T0);
  const double ai = sqrt(gamma * Rt * i.T);
  const double uit = dot(i.u, d);
  double ab, unb;
  {
    const double rim = uit - 2. * ai / gm1;
    const double a = (2. / gm1 + 1);
    const double b = (2. * rim);
    const double c = (0.5 * gm1 * rim * rim - a0 * a0);
    const double D = b * b - 4. * a * c;
    if (D < 0) {
      ab = a0;
    } else {
      ab = (-b + sqrt(D)) / (2. * a);
      ab = dmax(ab, .25 * a0);
    }
    unb = rim + 2. * ab / gm1;
  }
  const double maxMach = 1.0;
  V3 ub = {unb * d.x + 0.0, unb * d.y + 0.0, unb * d.z + 0.0};
  double M = sqrt(dot(ub, ub)) / ab;
  if (M > maxMach) {
    M = maxMach;
    const double utot = M * a0 / sqrt(1 + .5 * gm1 * M * M);
    const double un = sqrt(dmax(utot * utot - 0.0, 0.0));
    ub = {un * d.x + 0.0, un * d.y + 0.0, un * d.z + 0.0};
  }
  double Tb = T0 / (1 + .5 * gm1 * M * M);
  double pb = p0 * pow(1. + .5 * gm1 * M * M, -gamma / gm1);
  double f = (Tb - i.T) / dmax(i.T, Tb) - .1;
  if (f > 0) {
    f = pow(f, .5);
    ab = (1 - f) * ab + f * a0;
    unb = (1 - f) * unb + f * uit;
    ub = {unb * d.x + 0.0, unb * d.y + 0.0, unb * d.z + 0.0};
    double M2 = sqrt(dot(ub, ub)) / ab;
    if (M2 > maxMach) {
      M2 = maxMach;
      const double utot = M2 * a0 / sqrt(1 + .5 * gm1 * M2 * M2);
      const double un = sqrt(dmax(utot * utot - 0.0, 0.0));
      ub = {un * d.x + 0.0, un * d.y + 0.0, un * d.z + 0.0};
    }
    Tb = T0 / (1 + .5 * gm1 * M2 * M2);
    pb = p0 * pow(1. + .5 * gm1 * M2 * M2, -gamma / gm1);
  }
  o.u = ub;
  o.T = Tb;
  o.pg = pb - Pamb;
}

// mass_flowrate_bc, src/bcFixedMass.loci:85-136: the temperature that
// carries mass flux mflux at the cell's pressure pin with total
// temperature Tstag (fixed-point iteration from the cell's temperature),
// flowing along d (the inward normal)
__device__ inline void fixed_mass_state(Q5 &o, double pin, double Tin, double mflux, double Tstag, double gamma,
                                        double Rt, const V3 &d, const V3 &n, double Pamb) {
  double T = Tin;
  const double eps = Tin * 1e-9;
  for (int iter = 0; iter < 1000; ++iter) {
    const double rho = pin / (Rt * T);
    const double a = sqrt(gamma * Rt * T);
    const double u = mflux / rho;
    const double M = u / a;
    const double lamda = sqrt(((gamma + 1.0) * M * M) / ((gamma - 1.0) * M * M + 2));
    const double p0 = pin * pow((1. - ((gamma - 1.) / (gamma + 1.)) * lamda * lamda), -gamma / (gamma - 1.0));
    const double rho0 = p0 / (Rt * Tstag);
    const double rho_1 = rho0 * pow((1. - (gamma - 1.0) / (gamma + 1.0) * lamda * lamda), 1.0 / (gamma - 1.0));
    const double T_1 = pin / (rho_1 * Rt);
    const double dT = T_1 - T;
    if (fabs(dT) < eps) break;
    T = T_1;
  }
  const double rhob = pin / (Rt * T);
  const double vel = mflux / rhob;
  const double sc = -vel / dot(d, n);
  o.u = {d.x * sc + 0.0 + 0.0, d.y * sc + 0.0 + 0.0, d.z * sc + 0.0 + 0.0};
  o.T = T;
  o.pg = pin - Pamb;
}

// inviscidRiemannFlux: the flux option's Riemann flux
__device__ inline void riemann_flux(double *f, const Q5 &L, const Q5 &R, const V3 &an, double area, const Prm &p,
                                    double us_n, double eta) {
  if (p.flux == 1) roe_flux(f, L, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, eta);
  else hllc_flux(f, L, R, an, area, p.Pamb, p.Rt, p.gamma, us_n, eta);
}

// srcJ of a reflecting face by one-sided differences of its flux in the
// reconstructed left state, momentum and energy rows
// (src/bcImpermeable.loci:132-181, the NUMERICAL form the file compiles)
template <typename F>
__device__ inline void reflecting_jacobian(F *blk, const Q5 &L, const V3 &an, double area, const Prm &p,
                                           double us_n, double eta) {
  double f0[5];
  riemann_flux(f0, L, mirror(L, an, us_n), an, area, p, us_n, eta);
  const double a2 = p.gamma * p.Rt * L.T;
  const double eps = 1e-5;
  const double deltas[5] = {L.T * eps, sqrt(a2) * eps, sqrt(a2) * eps, sqrt(a2) * eps, a2 * eps};
  double vals[5] = {L.T, L.u.x, L.u.y, L.u.z, L.pg};
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const double tmp = vals[j];
    vals[j] += deltas[j];
    Q5 l;
    l.T = vals[0];
    l.u = {vals[1], vals[2], vals[3]};
    l.pg = vals[4];
    double f1[5];
    riemann_flux(f1, l, mirror(l, an, us_n), an, area, p, us_n, eta);
    const double rh = 1. / (tmp - vals[j]);
#pragma unroll
    for (int i = 1; i < 5; ++i) acc(blk[i + 5 * j], -((f0[i] - f1[i]) * rh));
    vals[j] = tmp;
  }
}

// srcJ of an isentropic-inflow (kind 11) or fixed-mass (kind 12) face by
// one-sided differences of the face state's Euler flux in the cell's
// primitives (src/bcIsentropic.loci:171-247, src/bcFixedMass.loci:166-228)
template <typename F>
__device__ inline void bc_state_jacobian(F *blk, int kind, const Q5 &ci, const double *bs, const V3 &an, double area,
                                         const Prm &p, double us_n) {
  const V3 d = {-1. * an.x, -1. * an.y, -1. * an.z};
  const double eps = 1e-5;
  const double ai = sqrt(p.gamma * p.Rt * ci.T);
  const double deltas[5] = {ci.T * eps, ai * eps, ai * eps, ai * eps, ai * ai * eps};
  double vals[5] = {ci.T, ci.u.x, ci.u.y, ci.u.z, ci.pg};
  Q5 o;
  if (kind == 11) p0t0_state(o, bs[6], bs[5], p.gamma, p.Rt, ci, d, p.Pamb);
  else fixed_mass_state(o, ci.pg + p.Pamb, ci.T, bs[7], bs[6], p.gamma, p.Rt, d, an, p.Pamb);
  double f0[5];
  euler_flux(f0, o, an, area, p.Pamb, p.Rt, p.gamma, us_n);
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const double tmp = vals[j];
    vals[j] += deltas[j];
    Q5 s;
    s.T = vals[0];
    s.u = {vals[1], vals[2], vals[3]};
    s.pg = vals[4];
    if (kind == 11) p0t0_state(o, bs[6], bs[5], p.gamma, p.Rt, s, d, p.Pamb);
    else fixed_mass_state(o, vals[4] + p.Pamb, vals[0], bs[7], bs[6], p.gamma, p.Rt, d, an, p.Pamb);
    double f1[5];
    euler_flux(f1, o, an, area, p.Pamb, p.Rt, p.gamma, us_n);
    const double rh = 1. / (tmp - vals[j]);
#pragma unroll
    for (int i = 0; i < 5; ++i) acc(blk[i + 5 * j], -((f0[i] - f1[i]) * rh));
    vals[j] = tmp;
  }
}

// Sutherland law, src/fluxViscous.loci:112-119 and :137-143
__device__ __forceinline__ double suth_mu(const Prm &p, double T) { return p.a1 * pow(T, p.a2) / (T + p.a3); }
__device__ __forceinline__ double suth_k(const Prm &p, double T) { return p.k1 * pow(T, p.k2) / (T + p.k3); }

// tau and vflux from face gradients (src/fluxViscous.loci:217-238, :279-292)
__device__ inline void viscous_flux(double *vf, const double gu[3][3], const V3 &gT, double mu_total,
                                    double ktotal, const V3 &uf, const V3 &an, double sada) {
  const double dudx = gu[0][0], dvdy = gu[1][1], dwdz = gu[2][2];
  const double dvdx = gu[1][0], dwdx = gu[2][0], dwdy = gu[2][1];
  const double dudy = gu[0][1], dudz = gu[0][2], dvdz = gu[1][2];
  const double divm = (dudx + dvdy + dwdz) * (1. / 3.);
  const double txx = 2.0 * mu_total * (dudx - divm);
  const double tyy = 2.0 * mu_total * (dvdy - divm);
  const double tzz = 2.0 * mu_total * (dwdz - divm);
  const double txy = mu_total * (dudy + dvdx);
  const double txz = mu_total * (dudz + dwdx);
  const double tyz = mu_total * (dvdz + dwdy);
  vf[0] = (txx * an.x + txy * an.y + txz * an.z) * sada;
  vf[1] = (txy * an.x + tyy * an.y + tyz * an.z) * sada;
  vf[2] = (txz * an.x + tyz * an.y + tzz * an.z) * sada;
  vf[3] = (uf.x * vf[0] + uf.y * vf[1] + uf.z * vf[2]);
  vf[3] += dot(gT, an) * ktotal * sada;
}

// Thin-layer viscous Jacobian terms (src/fluxViscous.loci:315-403 for
// fjp (s=+1) / fjm (s=-1); :405-448 for the boundary srcJ with s=-1).
template <typename F>
__device__ inline void viscous_jacobian(F *Fj, double s, const V3 &an, double sada, double dist,
                                        double mu_f, double k_f, const V3 &uf, const double *vf) {
  const double coef = mu_f * sada / dist;
  const double third = 1.0 / 3.0;
  const double n[3] = {an.x, an.y, an.z};
  const double u[3] = {uf.x, uf.y, uf.z};
#define FJ(i, j) Fj[(i) + 5 * (j)]
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double c = (i == j) ? coef * (1.0 + n[i] * n[i] * third) : coef * n[i] * n[j] * third;
      acc(FJ(1 + i, 1 + j), s * c);
    }
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      const double term = (i == j) ? u[i] * coef * (1.0 + n[i] * n[i] * third)
                                   : u[i] * coef * n[i] * n[j] * third;
      acc(FJ(4, 1 + j), s * term);
    }
  acc(FJ(4, 1), s * (.5 * vf[0]));
  acc(FJ(4, 2), s * (.5 * vf[1]));
  acc(FJ(4, 3), s * (.5 * vf[2]));
  const double factor = k_f * sada / dist;
  acc(FJ(4, 0), s * factor);
#undef FJ
}

// Temporal-damping Jacobian, src/fluxInviscid.loci:794-866: damping_fjp
// (s = +1, left face state) / damping_fjm (s = -1, right face state) add
// beta (|u.n - us_n| + a) A/2 dQ/dq of the face state; the intermediates are
// held in the Jacobian precision F as the rule body declares them.
template <typename F>
__device__ inline void damping_jacobian(F *Fj, double s, const Q5 &Q, const V3 &an, double area, double pamb,
                                        double Rt, double gamma, double us_n, double beta) {
  const F coef = (F)(0.5 * area);
  const F gm1 = (F)(gamma - 1);
  const F rr = (F)((Q.pg + pamb) / (Rt * Q.T));
  const F drdp = (F)(1. / (Rt * Q.T));
  const F drdT = (F)(-(double)rr / Q.T);
  const F uta = (F)dot(Q.u, an);
  const F ut = (F)((double)uta - us_n);
  const F ar = (F)sqrt(gamma * Rt * Q.T);
  const F aut = ut < (F)0 ? -ut : ut;
  const F sum = aut + ar;
  const double coefd = beta * (double)sum * (double)coef;
  const double uu = dot(Q.u, Q.u);
  const double dT = (double)drdT, dp = (double)drdp, r = (double)rr;
#define FJ(i, j) Fj[(i) + 5 * (j)]
  acc(FJ(0, 0), s * (coefd * dT));
  acc(FJ(1, 0), s * (coefd * dT * Q.u.x));
  acc(FJ(2, 0), s * (coefd * dT * Q.u.y));
  acc(FJ(3, 0), s * (coefd * dT * Q.u.z));
  acc(FJ(4, 0), s * (coefd * dT * 0.5 * uu));
  acc(FJ(0, 4), s * (coefd * dp));
  acc(FJ(1, 4), s * (coefd * dp * Q.u.x));
  acc(FJ(2, 4), s * (coefd * dp * Q.u.y));
  acc(FJ(3, 4), s * (coefd * dp * Q.u.z));
  acc(FJ(4, 4), s * (coefd * (1. / (double)gm1 + dp * 0.5 * uu)));
  acc(FJ(1, 1), s * (coefd * r));
  acc(FJ(2, 2), s * (coefd * r));
  acc(FJ(3, 3), s * (coefd * r));
  acc(FJ(4, 1), s * (coefd * r * Q.u.x));
  acc(FJ(4, 2), s * (coefd * r * Q.u.y));
  acc(FJ(4, 3), s * (coefd * r * Q.u.z));
#undef FJ
}

// Venkatakrishnan limiter function (AIAA 93-0880), eps2 = (Kl h)^3
__device__ __forceinline__ double venkat(double delp, double delm, double eps2) {
  const double num = ((delp * delp + eps2) * delm + 2.0 * delm * delm * delp);
  const double den = (delm * (delp * delp + 2.0 * delm * delm + delm * delp + eps2));
  const double e = (den >= 0.0 ? 1.0e-30 : -1.0e-30);
  return (num + e) / (den + e);
}

}  // namespace fpsi
