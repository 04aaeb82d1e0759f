/*
 * CPU oracle for the flowPsi implicit finite-volume Newton step.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product library.
 *
 * Each function restates the flowPsi rule body it cites (file:line in the
 * flowPsi sources).  Pieces that flowPsi takes from the Loci FVM module
 * (least-squares gradients, limiters, face reconstruction, face gradients,
 * block LU, the order in which apply-rules join) are not in the flowPsi
 * sources; they are defined here from textbook methods, named where they
 * occur, and the GPU library implements the same definitions.
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, so the operation
 * order written here is the order executed).  REAL_FJ_FLOAT selects
 * real_fj = float (flowPsi's default, include/flowTypes.h:153-159).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double real;
#ifdef REAL_FJ_FLOAT
typedef float real_fj;
#else
typedef double real_fj;
#endif

#define NB 5
#define M_(a, i, j) (a)[(i) + NB * (j)] /* Loci::Mat column-major element (i,j) */

static inline real dot3(const real *a, const real *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline real maxr(real a, real b) { return a > b ? a : b; }
static inline real minr(real a, real b) { return a < b ? a : b; }
static inline void addfj(real_fj *a, real v) { *a = (real_fj)((real)*a + v); }
static inline void subfj(real_fj *a, real v) { *a = (real_fj)((real)*a - v); }

/* Face and cell loops whose iterations are independent run in parallel in
 * the OpenMP build that times the CPU baseline (liboracle_omp_*.so); the
 * pragma is inert in the serial builds the tests use.  Sweeps stay
 * sequential, as in the reference. */
#define OG_PAR _Pragma("omp parallel for schedule(static)")
#define OG_BT(g, f) ((g)->btype[f] & 0xFF)        /* boundary type of a face */
#define OG_SET(g, f) (((g)->btype[f] >> 8) & 0xFF) /* its parameter set */

int og_real_fj_bytes(void) { return (int)sizeof(real_fj); }

/* include/inviscidFlux.h:25-45 */
void og_inviscid_flux(double *iflux, double pg, double T, const double *U, const double *an,
                      double area, double pambient, double Rt, double gamma, double Us_n) {
  real gm1 = gamma - 1;
  real rho = (pg + pambient) / (Rt * T);
  real h0 = Rt * T * gamma / gm1 + 0.5 * dot3(U, U);
  const real e0 = h0 - Rt * T;
  const real uta = dot3(U, an);
  const real ut = uta - Us_n;
  real mdot = area * rho * ut;
  real p = pg + pambient;
  iflux[0] = mdot;
  iflux[1] = mdot * U[0] + area * pg * an[0];
  iflux[2] = mdot * U[1] + area * pg * an[1];
  iflux[3] = mdot * U[2] + area * pg * an[2];
  iflux[4] = mdot * e0 + area * p * uta;
}

/* src/fluxInviscid.loci:157-256 (hllc_flux) */
void og_hllc_flux(double *iflux, double pgl, double Tl, const double *Ul, double pgr, double Tr,
                  const double *Ur, const double *an, double area, double pambient, double Rt,
                  double gamma, double Us_n, double Eta) {
  real gm1 = gamma - 1;
  real rl = (pgl + pambient) / (Rt * Tl);
  real rr = (pgr + pambient) / (Rt * Tr);
  real h0l = Rt * Tl * gamma / gm1 + 0.5 * dot3(Ul, Ul);
  real h0r = Rt * Tr * gamma / gm1 + 0.5 * dot3(Ur, Ur);
  const real e0l = h0l - Rt * Tl;
  const real e0r = h0r - Rt * Tr;
  const real srl = sqrt(rl), srr = sqrt(rr);
  const real roe_factor = 1. / (srl + srr);
  real U_[3];
  for (int k = 0; k < 3; ++k) U_[k] = (Ul[k] * srl + Ur[k] * srr) * roe_factor;
  const real ut_ = dot3(U_, an) - Us_n;
  const real h0_ = (h0l * srl + h0r * srr) * roe_factor;
  const real a_ = sqrt(maxr(gm1 * (h0_ - 0.5 * dot3(U_, U_)), 0.0));
  const real EtaP = Eta + 1., EtaM = Eta - 1.;
  const real sigma = sqrt(maxr(ut_ * ut_ * EtaM * EtaM + 4. * Eta * a_ * a_, 1e-30));
  const real evl_ = 0.5 * (ut_ * EtaP - sigma);
  const real evr_ = 0.5 * (ut_ * EtaP + sigma);
  const real utla = dot3(Ul, an), utra = dot3(Ur, an);
  const real utl = utla - Us_n, utr = utra - Us_n;
  const real al2 = gamma * Rt * Tl, ar2 = gamma * Rt * Tr;
  const real sigmar = sqrt(maxr(utr * utr * EtaM * EtaM + 4. * Eta * ar2, 0.0));
  const real sigmal = sqrt(maxr(utl * utl * EtaM * EtaM + 4. * Eta * al2, 0.0));
  real SL = minr(evl_, .5 * (utl * EtaP - sigmal));
  real SR = maxr(evr_, .5 * (utr * EtaP + sigmar));
  real lmdot = rl * utl, rmdot = rr * utr;
  real pl = pgl + pambient, pr = pgr + pambient;
  real coef = 0.5 * area;
  real AL = coef * ((SL > 0) ? (SR < 0 ? 1. : 2.) : (SR < 0 ? 0. : 1.));
  real AR = coef * ((SR < 0) ? (SL > 0 ? 1. : 2.) : (SL > 0 ? 0. : 1.));
  real AM = coef * ((SL > 0 || SR < 0) ? 0. : 1.);
  iflux[0] = AL * lmdot + AR * rmdot;
  iflux[1] = AL * lmdot * Ul[0] + AR * rmdot * Ur[0] + an[0] * (AL * pgl + AR * pgr);
  iflux[2] = AL * lmdot * Ul[1] + AR * rmdot * Ur[1] + an[1] * (AL * pgl + AR * pgr);
  iflux[3] = AL * lmdot * Ul[2] + AR * rmdot * Ur[2] + an[2] * (AL * pgl + AR * pgr);
  iflux[4] = AL * (lmdot * e0l + pl * utla) + AR * (rmdot * e0r + pr * utra);
  iflux[0] -= AM * (SL * rl + SR * rr);
  iflux[1] -= AM * (SL * rl * Ul[0] + SR * rr * Ur[0]);
  iflux[2] -= AM * (SL * rl * Ul[1] + SR * rr * Ur[1]);
  iflux[3] -= AM * (SL * rl * Ul[2] + SR * rr * Ur[2]);
  iflux[4] -= AM * (SL * rl * e0l + SR * rr * e0r);
  const real SM = ((rr * utr * (SR - utr) - rl * utl * (SL - utl) + pgl - pgr) /
                   (rr * (SR - utr) - rl * (SL - utl)));
  const real rl_star = rl * (SL - utl) / (SL - SM);
  const real rr_star = rr * (SR - utr) / (SR - SM);
  const real pg_star = rl * (utl - SL) * (utl - SM) + pgl;
  const real p_star = pg_star + pambient;
  const real rul_star = ((SL - utl) * rl * Ul[0] + (pg_star - pgl) * an[0]) / (SL - SM);
  const real rvl_star = ((SL - utl) * rl * Ul[1] + (pg_star - pgl) * an[1]) / (SL - SM);
  const real rwl_star = ((SL - utl) * rl * Ul[2] + (pg_star - pgl) * an[2]) / (SL - SM);
  const real rur_star = ((SR - utr) * rr * Ur[0] + (pg_star - pgr) * an[0]) / (SR - SM);
  const real rvr_star = ((SR - utr) * rr * Ur[1] + (pg_star - pgr) * an[1]) / (SR - SM);
  const real rwr_star = ((SR - utr) * rr * Ur[2] + (pg_star - pgr) * an[2]) / (SR - SM);
  const real re0l_star = ((SL - utl) * rl * e0l - pl * utl + p_star * SM) / (SL - SM);
  const real re0r_star = ((SR - utr) * rr * e0r - pr * utr + p_star * SM) / (SR - SM);
  const real CR = SR - fabs(SM);
  const real CL = SL + fabs(SM);
  iflux[0] += AM * (CR * rr_star + CL * rl_star);
  iflux[1] += AM * (CR * rur_star + CL * rul_star);
  iflux[2] += AM * (CR * rvr_star + CL * rvl_star);
  iflux[3] += AM * (CR * rwr_star + CL * rwl_star);
  iflux[4] += AM * (CR * re0r_star + CL * re0l_star);
}

/* src/fluxInviscid.loci:869-1014 (roeCT_flux, ENTROPY_FIX defined) */
void og_roe_flux(double *iflux, double pgl, double Tl, const double *Ul, double pgr, double Tr,
                 const double *Ur, const double *an, double area, double pambient, double Rt,
                 double gamma, double Us_n, double Eta) {
  const real pl = pgl + pambient, pr = pgr + pambient;
  const real rl = pl / (Rt * Tl), rr = pr / (Rt * Tr);
  const real gm1 = gamma - 1.;
  const real re0l = rl * (Rt * Tl / gm1 + 0.5 * dot3(Ul, Ul));
  const real re0r = rr * (Rt * Tr / gm1 + 0.5 * dot3(Ur, Ur));
  const real h0l = (re0l + pl) / rl, h0r = (re0r + pr) / rr;
  const real r_ = sqrt(rl * rr);
  const real srl = sqrt(rl), srr = sqrt(rr);
  const real roe_factor = 1. / (srl + srr);
  real U_[3];
  for (int k = 0; k < 3; ++k) U_[k] = (Ul[k] * srl + Ur[k] * srr) * roe_factor;
  const real ut_ = dot3(U_, an) - Us_n;
  const real uta_ = dot3(U_, an);
  const real q2_ = dot3(U_, U_);
  const real h0_ = (h0l * srl + h0r * srr) * roe_factor;
  const real h_ = h0_ - 0.5 * q2_;
  const real al2 = gamma * Rt * Tl, ar2 = gamma * Rt * Tr;
  const real utl = dot3(Ul, an) - Us_n, utr = dot3(Ur, an) - Us_n;
  const real utla = dot3(Ul, an), utra = dot3(Ur, an);
  if (fabs(pl - pr) / (pl + pr + 1e-33) > .1) Eta = 1.;
  real af2 = gm1 * h_;
  af2 = maxr(af2, 1e-30);
  const real EtaP = Eta + 1., EtaM = Eta - 1.;
  real a_ = sqrt(af2);
  const real sigma = sqrt(ut_ * ut_ * EtaM * EtaM + 4. * Eta * a_ * a_);
  real ev1 = fabs(ut_);
  real ev2 = fabs(ut_ * EtaP - sigma) / 2.;
  real ev3 = fabs(ut_ * EtaP + sigma) / 2.;
  /* Harten-Hyman entropy fix (src/fluxInviscid.loci:940-962) */
  real sigmal = sqrt(utl * utl * EtaM * EtaM + 4. * Eta * al2);
  real sigmar = sqrt(utr * utr * EtaM * EtaM + 4. * Eta * ar2);
  const real e2l = (utl * EtaP - sigmal) / 2., e2r = (utr * EtaP - sigmar) / 2.;
  const real e3l = (utl * EtaP + sigmal) / 2., e3r = (utr * EtaP + sigmar) / 2.;
  const real ev2s = (ut_ * EtaP - sigma) / 2., ev3s = (ut_ * EtaP + sigma) / 2.;
  const real eps2 = maxr(0.0, maxr(ev2s - e2l, e2r - ev2s));
  const real eps3 = maxr(0.0, maxr(ev3s - e3l, e3r - ev3s));
  ev2 = ev2 < eps2 ? 0.5 * ((ev2s * ev2s) / eps2 + eps2) : ev2;
  ev3 = ev3 < eps3 ? 0.5 * ((ev3s * ev3s) / eps3 + eps3) : ev3;
  real Uj[3];
  for (int k = 0; k < 3; ++k) Uj[k] = Ur[k] - Ul[k];
  const real utj = dot3(Uj, an);
  const real pj = pgr - pgl;
  const real Tj = Tr - Tl;
  const real BM = EtaM;
  const real T1 = ev2 - ev3;
  const real T2 = (ev2 + ev3 - 2 * ev1);
  const real T3 = r_ * (BM * BM * ut_ * ut_ - sigma * sigma) / (4. * sigma);
  const real T4 = T2 * sigma + BM * T1 * ut_;
  const real T5 = (T2 * sigma - BM * T1 * ut_) / (2. * sigma);
  const real T6 = ((ev2 + ev3) * sigma - BM * T1 * ut_) / (2. * sigma);
  const real Cp = gamma * Rt / gm1;
  const real rT_ = -gamma * Rt * r_ / (h_ * gm1);
  const real rbeta = 1. / Eta;
  const real ra2 = 1. / (a_ * a_);
  const real T7 = rbeta * T1 * T3 * ra2;
  const real T8 = (((rbeta - 1) + gamma) * T6 - gm1 * T5) * ra2;
  const real T9 = Cp * r_ + h0_ * rT_;
  const real x1 = gamma * h0_ * T1 * T3 * ra2 * rbeta;
  const real x2 = -rbeta * gm1 * T1 * T3 * T9 * ra2 / rT_;
  const real x3 = -T1 * T3 * rbeta;
  const real x4 = .5 * r_ * T4 * uta_ / sigma;
  const real T10 = (x1 + x2 + x3 + x4);
  const real T11 = (rbeta * ra2 * h0_ + (T9 - r_ * rbeta * Cp) * ra2 * gm1 / rT_ - rbeta) * T6 -
                   gm1 * ra2 * T9 * T5 / rT_ - uta_ * T1 / sigma;
  iflux[0] = pj * T8 + utj * T7 + ev1 * Tj * rT_;
  iflux[1] = U_[0] * T7 * utj + ev1 * (U_[0] * Tj * rT_ + r_ * Uj[0]) +
             pj * (T8 * U_[0] - an[0] * T1 / sigma) + an[0] * r_ * T4 * utj / (2. * sigma);
  iflux[2] = U_[1] * T7 * utj + ev1 * (U_[1] * Tj * rT_ + r_ * Uj[1]) +
             pj * (T8 * U_[1] - an[1] * T1 / sigma) + an[1] * r_ * T4 * utj / (2. * sigma);
  iflux[3] = U_[2] * T7 * utj + ev1 * (U_[2] * Tj * rT_ + r_ * Uj[2]) +
             pj * (T8 * U_[2] - an[2] * T1 / sigma) + an[2] * r_ * T4 * utj / (2. * sigma);
  iflux[4] = pj * T11 + T10 * utj + ev1 * (T9 * Tj + r_ * dot3(U_, Uj));
  const real Ah = 0.5 * area;
  iflux[0] = Ah * (rl * utl + rr * utr - iflux[0]);
  iflux[1] = Ah * (Ul[0] * rl * utl + Ur[0] * rr * utr + an[0] * (pgl + pgr) - iflux[1]);
  iflux[2] = Ah * (Ul[1] * rl * utl + Ur[1] * rr * utr + an[1] * (pgl + pgr) - iflux[2]);
  iflux[3] = Ah * (Ul[2] * rl * utl + Ur[2] * rr * utr + an[2] * (pgl + pgr) - iflux[3]);
  iflux[4] = Ah * ((re0l * utl + pl * utla) + (re0r * utr + pr * utra) - iflux[4]);
}

/* src/fluxInviscid.loci:1018-1090 (RoeJacobian); Fj accumulated (+=) in real_fj */
static void roe_jacobian_ev(real_fj *Fj, real ev1, real ev2, real ev3, const real *U, real un,
                            real a, real beta, real sigma, real gamma, real Rt, real rho, real T,
                            real h0, const real *n, real area) {
  (void)T;
  const int Ti = 0, Pi = 4, mi = 1, ei = 4;
  const real BM = beta - 1.;
  const real una = dot3(U, n);
  const real T1 = ev2 - ev3;
  const real T2 = (ev2 + ev3 - 2 * ev1);
  const real T3 = rho * (BM * BM * un * un - sigma * sigma) / (4. * sigma);
  const real T4 = T2 * sigma + BM * T1 * un;
  const real T5 = (T2 * sigma - BM * T1 * un) / (2. * sigma);
  const real T6 = ((ev2 + ev3) * sigma - BM * T1 * un) / (2. * sigma);
  const real gm1 = gamma - 1.0;
  const real Cp = gamma * Rt / gm1;
  const real rT = -gamma * Rt * rho / ((h0 - 0.5 * dot3(U, U)) * gm1);
  const real rbeta = 1. / beta;
  const real ra2 = 1. / (a * a);
  const real T7 = rbeta * T1 * T3 * ra2;
  const real T8 = (((rbeta - 1) + gamma) * T6 - gm1 * T5) * ra2;
  const real T9 = Cp * rho + h0 * rT;
  const real x1 = gamma * h0 * T1 * T3 * ra2 * rbeta;
  const real x2 = -rbeta * gm1 * T1 * T3 * T9 * ra2 / rT;
  const real x3 = -T1 * T3 * rbeta;
  const real x4 = .5 * rho * T4 * una / sigma;
  const real T10 = (x1 + x2 + x3 + x4);
  const real T11 = (rbeta * ra2 * h0 + (T9 - rho * rbeta * Cp) * ra2 * gm1 / rT - rbeta) * T6 -
                   gm1 * ra2 * T9 * T5 / rT - una * T1 / sigma;
  addfj(&M_(Fj, 0, Ti), area * ev1 * rT);
  addfj(&M_(Fj, 0, mi + 0), area * n[0] * T7);
  addfj(&M_(Fj, 0, mi + 1), area * n[1] * T7);
  addfj(&M_(Fj, 0, mi + 2), area * n[2] * T7);
  addfj(&M_(Fj, 0, Pi), area * T8);
  addfj(&M_(Fj, mi + 0, Ti), area * ev1 * rT * U[0]);
  addfj(&M_(Fj, mi + 1, Ti), area * ev1 * rT * U[1]);
  addfj(&M_(Fj, mi + 2, Ti), area * ev1 * rT * U[2]);
  addfj(&M_(Fj, mi + 0, mi + 0), area * (n[0] * (T7 * U[0] + n[0] * rho * T4 / (2. * sigma)) + ev1 * rho));
  addfj(&M_(Fj, mi + 0, mi + 1), area * n[1] * (T7 * U[0] + n[0] * rho * T4 / (2. * sigma)));
  addfj(&M_(Fj, mi + 0, mi + 2), area * n[2] * (T7 * U[0] + n[0] * rho * T4 / (2. * sigma)));
  addfj(&M_(Fj, mi + 1, mi + 0), area * n[0] * (T7 * U[1] + n[1] * rho * T4 / (2. * sigma)));
  addfj(&M_(Fj, mi + 1, mi + 1), area * (n[1] * (T7 * U[1] + n[1] * rho * T4 / (2. * sigma)) + ev1 * rho));
  addfj(&M_(Fj, mi + 1, mi + 2), area * n[2] * (T7 * U[1] + n[1] * rho * T4 / (2. * sigma)));
  addfj(&M_(Fj, mi + 2, mi + 0), area * n[0] * (T7 * U[2] + n[2] * rho * T4 / (2. * sigma)));
  addfj(&M_(Fj, mi + 2, mi + 1), area * n[1] * (T7 * U[2] + n[2] * rho * T4 / (2. * sigma)));
  addfj(&M_(Fj, mi + 2, mi + 2), area * (n[2] * (T7 * U[2] + n[2] * rho * T4 / (2. * sigma)) + ev1 * rho));
  addfj(&M_(Fj, mi + 0, Pi), area * (T8 * U[0] - n[0] * T1 / sigma));
  addfj(&M_(Fj, mi + 1, Pi), area * (T8 * U[1] - n[1] * T1 / sigma));
  addfj(&M_(Fj, mi + 2, Pi), area * (T8 * U[2] - n[2] * T1 / sigma));
  addfj(&M_(Fj, ei, Ti), area * ev1 * T9);
  addfj(&M_(Fj, ei, mi + 0), area * (n[0] * T10 + ev1 * rho * U[0]));
  addfj(&M_(Fj, ei, mi + 1), area * (n[1] * T10 + ev1 * rho * U[1]));
  addfj(&M_(Fj, ei, mi + 2), area * (n[2] * T10 + ev1 * rho * U[2]));
  addfj(&M_(Fj, ei, Pi), area * T11);
}

void og_roe_jacobian_ev(void *fj25, double ev1, double ev2, double ev3, const double *U, double un,
                        double a, double beta, double sigma, double gamma, double Rt, double rho,
                        double T, double h0, const double *n, double area) {
  roe_jacobian_ev((real_fj *)fj25, ev1, ev2, ev3, U, un, a, beta, sigma, gamma, Rt, rho, T, h0, n, area);
}

/* src/fluxInviscid.loci:1092-1168 (roeCT_jacobian, smoothed eigenvalues) */
static void roe_jacobian(real_fj *Fj, real pgl, real Tl, const real *Ul, real pgr, real Tr,
                         const real *Ur, const real *an, real area, real pambient, real Rt,
                         real gamma, real Us_n, real Eta, int dir) {
  real gm1 = gamma - 1;
  const real pl = pgl + pambient, pr = pgr + pambient;
  const real rl = pl / (Rt * Tl), rr = pr / (Rt * Tr);
  const real re0l = rl * (Rt * Tl / gm1 + 0.5 * dot3(Ul, Ul));
  const real re0r = rr * (Rt * Tr / gm1 + 0.5 * dot3(Ur, Ur));
  const real h0l = (re0l + pl) / rl, h0r = (re0r + pr) / rr;
  const real r_ = sqrt(rl * rr);
  const real srl = sqrt(rl), srr = sqrt(rr);
  const real roe_factor = 1. / (srl + srr);
  real U_[3];
  for (int k = 0; k < 3; ++k) U_[k] = (Ul[k] * srl + Ur[k] * srr) * roe_factor;
  const real ut_ = dot3(U_, an) - Us_n;
  const real q2_ = dot3(U_, U_);
  const real h0_ = (h0l * srl + h0r * srr) * roe_factor;
  const real h_ = h0_ - 0.5 * q2_;
  if (fabs(pl - pr) / (pl + pr + 1e-33) > .1) Eta = 1.;
  real af2 = gm1 * h_;
  af2 = maxr(af2, 1e-30);
  const real T_ = h_ * gm1 / (Rt * gamma);
  const real EtaP = Eta + 1., EtaM = Eta - 1.;
  real a_ = sqrt(af2);
  const real sigma = sqrt(ut_ * ut_ * EtaM * EtaM + 4. * Eta * a_ * a_);
  real ev1 = ut_;
  real ev2 = (ut_ * EtaP - sigma) / 2.;
  real ev3 = (ut_ * EtaP + sigma) / 2.;
  real d = (dir < 0) ? -1 : 1;
  const real mxv = maxr(ev1, maxr(ev2, ev3));
  const real mnv = minr(ev1, minr(ev2, ev3));
  const real smooth_factor = 0.02;
  const real eps = smooth_factor * (mxv - mnv);
  const real eps2 = eps * eps;
  ev1 = .5 * (ev1 + d * sqrt(ev1 * ev1 + eps2));
  ev2 = .5 * (ev2 + d * sqrt(ev2 * ev2 + eps2));
  ev3 = .5 * (ev3 + d * sqrt(ev3 * ev3 + eps2));
  roe_jacobian_ev(Fj, ev1, ev2, ev3, U_, ut_, a_, Eta, sigma, gamma, Rt, r_, T_, h0_, an, area);
}

void og_roe_jacobian(void *fj25, double pgl, double Tl, const double *Ul, double pgr, double Tr,
                     const double *Ur, const double *an, double area, double pamb, double Rt,
                     double gamma, double us_n, double eta, int dir) {
  roe_jacobian((real_fj *)fj25, pgl, Tl, Ul, pgr, Tr, Ur, an, area, pamb, Rt, gamma, us_n, eta, dir);
}

/* include/bcTools.h:25-51 */
void og_cv_inflow(double *pgb, double *Tb, double *ub, double pgi, double Ti, const double *ui,
                  double pg0, double T0, const double *u0, const double *n, double us_n,
                  double Pambient, double gamma, double Rtilde, double Beta) {
  (void)Ti;
  const real uit = dot3(ui, n) - us_n;
  const real u0t = dot3(u0, n) - us_n;
  const real a02 = gamma * Rtilde * T0;
  const real a0 = sqrt(a02);
  const real r0 = (pg0 + Pambient) / (Rtilde * T0);
  const real MinPressure = 0.1 * (pg0 + Pambient) - Pambient;
  const real BetaM = 0.5 * (1.0 - Beta);
  const real uitBeta = uit * (1. - Beta);
  const real sigma2 = uitBeta * uitBeta + 4. * Beta * a0 * a0;
  const real sigma = sqrt(sigma2);
  real p = 0.5 * (pg0 + pgi - 0.5 * r0 * sigma * (u0t - uit)) -
           ((pgi - pg0) + r0 * (uit - u0t) * uit * BetaM) * uit * BetaM / (2. * sigma);
  p = maxr(p, MinPressure);
  for (int k = 0; k < 3; ++k) ub[k] = u0[k] + 2. * n[k] * (p - pg0) / (r0 * (sigma + uit * BetaM));
  real rhob = r0 + (p - pg0) * (sigma - uit * BetaM) / (Beta * a0 * a0 * (sigma + uit * BetaM));
  *Tb = (p + Pambient) / (Rtilde * rhob);
  *pgb = p;
}

/* include/bcTools.h:52-80 */
void og_cv_outflow(double *pgb, double *Tb, double *ub, double pgi, double Ti, const double *ui,
                   double pg0, double T0, const double *u0, const double *n, double us_n,
                   double Pambient, double gamma, double Rtilde, double Beta) {
  (void)Ti;
  const real uit = dot3(ui, n) - us_n;
  const real u0t = dot3(u0, n) - us_n;
  const real a02 = gamma * Rtilde * T0;
  const real a0 = sqrt(a02);
  const real r0 = (pg0 + Pambient) / (Rtilde * T0);
  const real MinPressure = 0.1 * (pgi + Pambient) - Pambient;
  const real BetaM = 0.5 * (1.0 - Beta);
  const real uitBeta = uit * (1. - Beta);
  const real sigma2 = uitBeta * uitBeta + 4. * Beta * a0 * a0;
  const real sigma = sqrt(sigma2);
  real p = .5 * (pgi + pg0 - 0.5 * r0 * sigma * (u0t - uit)) +
           ((pgi - pg0) + r0 * (uit - u0t) * uit * BetaM) * uit * BetaM / (2. * sigma);
  p = maxr(p, MinPressure);
  for (int k = 0; k < 3; ++k) ub[k] = ui[k] + 2. * n[k] * (pgi - p) / (r0 * (sigma + uit * BetaM));
  const real rhob = r0 + (p - pgi) * (sigma - uit * BetaM) / (Beta * a0 * a0 * (sigma + uit * BetaM));
  *Tb = (p + Pambient) / (Rtilde * rhob);
  *pgb = p;
}

/* ------------------------------------------------------------------ */
/* grid helpers                                                        */
/* ------------------------------------------------------------------ */

/* Per-cell face lists: lower (cr == c), upper (cl == c, periodic faces
 * included, as their ghost sits on the cr side), boundary (cl == c, not
 * periodic); each ascending in face index. */
typedef struct {
  int *lo_off, *lo, *up_off, *up, *bd_off, *bd;
} og_adj;

static void build_adj(const og_grid *g, og_adj *a) {
  int nc = g->nc, nf = g->nf;
  a->lo_off = calloc(nc + 1, sizeof(int));
  a->up_off = calloc(nc + 1, sizeof(int));
  a->bd_off = calloc(nc + 1, sizeof(int));
  for (int f = 0; f < nf; ++f) {
    int t = OG_BT(g, f);
    if (t == OG_INTERIOR) {
      a->up_off[g->cl[f] + 1]++;
      a->lo_off[g->cr[f] + 1]++;
    } else if (t == OG_PERIODIC) {
      a->up_off[g->cl[f] + 1]++;
    } else {
      a->bd_off[g->cl[f] + 1]++;
    }
  }
  for (int c = 0; c < nc; ++c) {
    a->lo_off[c + 1] += a->lo_off[c];
    a->up_off[c + 1] += a->up_off[c];
    a->bd_off[c + 1] += a->bd_off[c];
  }
  a->lo = malloc(sizeof(int) * (a->lo_off[nc] + 1));
  a->up = malloc(sizeof(int) * (a->up_off[nc] + 1));
  a->bd = malloc(sizeof(int) * (a->bd_off[nc] + 1));
  int *lp = calloc(nc, sizeof(int)), *upp = calloc(nc, sizeof(int)), *bp = calloc(nc, sizeof(int));
  for (int f = 0; f < nf; ++f) {
    int t = OG_BT(g, f);
    if (t == OG_INTERIOR) {
      int c = g->cl[f];
      a->up[a->up_off[c] + upp[c]++] = f;
      c = g->cr[f];
      a->lo[a->lo_off[c] + lp[c]++] = f;
    } else if (t == OG_PERIODIC) {
      int c = g->cl[f];
      a->up[a->up_off[c] + upp[c]++] = f;
    } else {
      int c = g->cl[f];
      a->bd[a->bd_off[c] + bp[c]++] = f;
    }
  }
  free(lp);
  free(upp);
  free(bp);
}

static void free_adj(og_adj *a) {
  free(a->lo_off); free(a->lo); free(a->up_off); free(a->up); free(a->bd_off); free(a->bd);
}

static inline void fnormal(const og_grid *g, int f, real *n) {
  n[0] = g->n[f]; n[1] = g->n[g->nf + f]; n[2] = g->n[2 * g->nf + f];
}
static inline void fcenter(const og_grid *g, int f, real *x) {
  x[0] = g->fc[f]; x[1] = g->fc[g->nf + f]; x[2] = g->fc[2 * g->nf + f];
}
static inline void ccenter(const og_grid *g, int c, real *x) {
  x[0] = g->cc[c]; x[1] = g->cc[g->nc + c]; x[2] = g->cc[2 * g->nc + c];
}
/* cell on the cr side of an interior or periodic face, and its centre
 * (src/solverPeriodic.loci:31-35: ghost centre = transformed partner centre) */
static inline int rcell(const og_grid *g, int f, real *x) {
  if (OG_BT(g, f) == OG_PERIODIC) {
    int c = g->cl[g->pface[f]];
    ccenter(g, c, x);
    for (int d = 0; d < 3; ++d) x[d] += g->pshift[d * g->nf + f];
    return c;
  }
  ccenter(g, g->cr[f], x);
  return g->cr[f];
}

/* variables per cell: T, u, v, w, pg (+ nuTilde, or + k, omega) */
static inline int nq_of(const og_params *p) { return p->turb == 2 ? 7 : p->turb ? 6 : 5; }

/* Neighbour of cell c across face f for the gradient stencil: its position
 * and primitive values (cell, periodic ghost, or boundary-face state). */
static void stencil_nb(const og_grid *g, const double *q, const double *qf, int c, int f, int nq,
                       real *x, real *v) {
  int nc = g->nc, nf = g->nf;
  int t = OG_BT(g, f);
  if (t == OG_INTERIOR) {
    int o = (g->cl[f] == c) ? g->cr[f] : g->cl[f];
    ccenter(g, o, x);
    for (int k = 0; k < nq; ++k) v[k] = q[k * nc + o];
  } else if (t == OG_PERIODIC) {
    int o = rcell(g, f, x);
    for (int k = 0; k < nq; ++k) v[k] = q[k * nc + o];
  } else {
    fcenter(g, f, x);
    for (int k = 0; k < nq; ++k) v[k] = qf[k * nf + f];
  }
}

static real tmu_cell(const og_grid *g, const og_params *p, const double *q, const double *grad, int c);

/* ------------------------------------------------------------------ */
/* low-Mach preconditioning parameter, src/fluxInviscid.loci:35-155     */
/* ------------------------------------------------------------------ */

static inline real suth_mu_(const og_params *p, real T) { return p->suth[0] * pow(T, p->suth[1]) / (T + p->suth[2]); }

/* Mlim2 (:98-104) */
static real mlim2(const og_params *p) {
  const real M2 = p->Minf * p->Minf;
  real r = 1.0;
  if (p->Minf < 0.5) r = p->etaT * 2. * M2 / (1. - 2. * M2);
  return r;
}

/* Eta_pf / Eta_bc of a boundary face from its interior cell (:145-155) */
double og_eta_face(const og_params *p, double T, const double *U) {
  const real a = sqrt(p->gamma * p->Rt * T);
  const real M = sqrt(dot3(U, U)) / a;
  const real M2 = 2. * minr(0.25, M * M);
  return minr(1.0, maxr(M2 / (1. - M2), mlim2(p)));
}

/* Eta_p per cell: Mref = |u|/a (:35-37) raised by the pressure-jump and
 * laminar viscous limits over every interior and periodic face (:43-79;
 * the periodic ghost carries its partner's state and volume,
 * src/solverPeriodic.loci:39-66), Eta_l (:106-110), then the maximum over
 * the face neighbours twice (Eta_ll :116-126, Eta_p :132-142).  Maxima are
 * exact, so the join order does not matter. */
static void eta_impl(const og_grid *g, const og_params *p, const double *q, const double *grad, double *eta_p) {
  int nc = g->nc, nf = g->nf;
  const real ml2 = mlim2(p);
  if (ml2 >= 1.0) { /* Eta_l = max(M2/(1-M2), 1) = 1 for M2 <= 1/2 */
    for (int c = 0; c < nc; ++c) eta_p[c] = 1.0;
    return;
  }
  real *mref = malloc(sizeof(real) * nc), *el = malloc(sizeof(real) * nc), *ell = malloc(sizeof(real) * nc);
  real *rho = malloc(sizeof(real) * nc), *a = malloc(sizeof(real) * nc), *mu = malloc(sizeof(real) * nc);
  for (int c = 0; c < nc; ++c) {
    const real T = q[c], pg = q[4 * nc + c];
    real U[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
    rho[c] = (pg + p->Pambient) / (p->Rt * T);
    a[c] = sqrt(p->gamma * p->Rt * T);
    mu[c] = suth_mu_(p, T);
    mref[c] = sqrt(dot3(U, U)) / a[c];
  }
  for (int f = 0; f < nf; ++f) {
    const int t = OG_BT(g, f);
    if (t != OG_INTERIOR && t != OG_PERIODIC) continue;
    const int l = g->cl[f];
    const int r = t == OG_INTERIOR ? g->cr[f] : g->cl[g->pface[f]];
    const real dP = fabs(q[4 * nc + l] - q[4 * nc + r]);
    /* cl side; the cr side of an interior face (a periodic ghost's own
     * value is never used: each partner joins from its own face) */
    mref[l] = maxr(mref[l], p->PLimPC * dP / (rho[l] * a[l] * a[l]));
    if (t == OG_INTERIOR) mref[r] = maxr(mref[r], p->PLimPC * dP / (rho[r] * a[r] * a[r]));
    if (p->viscous) {
      const real area = maxr(g->area[f], 1e-30);
      const real rhom = 0.5 * (rho[l] + rho[r]);
      if (!p->turb) {
        const real mum = maxr(mu[l], mu[r]);
        const real nu = mum / rhom;
        mref[l] = maxr(mref[l], nu / ((g->vol[l] / area) * a[l]));
        if (t == OG_INTERIOR) mref[r] = maxr(mref[r], nu / ((g->vol[r] / area) * a[r]));
      } else { /* :81-96; the cl side takes max(tmu_cl, tmu_cl) as written */
        const real tl = tmu_cell(g, p, q, grad, l), tr = tmu_cell(g, p, q, grad, r);
        const real nul = (maxr(mu[l], mu[r]) + maxr(tl, tl)) / rhom;
        mref[l] = maxr(mref[l], nul / ((g->vol[l] / area) * a[l]));
        if (t == OG_INTERIOR) {
          const real nur = (maxr(mu[l], mu[r]) + maxr(tl, tr)) / rhom;
          mref[r] = maxr(mref[r], nur / ((g->vol[r] / area) * a[r]));
        }
      }
    }
  }
  for (int c = 0; c < nc; ++c) {
    const real M2 = 2. * minr(0.25, mref[c] * mref[c]);
    el[c] = maxr(M2 / (1. - M2), ml2);
  }
  real *src = el, *dst = ell;
  for (int pass = 0; pass < 2; ++pass) {
    for (int c = 0; c < nc; ++c) dst[c] = src[c];
    for (int f = 0; f < nf; ++f) {
      const int t = OG_BT(g, f);
      if (t != OG_INTERIOR && t != OG_PERIODIC) continue;
      const int l = g->cl[f];
      const int r = t == OG_INTERIOR ? g->cr[f] : g->cl[g->pface[f]];
      dst[l] = maxr(dst[l], src[r]);
      if (t == OG_INTERIOR) dst[r] = maxr(dst[r], src[l]);
    }
    if (pass == 0) { src = ell; dst = eta_p; }
  }
  free(mref); free(el); free(ell); free(rho); free(a); free(mu);
}

void og_face_states(const og_grid *g, const og_params *p, const double *q, double *qf);
void og_gradients(const og_grid *g, const og_params *p, const double *q, const double *qf, double *grad,
                  double *phi);
/* Eta_p of a state; under SST the eddy viscosity needs the velocity
 * gradient, so the face states and gradients are made first */
void og_eta(const og_grid *g, const og_params *p, const double *q, double *eta_p) {
  if (p->turb != 2) {
    eta_impl(g, p, q, NULL, eta_p);
    return;
  }
  const int nq = nq_of(p);
  double *qf = calloc((size_t)nq * g->nf, sizeof(double));
  double *grad = malloc(sizeof(double) * 3 * nq * (size_t)g->nc), *phi = malloc(sizeof(double) * nq * (size_t)g->nc);
  og_face_states(g, p, q, qf);
  og_gradients(g, p, q, qf, grad, phi);
  eta_impl(g, p, q, grad, eta_p);
  free(qf); free(grad); free(phi);
}

/* ------------------------------------------------------------------ */
/* boundary conditions with parameter sets (types 5-12)                */
/* ------------------------------------------------------------------ */

/* inviscidFlux, include/inviscidFlux.h:25-45 */
static void euler_flux_(real *iflux, real pg, real T, const real *U, const real *an, real area,
                        real pambient, real Rt, real gamma, real Us_n) {
  real gm1 = gamma - 1;
  real rho = (pg + pambient) / (Rt * T);
  real h0 = Rt * T * gamma / gm1 + 0.5 * dot3(U, U);
  const real e0 = h0 - Rt * T;
  const real uta = dot3(U, an);
  const real ut = uta - Us_n;
  real mdot = area * rho * ut;
  real pabs = pg + pambient;
  iflux[0] = mdot;
  iflux[1] = mdot * U[0] + area * pg * an[0];
  iflux[2] = mdot * U[1] + area * pg * an[1];
  iflux[3] = mdot * U[2] + area * pg * an[2];
  iflux[4] = mdot * e0 + area * pabs * uta;
}

/* the option's Riemann flux (inviscidRiemannFlux) */
static void riemann_(const og_params *p, real *fl, real pgl, real Tl, const real *ul, real pgr, real Tr,
                     const real *ur, const real *an, real area, real us_n, real eta) {
  if (p->flux == 1) og_roe_flux(fl, pgl, Tl, ul, pgr, Tr, ur, an, area, p->Pambient, p->Rt, p->gamma, us_n, eta);
  else og_hllc_flux(fl, pgl, Tl, ul, pgr, Tr, ur, an, area, p->Pambient, p->Rt, p->gamma, us_n, eta);
}

/* src/bcImpermeable.loci:119-130: the flux against the mirrored left state */
static void reflecting_flux_(const og_params *p, real *fl, real pgl, real Tl, const real *ul, const real *n,
                             real area, real us_n, real eta) {
  const real uit = dot3(ul, n) - us_n;
  real ur[3];
  for (int k = 0; k < 3; ++k) ur[k] = ul[k] - 2. * uit * n[k];
  riemann_(p, fl, pgl, Tl, ul, pgl, Tl, ur, n, area, us_n, eta);
}

/* fixed_p0t0_bc, src/bcIsentropic.loci:66-89 */
static void fixed_p0t0_bc_(real *unb, real *ab, real uni, real ai, real a0, real gamma0) {
  const real gm1 = gamma0 - 1.;
  const real rim = uni - 2. * ai / gm1;
  real a = (2. / gm1 + 1);
  real b = (2. * rim);
  real c = (0.5 * gm1 * rim * rim - a0 * a0);
  real D = b * b - 4. * a * c;
  if (D < 0) {
    *ab = a0;
    *unb = rim + 2. * (*ab) / gm1;
    return;
  }
  *ab = (-b + sqrt(D)) / (2. * a);
  *ab = maxr(*ab, .25 * a0);
  *unb = rim + 2. * (*ab) / gm1;
}

/* total_pressure_bc_imp, src/bcIsentropic.loci:91-144, no rigid rotation */
static void total_pressure_bc_(real *pb, real *Tb, real *ub, real T0, real p0, real gamma0, real a0,
                               const real *ui, real ai, real ti, const real *n) {
  const real gm1 = gamma0 - 1.;
  const real uit = dot3(ui, n);
  real unb = uit, ab = a0;
  fixed_p0t0_bc_(&unb, &ab, uit, ai, a0, gamma0);
  for (int k = 0; k < 3; ++k) ub[k] = unb * n[k] + 0.0;
  real M = sqrt(dot3(ub, ub)) / ab;
  const real maxMach = 1.0;
  if (M > maxMach) {
    M = maxMach;
    const real utot = M * a0 / sqrt(1 + .5 * gm1 * M * M);
    const real un = sqrt(maxr(utot * utot - 0.0, 0.0));
    for (int k = 0; k < 3; ++k) ub[k] = un * n[k] + 0.0;
  }
  *Tb = T0 / (1 + .5 * gm1 * M * M);
  *pb = p0 * pow(1. + .5 * gm1 * M * M, -gamma0 / gm1);
  real f = (*Tb - ti) / maxr(ti, *Tb) - .1;
  if (f > 0) {
    f = pow(f, .5);
    ab = (1 - f) * ab + f * a0;
    unb = (1 - f) * unb + f * uit;
    for (int k = 0; k < 3; ++k) ub[k] = unb * n[k] + 0.0;
    real M2 = sqrt(dot3(ub, ub)) / ab;
    if (M2 > maxMach) {
      M2 = maxMach;
      const real utot = M2 * a0 / sqrt(1 + .5 * gm1 * M2 * M2);
      const real un = sqrt(maxr(utot * utot - 0.0, 0.0));
      for (int k = 0; k < 3; ++k) ub[k] = un * n[k] + 0.0;
    }
    *Tb = T0 / (1 + .5 * gm1 * M2 * M2);
    *pb = p0 * pow(1. + .5 * gm1 * M2 * M2, -gamma0 / gm1);
  }
}

/* mass_flowrate_bc, src/bcFixedMass.loci:85-136 */
static void mass_flowrate_bc_(real *ub, real *Tb, real *Pb, real pin, real Tin, real massflow,
                              const real *flow_dir, real Tstag, real Pambient, real gamma, real Rtilde,
                              const real *n) {
  real T = Tin;
  const real eps = Tin * 1e-9;
  for (int iter = 0; iter < 1000; ++iter) {
    const real rho = pin / (Rtilde * T);
    real a = sqrt(gamma * Rtilde * T);
    real u = massflow / rho;
    real M = u / a;
    real lamda = sqrt(((gamma + 1.0) * M * M) / ((gamma - 1.0) * M * M + 2));
    real p0 = pin * pow((1. - ((gamma - 1.) / (gamma + 1.)) * lamda * lamda), -gamma / (gamma - 1.0));
    real rho0 = p0 / (Rtilde * Tstag);
    const real rho_1 = rho0 * pow((1. - (gamma - 1.0) / (gamma + 1.0) * lamda * lamda), 1.0 / (gamma - 1.0));
    real T_1 = pin / (rho_1 * Rtilde);
    real dT = T_1 - T;
    if (fabs(dT) < eps) break;
    T = T_1;
  }
  const real rhob = pin / (Rtilde * T);
  const real vel = massflow / rhob;
  const real sc = -vel / dot3(flow_dir, n);
  for (int k = 0; k < 3; ++k) ub[k] = flow_dir[k] * sc + 0.0 + 0.0; /* += rigid_u, += us_n n (zero) */
  *Tb = T;
  *Pb = pin - Pambient;
}

/* massFlux of a set: mdot over the summed area of its fixed-mass faces
 * (src/bcFixedMass.loci:230-232: boundary_area of the boundary), summed in
 * face order */
static real set_mass_flux(const og_grid *g, const og_params *p, int set) {
  real area = 0;
  for (int f = 0; f < g->nf; ++f)
    if (OG_BT(g, f) == OG_FIXED_MASS && OG_SET(g, f) == set) area += g->area[f];
  return area > 0 ? p->bcs[set][7] / area : 0.0;
}

/* face state of an isentropic-inflow or fixed-mass face from a cell state
 * (T, u, pg); src/bcIsentropic.loci:146-166, src/bcFixedMass.loci:140-153 */
static void state_bc_(const og_params *p, int t, const double *bs, real mflux, real Ti, const real *ui, real pgi,
                      const real *n, real *Tb, real *ub, real *pgb) {
  const real d[3] = {-1. * n[0], -1. * n[1], -1. * n[2]}; /* flow_direction = -area.n */
  if (t == OG_ISENTROPIC_INFLOW) {
    const real a0 = sqrt(p->gamma * p->Rt * bs[6]);
    const real ai = sqrt(p->gamma * p->Rt * Ti);
    real pb;
    total_pressure_bc_(&pb, Tb, ub, bs[6], bs[5], p->gamma, a0, ui, ai, Ti, d);
    *pgb = pb - p->Pambient;
  } else {
    mass_flowrate_bc_(ub, Tb, pgb, pgi + p->Pambient, Ti, mflux, d, bs[6], p->Pambient, p->gamma, p->Rt, n);
  }
}

/* ------------------------------------------------------------------ */
/* boundary-face states                                                */
/* ------------------------------------------------------------------ */

/* farfield: src/bcInflow.loci:311-340; viscous wall: src/bcViscousWall.loci:205-214
 * (stationary wall, us = 0); symmetry: src/bcImpermeable.loci:48-64 */
void og_face_states(const og_grid *g, const og_params *p, const double *q, double *qf) {
  int nc = g->nc, nf = g->nf;
  real mflux[8];
  for (int k = 0; k < 8; ++k) mflux[k] = set_mass_flux(g, p, k);
  OG_PAR
  for (int f = 0; f < nf; ++f) {
    int t = OG_BT(g, f);
    if (t == OG_INTERIOR || t == OG_PERIODIC) continue;
    int c = g->cl[f];
    real n[3];
    fnormal(g, f, n);
    real Ti = q[c], ui[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]}, pgi = q[4 * nc + c];
    real Tb, pgb, ub[3];
    if (t == OG_FARFIELD) {
      const real us_n = 0.0, eta_pf = og_eta_face(p, Ti, ui);
      const real uit = dot3(ui, n) - us_n;
      const real ai = sqrt(p->gamma * p->Rt * Ti);
      const real a0 = sqrt(p->gamma * p->Rt * p->T0);
      if (uit > 0) {
        if (uit > ai) {
          pgb = pgi; Tb = Ti; ub[0] = ui[0]; ub[1] = ui[1]; ub[2] = ui[2];
        } else {
          og_cv_outflow(&pgb, &Tb, ub, pgi, Ti, ui, p->pg0, p->T0, p->u0, n, us_n,
                        p->Pambient, p->gamma, p->Rt, eta_pf);
        }
      } else {
        if (-uit > a0) {
          pgb = p->pg0; Tb = p->T0; ub[0] = p->u0[0]; ub[1] = p->u0[1]; ub[2] = p->u0[2];
        } else {
          og_cv_inflow(&pgb, &Tb, ub, pgi, Ti, ui, p->pg0, p->T0, p->u0, n, us_n,
                       p->Pambient, p->gamma, p->Rt, eta_pf);
        }
      }
    } else if (t == OG_WALL) {
      /* u_f = uwall - (uwall.n) n + us with uwall = 0, us = 0 */
      ub[0] = ub[1] = ub[2] = 0.0;
      pgb = pgi;
      Tb = p->Twall;
    } else if (t == OG_SYMMETRY) {
      const real us_n = 0.0;
      const real uit = dot3(ui, n) - us_n;
      for (int k = 0; k < 3; ++k) ub[k] = ui[k] - uit * n[k];
      const real rho = (pgi + p->Pambient) / (p->Rt * Ti);
      const real U = sqrt(dot3(ui, ui));
      pgb = pgi + 0.5 * rho * U * uit;
      Tb = Ti;
    } else if (t == OG_REFLECTING) { /* src/bcImpermeable.loci:102-117 */
      const real us_n = 0.0;
      const real uit = dot3(ui, n) - us_n;
      for (int k = 0; k < 3; ++k) ub[k] = ui[k] - uit * n[k];
      const real rho = (pgi + p->Pambient) / (p->Rt * Ti);
      const real Pmin = 0.5 * (pgi + p->Pambient) - p->Pambient;
      pgb = maxr(Pmin, pgi + 0.5 * rho * uit * fabs(uit));
      Tb = Ti;
    } else if (t == OG_INFLOW || t == OG_SUPERSONIC_INFLOW) { /* src/bcInflow.loci:270-278, 377-402 */
      const double *bs = p->bcs[OG_SET(g, f)];
      const real us_n = 0.0;
      const real u0[3] = {bs[2], bs[3], bs[4]};
      const real uit = dot3(ui, n) - us_n;
      const real a0 = sqrt(p->gamma * p->Rt * bs[0]);
      pgb = bs[1]; Tb = bs[0]; ub[0] = u0[0]; ub[1] = u0[1]; ub[2] = u0[2];
      if (t == OG_INFLOW && !(uit > 0) && !(-uit > a0))
        og_cv_inflow(&pgb, &Tb, ub, pgi, Ti, ui, bs[1], bs[0], u0, n, us_n, p->Pambient, p->gamma, p->Rt,
                     og_eta_face(p, Ti, ui));
    } else if (t == OG_OUTFLOW) { /* src/bcOutflow.loci:206-240, option p */
      const double *bs = p->bcs[OG_SET(g, f)];
      const real us_n = 0.0;
      const real pref = bs[1];
      const real uit = dot3(ui, n) - us_n;
      const real ai = sqrt(p->gamma * p->Rt * Ti);
      const real rhoi = (pgi + p->Pambient) / (p->Rt * Ti);
      if (uit >= 0) {
        if (uit > ai) {
          for (int k = 0; k < 3; ++k) ub[k] = ui[k];
          Tb = Ti;
          pgb = pgi;
        } else {
          pgb = pref;
          for (int k = 0; k < 3; ++k) ub[k] = ui[k];
          const real pa = pgb + p->Pambient;
          Tb = pa / (p->Rt * rhoi);
        }
      } else {
        for (int k = 0; k < 3; ++k) ub[k] = us_n * n[k];
        pgb = pgi - .05 * (pgi + p->Pambient);
        Tb = Ti;
      }
    } else if (t == OG_SUPERSONIC_OUTFLOW) { /* src/bcOutflow.loci:127-141 */
      pgb = pgi;
      const real ai = sqrt(p->gamma * p->Rt * Ti);
      const real ut = dot3(n, ui) - 0.0;
      if (ut < ai) pgb -= (pgi + p->Pambient) * .1;
      Tb = Ti;
      for (int k = 0; k < 3; ++k) ub[k] = ui[k];
    } else if (t == OG_EXTRAPOLATE) { /* src/bcOutflow.loci:84-94 */
      pgb = pgi; Tb = Ti;
      for (int k = 0; k < 3; ++k) ub[k] = ui[k];
    } else { /* isentropic inflow, fixed mass */
      state_bc_(p, t, p->bcs[OG_SET(g, f)], mflux[OG_SET(g, f)], Ti, ui, pgi, n, &Tb, ub, &pgb);
    }
    if (p->turb == 2) {
      /* k_f, w_f: inflow/outflow switch on u_f.n (turbulence/KOmega/sst_viscous.loci:131-144),
       * Menter's wall values k = 0, omega = 60 mu/(rho_f beta1 d^2) with the
       * cell's mu and wall distance (:236-243), symmetry copies the cell
       * (src/solverScalarTransport.loci:423-427) */
      const real ki = q[5 * nc + c], wi = q[6 * nc + c];
      real kb = ki, wb = wi;
      if (t == OG_WALL) {
        const real rho_f = (pgb + p->Pambient) / (p->Rt * Tb);
        const real d = g->dist[c];
        kb = 0.0;
        wb = 10.0 * 6.0 * suth_mu_(p, Ti) / (rho_f * 0.075 * pow(d, 2.0));
      } else if (t == OG_FARFIELD) {
        kb = p->k_bc;
        wb = p->w_bc;
        if (dot3(ub, n) >= -0.0) { kb = ki; wb = wi; }
      }
      qf[5 * nf + f] = kb;
      qf[6 * nf + f] = wb;
    } else if (p->turb) { /* nuTilde_f: SA_turb.loci:181-183 (wall), solverScalarTransport.loci:423-463 */
      const real Xi = q[5 * nc + c];
      real Xb = Xi;
      if (t == OG_WALL) Xb = 0.0;
      else if (t == OG_FARFIELD) Xb = (dot3(n, ub) - 0.0 >= -0.0) ? Xi : p->nut_bc;
      qf[5 * nf + f] = Xb;
    }
    qf[f] = Tb;
    qf[nf + f] = ub[0];
    qf[2 * nf + f] = ub[1];
    qf[3 * nf + f] = ub[2];
    qf[4 * nf + f] = pgb;
  }
}

/* ------------------------------------------------------------------ */
/* gradients and limiters (Loci FVM pieces: textbook definitions)      */
/* ------------------------------------------------------------------ */

/* Inverse of the symmetric 3x3 least-squares normal matrix
 * A = sum dx dx^T over the face stencil, by the adjugate. */
static void lsq_inverse(const og_grid *g, const og_adj *a, int c, real *Ai /* 6: xx xy xz yy yz zz */) {
  real A[6] = {0, 0, 0, 0, 0, 0};
  real xc[3];
  ccenter(g, c, xc);
  int lists[3][2] = {{a->lo_off[c], a->lo_off[c + 1]}, {a->up_off[c], a->up_off[c + 1]}, {a->bd_off[c], a->bd_off[c + 1]}};
  const int *arr[3] = {a->lo, a->up, a->bd};
  for (int L = 0; L < 3; ++L)
    for (int j = lists[L][0]; j < lists[L][1]; ++j) {
      int f = arr[L][j];
      real x[3];
      int t = OG_BT(g, f);
      if (t == OG_INTERIOR) {
        int o = (g->cl[f] == c) ? g->cr[f] : g->cl[f];
        ccenter(g, o, x);
      } else if (t == OG_PERIODIC) {
        rcell(g, f, x);
      } else {
        fcenter(g, f, x);
      }
      real dx = x[0] - xc[0], dy = x[1] - xc[1], dz = x[2] - xc[2];
      A[0] += dx * dx; A[1] += dx * dy; A[2] += dx * dz;
      A[3] += dy * dy; A[4] += dy * dz; A[5] += dz * dz;
    }
  real c00 = A[3] * A[5] - A[4] * A[4];
  real c01 = A[2] * A[4] - A[1] * A[5];
  real c02 = A[1] * A[4] - A[2] * A[3];
  real c11 = A[0] * A[5] - A[2] * A[2];
  real c12 = A[1] * A[2] - A[0] * A[4];
  real c22 = A[0] * A[3] - A[1] * A[1];
  real det = A[0] * c00 + A[1] * c01 + A[2] * c02;
  real rdet = 1.0 / det;
  Ai[0] = c00 * rdet; Ai[1] = c01 * rdet; Ai[2] = c02 * rdet;
  Ai[3] = c11 * rdet; Ai[4] = c12 * rdet; Ai[5] = c22 * rdet;
}

/* Venkatakrishnan limiter function (Venkatakrishnan, AIAA 93-0880) with the
 * threshold eps2 = (Kl h)^3, h = vol^(1/3). */
static inline real venkat(real delp, real delm, real eps2) {
  const real num = ((delp * delp + eps2) * delm + 2.0 * delm * delm * delp);
  const real den = (delm * (delp * delp + 2.0 * delm * delm + delm * delp + eps2));
  const real e = (den >= 0.0 ? 1.0e-30 : -1.0e-30);
  return (num + e) / (den + e);
}

void og_gradients(const og_grid *g, const og_params *p, const double *q, const double *qf,
                  double *grad, double *phi) {
  int nc = g->nc;
  og_adj a;
  build_adj(g, &a);
  OG_PAR
  for (int c = 0; c < nc; ++c) {
    real Ai[6];
    lsq_inverse(g, &a, c, Ai);
    real xc[3];
    ccenter(g, c, xc);
    const int nq = nq_of(p);
    real qc[7];
    for (int k = 0; k < nq; ++k) qc[k] = q[k * nc + c];
    real b[7][3];
    real qmin[7], qmax[7];
    for (int k = 0; k < nq; ++k) { b[k][0] = b[k][1] = b[k][2] = 0; qmin[k] = qmax[k] = qc[k]; }
    int lists[3][2] = {{a.lo_off[c], a.lo_off[c + 1]}, {a.up_off[c], a.up_off[c + 1]}, {a.bd_off[c], a.bd_off[c + 1]}};
    const int *arr[3] = {a.lo, a.up, a.bd};
    for (int L = 0; L < 3; ++L)
      for (int j = lists[L][0]; j < lists[L][1]; ++j) {
        int f = arr[L][j];
        real x[3], v[7];
        stencil_nb(g, q, qf, c, f, nq, x, v);
        real dx = x[0] - xc[0], dy = x[1] - xc[1], dz = x[2] - xc[2];
        for (int k = 0; k < nq; ++k) {
          real dq = v[k] - qc[k];
          b[k][0] += dx * dq; b[k][1] += dy * dq; b[k][2] += dz * dq;
          qmin[k] = minr(qmin[k], v[k]);
          qmax[k] = maxr(qmax[k], v[k]);
        }
      }
    real gr[7][3];
    for (int k = 0; k < nq; ++k) {
      gr[k][0] = Ai[0] * b[k][0] + Ai[1] * b[k][1] + Ai[2] * b[k][2];
      gr[k][1] = Ai[1] * b[k][0] + Ai[3] * b[k][1] + Ai[4] * b[k][2];
      gr[k][2] = Ai[2] * b[k][0] + Ai[4] * b[k][1] + Ai[5] * b[k][2];
      for (int d = 0; d < 3; ++d) grad[(3 * k + d) * nc + c] = gr[k][d];
    }
    /* limiter */
    real ph[7];
    for (int k = 0; k < nq; ++k) ph[k] = 1.0;
    if (p->limiter == 2) {
      for (int k = 0; k < nq; ++k) ph[k] = 0.0;
    } else if (p->limiter == 1 || p->limiter == 3) {
      const real h3 = g->vol[c];
      const real eps2 = p->Kl * p->Kl * p->Kl * h3;
      for (int L = 0; L < 3; ++L)
        for (int j = lists[L][0]; j < lists[L][1]; ++j) {
          int f = arr[L][j];
          real xf[3];
          fcenter(g, f, xf);
          real dx = xf[0] - xc[0], dy = xf[1] - xc[1], dz = xf[2] - xc[2];
          for (int k = 0; k < nq; ++k) {
            real qdif = gr[k][0] * dx + gr[k][1] * dy + gr[k][2] * dz;
            real lim;
            if (p->limiter == 1) {
              real delp = (qdif >= 0.0) ? qmax[k] - qc[k] : qmin[k] - qc[k];
              lim = venkat(delp, qdif, eps2);
            } else {
              if (qdif > 0) lim = minr(1.0, (qmax[k] - qc[k]) / qdif);
              else if (qdif < 0) lim = minr(1.0, (qmin[k] - qc[k]) / qdif);
              else lim = 1.0;
            }
            ph[k] = minr(ph[k], lim);
          }
        }
      for (int k = 0; k < nq; ++k) ph[k] = maxr(0.0, minr(1.0, ph[k]));
    }
    for (int k = 0; k < nq; ++k) phi[k * nc + c] = ph[k];
  }
  free_adj(&a);
}

/* Limited reconstruction q_f = q_c + phi_c (grad q_c . (x_f - x_c)), with the
 * floors of leftsP(temperature,Zero) and leftsP(gagePressure,minPg). */
static void recon(const og_grid *g, const og_params *p, const double *q, const double *grad,
                  const double *phi, int c, const real *xcell, const real *xf, real *out) {
  int nc = g->nc;
  real d[3] = {xf[0] - xcell[0], xf[1] - xcell[1], xf[2] - xcell[2]};
  for (int k = 0; k < 5; ++k) {
    real gd = grad[(3 * k) * nc + c] * d[0] + grad[(3 * k + 1) * nc + c] * d[1] +
              grad[(3 * k + 2) * nc + c] * d[2];
    out[k] = q[k * nc + c] + phi[k * nc + c] * gd;
  }
  out[0] = maxr(out[0], 0.0);
  out[4] = maxr(out[4], p->minPg);
}

/* Sutherland law: src/fluxViscous.loci:112-119 (mu), :137-143 (kconduct) */
static inline real suth_mu(const og_params *p, real T) { return p->suth[0] * pow(T, p->suth[1]) / (T + p->suth[2]); }
static inline real suth_k(const og_params *p, real T) { return p->suth[3] * pow(T, p->suth[4]) / (T + p->suth[5]); }

/* Face gradient of variable k (gradv3d_f / grads_f, Loci FVM):
 * interior/periodic: average of the two cell gradients corrected along the
 * cell-centre line; boundary: cell gradient corrected along x_f - x_c. */
static void face_grad(const og_grid *g, const double *q, const double *qf, const double *grad,
                      int f, int k, real *gout) {
  int nc = g->nc, nf = g->nf;
  int t = OG_BT(g, f);
  int c = g->cl[f];
  real xl[3];
  ccenter(g, c, xl);
  if (t == OG_INTERIOR || t == OG_PERIODIC) {
    real xr[3];
    int o = rcell(g, f, xr);
    real d[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
    real ga[3];
    for (int j = 0; j < 3; ++j) ga[j] = 0.5 * (grad[(3 * k + j) * nc + c] + grad[(3 * k + j) * nc + o]);
    real corr = (q[k * nc + o] - q[k * nc + c] - dot3(ga, d)) / dot3(d, d);
    for (int j = 0; j < 3; ++j) gout[j] = ga[j] + corr * d[j];
  } else {
    real xf[3];
    fcenter(g, f, xf);
    real d[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
    real ga[3];
    for (int j = 0; j < 3; ++j) ga[j] = grad[(3 * k + j) * nc + c];
    real corr = (qf[k * nf + f] - q[k * nc + c] - dot3(ga, d)) / dot3(d, d);
    for (int j = 0; j < 3; ++j) gout[j] = ga[j] + corr * d[j];
  }
}

/* Spalart-Allmaras eddy viscosity of a cell: xcoeff, fv1, tmu
 * (turbulence/SA/SA_turb.loci:93-107). */
static real sa_fv1(const og_params *p, real x) {
  return pow(x, 3.0) / (pow(x, 3.0) + pow(p->sa_cv1, 3.0));
}
static real sst_tmu_cell(const og_grid *g, const og_params *p, const double *q, const double *grad, int c);
static real tmu_cell(const og_grid *g, const og_params *p, const double *q, const double *grad, int c) {
  if (p->turb == 2) return sst_tmu_cell(g, p, q, grad, c);
  const int nc = g->nc;
  const real T = q[c], pg = q[4 * nc + c], nut = q[5 * nc + c];
  const real rho = (pg + p->Pambient) / (p->Rt * T);
  const real mu = suth_mu(p, T);
  const real x = rho * nut / mu;
  return nut * sa_fv1(p, x) * rho;
}
/* tmu_f: volume-weighted interior average, the interior cell on boundary
 * faces, zero on no-slip walls (src/fluxViscous.loci:192-201) */
static real tmu_face(const og_grid *g, const og_params *p, const double *q, const double *grad, int f) {
  const int t = OG_BT(g, f);
  const int c = g->cl[f];
  if (t == OG_INTERIOR || t == OG_PERIODIC) {
    real xr[3];
    const int o = rcell(g, f, xr);
    const real vl = g->vol[c], vr = g->vol[o];
    return (tmu_cell(g, p, q, grad, o) * vl + tmu_cell(g, p, q, grad, c) * vr) / (vr + vl);
  }
  if (t == OG_WALL) return 0.0;
  return tmu_cell(g, p, q, grad, c);
}

/* Face viscous quantities: muTotal_f, kconductTotal, u_f and vflux[4]
 * (src/fluxViscous.loci:121-133,145-151,184-197,217-238,272-292). */
static void viscous_face(const og_grid *g, const og_params *p, const double *q, const double *qf,
                         const double *grad, int f, real *mu_f, real *k_f, real *uf, real *vflux) {
  int nc = g->nc, nf = g->nf;
  int t = OG_BT(g, f);
  real an[3];
  fnormal(g, f, an);
  const real sada = g->area[f];
  if (t == OG_INTERIOR || t == OG_PERIODIC) {
    int c = g->cl[f];
    real xr[3];
    int o = rcell(g, f, xr);
    const real vl = g->vol[c], vr = g->vol[o];
    real mul = suth_mu(p, q[c]), mur = suth_mu(p, q[o]);
    real kl = suth_k(p, q[c]), kr = suth_k(p, q[o]);
    *mu_f = (vl * mur + vr * mul) / (vl + vr);
    *k_f = (vl * kr + vr * kl) / (vl + vr);
    const real factor = 1. / (vl + vr);
    for (int j = 0; j < 3; ++j) uf[j] = factor * (vl * q[(1 + j) * nc + o] + vr * q[(1 + j) * nc + c]);
  } else {
    real Tf = qf[f];
    *mu_f = suth_mu(p, Tf);
    *k_f = suth_k(p, Tf);
    for (int j = 0; j < 3; ++j) uf[j] = qf[(1 + j) * nf + f];
  }
  if (t == OG_SYMMETRY) { /* src/fluxViscous.loci:294-297 */
    if (p->turb) {
      const real tmf = tmu_face(g, p, q, grad, f);
      *k_f = *k_f + tmf * (p->Rt * p->gamma / (p->gamma - 1.0)) / p->Prt;
      *mu_f = *mu_f + tmf;
    }
    vflux[0] = vflux[1] = vflux[2] = vflux[3] = 0;
    return;
  }
  real gu[3][3], gT[3];
  for (int i = 0; i < 3; ++i) face_grad(g, q, qf, grad, f, 1 + i, gu[i]);
  face_grad(g, q, qf, grad, f, 0, gT);
  real mu_total = *mu_f;
  if (p->turb) { /* muTotal_f, kconductTotal: src/fluxViscous.loci:182-200,275-278 */
    const real tmf = tmu_face(g, p, q, grad, f);
    mu_total = *mu_f + tmf;
    *k_f = *k_f + tmf * (p->Rt * p->gamma / (p->gamma - 1.0)) / p->Prt;
    *mu_f = mu_total;
  }
  real dudx = gu[0][0], dvdy = gu[1][1], dwdz = gu[2][2];
  real dvdx = gu[1][0], dwdx = gu[2][0], dwdy = gu[2][1];
  real dudy = gu[0][1], dudz = gu[0][2], dvdz = gu[1][2];
  real divm = (dudx + dvdy + dwdz) * (1. / 3.);
  real txx = 2.0 * mu_total * (dudx - divm);
  real tyy = 2.0 * mu_total * (dvdy - divm);
  real tzz = 2.0 * mu_total * (dwdz - divm);
  real txy = mu_total * (dudy + dvdx);
  real txz = mu_total * (dudz + dwdx);
  real tyz = mu_total * (dvdz + dwdy);
  vflux[0] = (txx * an[0] + txy * an[1] + txz * an[2]) * sada;
  vflux[1] = (txy * an[0] + tyy * an[1] + tyz * an[2]) * sada;
  vflux[2] = (txz * an[0] + tyz * an[1] + tzz * an[2]) * sada;
  vflux[3] = (uf[0] * vflux[0] + uf[1] * vflux[1] + uf[2] * vflux[2]);
  const real ktotal = *k_f;
  vflux[3] += dot3(gT, an) * ktotal * sada;
}

/* Face fluxes: interior/periodic src/fluxInviscid.loci:1170-1188; farfield
 * src/bcInflow.loci:342-355; wall src/bcViscousWall.loci:216-229; symmetry
 * src/bcImpermeable.loci:66-73; maxev_f src/solverSetup.loci:83-105. */
void og_fluxes(const og_grid *g, const og_params *p, const double *q, const double *qf,
               const double *grad, const double *phi, double *iflux, double *vflux,
               double *maxev) {
  int nc = g->nc, nf = g->nf;
  const real us_n = 0.0;
  real *eta_p = malloc(sizeof(real) * nc);
  eta_impl(g, p, q, grad, eta_p);
  OG_PAR
  for (int f = 0; f < nf; ++f) {
    int t = OG_BT(g, f);
    real an[3], xf[3], xl[3];
    fnormal(g, f, an);
    fcenter(g, f, xf);
    int c = g->cl[f];
    ccenter(g, c, xl);
    /* interior and periodic: max of the two cells (:1176-1187, the periodic
     * ghost holds its partner's Eta_p); boundary faces: the interior cell */
    real eta = eta_p[c];
    if (t == OG_INTERIOR) eta = maxr(eta_p[c], eta_p[g->cr[f]]);
    else if (t == OG_PERIODIC) eta = maxr(eta_p[c], eta_p[g->cl[g->pface[f]]]);
    real fl[5];
    real qL[5];
    recon(g, p, q, grad, phi, c, xl, xf, qL);
    real UL[3] = {qL[1], qL[2], qL[3]};
    const int riemann_face = t == OG_FARFIELD || (t >= OG_INFLOW && t <= OG_EXTRAPOLATE);
    if (t == OG_INTERIOR || t == OG_PERIODIC || riemann_face) {
      real qR[5];
      if (riemann_face) {
        for (int k = 0; k < 5; ++k) qR[k] = qf[k * nf + f];
        if (t == OG_SUPERSONIC_INFLOW) { /* Eta_pf, src/bcInflow.loci:280-289 */
          const real Uc[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
          eta = og_eta_face(p, q[c], Uc);
        }
      } else {
        real xr[3];
        int o = rcell(g, f, xr);
        recon(g, p, q, grad, phi, o, xr, xf, qR);
      }
      real UR[3] = {qR[1], qR[2], qR[3]};
      if (p->flux == 1)
        og_roe_flux(fl, qL[4], qL[0], UL, qR[4], qR[0], UR, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n, eta);
      else
        og_hllc_flux(fl, qL[4], qL[0], UL, qR[4], qR[0], UR, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n, eta);
    } else if (t == OG_REFLECTING) { /* src/bcImpermeable.loci:119-131 */
      reflecting_flux_(p, fl, qL[4], qL[0], UL, an, g->area[f], us_n, eta);
      fl[0] = 0;
    } else if (t == OG_ISENTROPIC_INFLOW || t == OG_FIXED_MASS) { /* the face state's Euler flux */
      const real uf[3] = {qf[nf + f], qf[2 * nf + f], qf[3 * nf + f]};
      euler_flux_(fl, qf[4 * nf + f], qf[f], uf, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n);
    } else if (t == OG_WALL) {
      const real pg = qf[4 * nf + f];
      const real A = g->area[f];
      fl[0] = 0;
      fl[1] = pg * (an[0] * A);
      fl[2] = pg * (an[1] * A);
      fl[3] = pg * (an[2] * A);
      const real pw = pg + p->Pambient;
      fl[4] = pw * A * us_n;
    } else { /* symmetry */
      const real pg = qf[4 * nf + f];
      real uf[3] = {qf[nf + f], qf[2 * nf + f], qf[3 * nf + f]};
      fl[0] = 0;
      const real moment = pg * g->area[f];
      fl[1] = an[0] * moment;
      fl[2] = an[1] * moment;
      fl[3] = an[2] * moment;
      fl[4] = g->area[f] * pg * dot3(uf, an);
    }
    for (int k = 0; k < 5; ++k) iflux[k * nf + f] = fl[k];
    /* viscous */
    real vf[4] = {0, 0, 0, 0};
    if (p->viscous) {
      real mu_f, k_f, uf[3];
      viscous_face(g, p, q, qf, grad, f, &mu_f, &k_f, uf, vf);
    }
    for (int k = 0; k < 4; ++k) vflux[k * nf + f] = vf[k];
    /* maxev_f */
    const real Eta_P = eta + 1., Eta_M = eta - 1.;
    if (t == OG_INTERIOR || t == OG_PERIODIC) {
      real xr[3];
      int o = rcell(g, f, xr);
      real ul[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
      real ur[3] = {q[nc + o], q[2 * nc + o], q[3 * nc + o]};
      const real utl = dot3(ul, an) - us_n, utr = dot3(ur, an) - us_n;
      const real al = sqrt(p->gamma * p->Rt * q[c]), ar = sqrt(p->gamma * p->Rt * q[o]);
      const real sigmal = sqrt(utl * utl * Eta_M * Eta_M + 4. * eta * al * al);
      const real sigmar = sqrt(utr * utr * Eta_M * Eta_M + 4. * eta * ar * ar);
      const real mevl = 0.5 * maxr(fabs(utl * Eta_P - sigmal), fabs(utl * Eta_P + sigmal));
      const real mevr = 0.5 * maxr(fabs(utr * Eta_P - sigmar), fabs(utr * Eta_P + sigmar));
      maxev[f] = maxr(mevl, mevr);
    } else {
      real uf[3] = {qf[nf + f], qf[2 * nf + f], qf[3 * nf + f]};
      const real utf = dot3(uf, an) - us_n;
      const real af = sqrt(p->gamma * p->Rt * qf[f]);
      const real sigmaf = sqrt(utf * utf * Eta_M * Eta_M + 4. * eta * af * af);
      maxev[f] = 0.5 * maxr(fabs(utf * Eta_P - sigmaf), fabs(utf * Eta_P + sigmaf));
    }
  }
  free(eta_p);
}

/* ------------------------------------------------------------------ */
/* residual gather and norms                                           */
/* ------------------------------------------------------------------ */

/* src join: src/fluxInviscid.loci:1253-1268, src/fluxViscous.loci:299-310.
 * Per cell the faces join in the order lower, upper, boundary (each
 * ascending), each face adding its inviscid then its viscous part. */
static void gather_src(const og_grid *g, const og_params *p, const og_adj *a, const double *iflux,
                       const double *vflux, double *src) {
  int nc = g->nc, nf = g->nf;
  OG_PAR
  for (int c = 0; c < nc; ++c) {
    real s[5] = {0, 0, 0, 0, 0};
    for (int j = a->lo_off[c]; j < a->lo_off[c + 1]; ++j) {
      int f = a->lo[j];
      for (int k = 0; k < 5; ++k) s[k] += iflux[k * nf + f];
      if (p->viscous)
        for (int k = 0; k < 4; ++k) s[k + 1] -= vflux[k * nf + f];
    }
    for (int j = a->up_off[c]; j < a->up_off[c + 1]; ++j) {
      int f = a->up[j];
      for (int k = 0; k < 5; ++k) s[k] -= iflux[k * nf + f];
      if (p->viscous)
        for (int k = 0; k < 4; ++k) s[k + 1] += vflux[k * nf + f];
    }
    for (int j = a->bd_off[c]; j < a->bd_off[c + 1]; ++j) {
      int f = a->bd[j];
      for (int k = 0; k < 5; ++k) s[k] -= iflux[k * nf + f];
      if (p->viscous)
        for (int k = 0; k < 4; ++k) s[k + 1] += vflux[k * nf + f];
    }
    for (int k = 0; k < 5; ++k) src[k * nc + c] = s[k];
  }
}

/* src/solverResidual.loci:31-112 (steady residual, resid_join), in cell order */
static void residual_norms(const og_grid *g, const double *src, og_norms *r) {
  int nc = g->nc;
  real vtot = 0;
  for (int c = 0; c < nc; ++c) vtot += g->vol[c];
  memset(r, 0, sizeof(*r));
  for (int c = 0; c < nc; ++c) {
    const int mi = 1, ei = 4;
    real rvol = 1. / g->vol[c];
    real weight = g->vol[c] / vtot;
    real s0 = src[c];
    real rtmax = 0, cr3[3] = {0, 0, 0};
    if (fabs(rtmax) < fabs(rvol * s0)) { rtmax = rvol * s0; ccenter(g, c, cr3); }
    real rtrms = 0;
    rtrms += (rvol * s0) * (rvol * s0);
    rtrms *= weight;
    real mtrms = 0, mtmax = 0;
    for (int i = 0; i < 3; ++i) {
      real v = rvol * src[(mi + i) * nc + c];
      if (fabs(mtmax) < fabs(v)) mtmax = v;
      mtrms += v * v;
    }
    mtrms *= weight / 3.0;
    real ve = rvol * src[ei * nc + c];
    real etrms = weight * ve * ve;
    real etmax = ve;
    r->rtrms += rtrms; r->etrms += etrms; r->mtrms += mtrms;
    if (fabs(r->etmax) < fabs(etmax)) { r->etmax = etmax; ccenter(g, c, r->ccmaxe); }
    if (fabs(r->rtmax) < fabs(rtmax)) { r->rtmax = rtmax; memcpy(r->ccmaxr, cr3, sizeof(cr3)); }
    if (fabs(r->mtmax) < fabs(mtmax)) r->mtmax = mtmax;
  }
}

void og_residual(const og_grid *g, const og_params *p, const double *q, double *src,
                 double *iflux, double *vflux, og_norms *nrm) {
  int nc = g->nc, nf = g->nf;
  double *qf = calloc((size_t)7 * nf, sizeof(double));
  double *grad = malloc(sizeof(double) * 21 * nc), *phi = malloc(sizeof(double) * 7 * nc);
  double *maxev = malloc(sizeof(double) * nf);
  og_face_states(g, p, q, qf);
  og_gradients(g, p, q, qf, grad, phi);
  og_fluxes(g, p, q, qf, grad, phi, iflux, vflux, maxev);
  og_adj a;
  build_adj(g, &a);
  gather_src(g, p, &a, iflux, vflux, src);
  if (nrm) residual_norms(g, src, nrm);
  free_adj(&a);
  free(qf); free(grad); free(phi); free(maxev);
}

/* ------------------------------------------------------------------ */
/* block LU (Loci Tools Mat; textbook Doolittle with partial pivoting)  */
/* ------------------------------------------------------------------ */

/* In place: unit-lower L below the diagonal, U on and above it; rows are
 * swapped, pivot[i] = original row now at position i.  Returns 0 on success. */
static int lu_decompose(real_fj *a, int *pivot, int nb) {
  for (int i = 0; i < nb; ++i) pivot[i] = i;
  for (int k = 0; k < nb; ++k) {
    int pr = k;
    real_fj amax = (real_fj)fabs((double)a[k + nb * k]);
    for (int i = k + 1; i < nb; ++i) {
      real_fj v = (real_fj)fabs((double)a[i + nb * k]);
      if (v > amax) { amax = v; pr = i; }
    }
    if (pr != k) {
      for (int j = 0; j < nb; ++j) {
        real_fj t = a[k + nb * j]; a[k + nb * j] = a[pr + nb * j]; a[pr + nb * j] = t;
      }
      int t = pivot[k]; pivot[k] = pivot[pr]; pivot[pr] = t;
    }
    if (a[k + nb * k] == 0) return -1;
    for (int i = k + 1; i < nb; ++i) {
      real_fj l = a[i + nb * k] / a[k + nb * k];
      a[i + nb * k] = l;
      for (int j = k + 1; j < nb; ++j) a[i + nb * j] = a[i + nb * j] - l * a[k + nb * j];
    }
  }
  return 0;
}

/* forward/back substitution in real (double) with real_fj factors
 * (the order of src/linearSolverFSGS.loci:73-106, with the row permutation) */
static void lu_solve(const real_fj *a, const int *pivot, const real *b, real *x, int nb) {
  for (int i = 0; i < nb; ++i) {
    real tmp = b[pivot[i]];
    for (int j = 0; j < i; ++j) tmp -= a[i + nb * j] * x[j];
    x[i] = tmp;
  }
  for (int i = nb - 1; i >= 0; --i) {
    real tmp = x[i];
    for (int j = i + 1; j < nb; ++j) tmp -= a[i + nb * j] * x[j];
    x[i] = tmp / a[i + nb * i];
  }
}

int og_lu_decompose(void *a25, int *pivot, int nb) { return lu_decompose((real_fj *)a25, pivot, nb); }
void og_lu_solve(const void *a25, const int *pivot, const double *b, double *x, int nb) {
  lu_solve((const real_fj *)a25, pivot, b, x, nb);
}

/* t[i] += sum_j M(i,j) v[j], columns outer (src/linearSolverFSGS.loci:53-70) */
static inline void dotprod_accum(const real_fj *M, const real *v, real *t) {
  for (int j = 0; j < NB; ++j) {
    const real in = v[j];
    for (int i = 0; i < NB; ++i) t[i] += M[i + NB * j] * in;
  }
}

/* ------------------------------------------------------------------ */
/* Jacobians and assembly                                              */
/* ------------------------------------------------------------------ */

/* thin-layer viscous Jacobian terms (src/fluxViscous.loci:315-403), added
 * with sign s (+1 for fjp, -1 for fjm) */
static void visc_jac(real_fj *F, real s, const real *an, real sada, real dist, real mu_f, real k_f,
                     const real *uf, const real *vf) {
  const int mi = 1, ei = 4, ru = mi, rv = mi + 1, rw = mi + 2;
  const real coef = mu_f * sada / dist;
  const real third = 1.0 / 3.0;
  real c_[3][3];
  c_[0][0] = coef * (1.0 + an[0] * an[0] * third);
  c_[0][1] = coef * an[0] * an[1] * third;
  c_[0][2] = coef * an[0] * an[2] * third;
  c_[1][0] = coef * an[1] * an[0] * third;
  c_[1][1] = coef * (1.0 + an[1] * an[1] * third);
  c_[1][2] = coef * an[1] * an[2] * third;
  c_[2][0] = coef * an[2] * an[0] * third;
  c_[2][1] = coef * an[2] * an[1] * third;
  c_[2][2] = coef * (1.0 + an[2] * an[2] * third);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) addfj(&M_(F, mi + i, mi + j), s * c_[i][j]);
  /* energy row, literal operand order of the rule bodies:
   * diagonal u_i*coef*(1 + n_i n_i/3), off-diagonal u_i*coef*n_i*n_j/3 */
  const int cols[3] = {ru, rv, rw};
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      real term = (i == j) ? uf[i] * coef * (1.0 + an[i] * an[i] * third)
                           : uf[i] * coef * an[i] * an[j] * third;
      addfj(&M_(F, ei, cols[j]), s * term);
    }
  addfj(&M_(F, ei, ru), s * (.5 * vf[0]));
  addfj(&M_(F, ei, rv), s * (.5 * vf[1]));
  addfj(&M_(F, ei, rw), s * (.5 * vf[2]));
  const real factor = k_f * sada / dist;
  addfj(&M_(F, ei, 0), s * factor);
}

/* Temporal-damping Jacobian damping_fjp (s = +1, left face state) /
 * damping_fjm (s = -1, right face state), src/fluxInviscid.loci:794-866:
 * beta (|u.n - us_n| + a) A/2 times dQ/dq of the face state, with the
 * intermediates held in real_fj as the rule body declares them. */
static void damping_jac(real_fj *fj, real s, real pg, real T, const real *U, const real *an, real area,
                        real pamb, real Rt, real gamma, real us_n, real beta) {
  const real_fj coef = (real_fj)(0.5 * area);
  const real_fj gm1 = (real_fj)(gamma - 1);
  const real_fj rr = (real_fj)((pg + pamb) / (Rt * T));
  const real_fj drdp = (real_fj)(1. / (Rt * T));
  const real_fj drdT = (real_fj)(-rr / T);
  const real_fj uta = (real_fj)dot3(U, an);
  const real_fj ut = (real_fj)(uta - us_n);
  const real_fj ar = (real_fj)sqrt(gamma * Rt * T);
  const real_fj aut = (real_fj)fabs(ut);
  const real coefd = beta * (real_fj)(aut + ar) * coef;
  const real uu = dot3(U, U);
  addfj(&M_(fj, 0, 0), s * (coefd * drdT));
  addfj(&M_(fj, 1, 0), s * (coefd * drdT * U[0]));
  addfj(&M_(fj, 2, 0), s * (coefd * drdT * U[1]));
  addfj(&M_(fj, 3, 0), s * (coefd * drdT * U[2]));
  addfj(&M_(fj, 4, 0), s * (coefd * drdT * 0.5 * uu));
  addfj(&M_(fj, 0, 4), s * (coefd * drdp));
  addfj(&M_(fj, 1, 4), s * (coefd * drdp * U[0]));
  addfj(&M_(fj, 2, 4), s * (coefd * drdp * U[1]));
  addfj(&M_(fj, 3, 4), s * (coefd * drdp * U[2]));
  addfj(&M_(fj, 4, 4), s * (coefd * (1. / gm1 + drdp * 0.5 * uu)));
  addfj(&M_(fj, 1, 1), s * (coefd * rr));
  addfj(&M_(fj, 2, 2), s * (coefd * rr));
  addfj(&M_(fj, 3, 3), s * (coefd * rr));
  addfj(&M_(fj, 4, 1), s * (coefd * rr * U[0]));
  addfj(&M_(fj, 4, 2), s * (coefd * rr * U[1]));
  addfj(&M_(fj, 4, 3), s * (coefd * rr * U[2]));
}

/* Boundary-face Jacobian block bj (real_fj, zeroed by the caller), i.e. the
 * face's whole contribution to srcJ of its interior cell:
 *   farfield src/bcInflow.loci:357-375, wall src/bcViscousWall.loci:230-240,
 *   symmetry src/bcImpermeable.loci:75-100, viscous src/fluxViscous.loci:405-448. */
static void boundary_block(const og_grid *g, const og_params *p, const double *q, const double *qf,
                           const double *grad, const double *phi, int f, real eta_c, real_fj *bj) {
  int nc = g->nc, nf = g->nf;
  int t = OG_BT(g, f);
  int c = g->cl[f];
  real an[3], xf[3], xl[3];
  fnormal(g, f, an);
  fcenter(g, f, xf);
  ccenter(g, c, xl);
  const real sada = g->area[f], us_n = 0.0;
  if (t == OG_FARFIELD || (t >= OG_INFLOW && t <= OG_EXTRAPOLATE)) {
    /* srcJ -= fjp(left, face state): farfield; inflow, supersonic inflow
     * (src/bcInflow.loci:291-308,415-430), outflow, supersonic outflow,
     * extrapolate (src/bcOutflow.loci:107-125,154-171,254-274) */
    real qL[5];
    recon(g, p, q, grad, phi, c, xl, xf, qL);
    real UL[3] = {qL[1], qL[2], qL[3]};
    real UR[3] = {qf[nf + f], qf[2 * nf + f], qf[3 * nf + f]};
    real eta = eta_c;
    if (t == OG_SUPERSONIC_INFLOW) {
      const real Uc[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
      eta = og_eta_face(p, q[c], Uc);
    }
    real_fj fjp[25];
    for (int i = 0; i < 25; ++i) fjp[i] = 0;
    roe_jacobian(fjp, qL[4], qL[0], UL, qf[4 * nf + f], qf[f], UR, an, sada, p->Pambient, p->Rt,
                 p->gamma, us_n, eta, 1);
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) M_(bj, i, j) -= M_(fjp, i, j);
  } else if (t == OG_REFLECTING) {
    /* src/bcImpermeable.loci:132-181 (NUMERICAL): one-sided differences
     * of the flux in the reconstructed left state, rows 1-4 */
    real qL[5];
    recon(g, p, q, grad, phi, c, xl, xf, qL);
    real Tl = qL[0], Pgl = qL[4], ul[3] = {qL[1], qL[2], qL[3]};
    real f0[5], f1[5], vals[5], deltas[5];
    reflecting_flux_(p, f0, Pgl, Tl, ul, an, sada, us_n, eta_c);
    const real a2 = p->gamma * p->Rt * Tl;
    const real eps = 1e-5;
    deltas[0] = Tl * eps;
    deltas[1] = sqrt(a2) * eps;
    deltas[2] = deltas[1];
    deltas[3] = deltas[1];
    deltas[4] = a2 * eps;
    vals[0] = Tl; vals[1] = ul[0]; vals[2] = ul[1]; vals[3] = ul[2]; vals[4] = Pgl;
    for (int j = 0; j < 5; ++j) {
      real tmp = vals[j];
      vals[j] += deltas[j];
      real u2[3] = {vals[1], vals[2], vals[3]};
      reflecting_flux_(p, f1, vals[4], vals[0], u2, an, sada, us_n, eta_c);
      const real rh = 1. / (tmp - vals[j]);
      for (int i = 1; i < 5; ++i) subfj(&M_(bj, i, j), (f0[i] - f1[i]) * rh);
      vals[j] = tmp;
    }
  } else if (t == OG_ISENTROPIC_INFLOW || t == OG_FIXED_MASS) {
    /* src/bcIsentropic.loci:171-247, src/bcFixedMass.loci:166-228:
     * one-sided differences of the face state's Euler flux in the cell's
     * primitives, all rows */
    const double *bs = p->bcs[OG_SET(g, f)];
    const real mflux = t == OG_FIXED_MASS ? set_mass_flux(g, p, OG_SET(g, f)) : 0.0;
    real vals[5] = {q[c], q[nc + c], q[2 * nc + c], q[3 * nc + c], q[4 * nc + c]};
    real Tb, pgb, ub[3], f0[5], f1[5], deltas[5];
    state_bc_(p, t, bs, mflux, vals[0], vals + 1, vals[4], an, &Tb, ub, &pgb);
    euler_flux_(f0, pgb, Tb, ub, an, sada, p->Pambient, p->Rt, p->gamma, us_n);
    const real eps = 1e-5;
    const real ai = sqrt(p->gamma * p->Rt * vals[0]);
    deltas[0] = vals[0] * eps;
    deltas[1] = ai * eps;
    deltas[2] = deltas[1];
    deltas[3] = deltas[1];
    deltas[4] = ai * ai * eps;
    for (int j = 0; j < 5; ++j) {
      real tmp = vals[j];
      vals[j] += deltas[j];
      state_bc_(p, t, bs, mflux, vals[0], vals + 1, vals[4], an, &Tb, ub, &pgb);
      euler_flux_(f1, pgb, Tb, ub, an, sada, p->Pambient, p->Rt, p->gamma, us_n);
      const real rh = 1. / (tmp - vals[j]);
      for (int i = 0; i < 5; ++i) subfj(&M_(bj, i, j), (f0[i] - f1[i]) * rh);
      vals[j] = tmp;
    }
  } else if (t == OG_WALL) {
    const int mi = 1, ei = 4;
    real uf[3] = {qf[nf + f], qf[2 * nf + f], qf[3 * nf + f]};
    subfj(&M_(bj, mi + 0, ei), an[0] * sada);
    subfj(&M_(bj, mi + 1, ei), an[1] * sada);
    subfj(&M_(bj, mi + 2, ei), an[2] * sada);
    subfj(&M_(bj, ei, ei), sada * dot3(uf, an));
  } else if (t == OG_SYMMETRY) {
    const real pp = q[4 * nc + c] + p->Pambient;
    const real T = q[c];
    const real rho = pp / (p->Rt * T);
    const real drdt = -rho / T;
    const real drdp = rho / pp;
    const int mi = 1;
    real ui[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
    const real U = sqrt(dot3(ui, ui));
    const real uit = dot3(an, ui) - us_n;
    const real coef1 = -0.5 * rho * sada;
    const real coef2 = -0.5 * sada;
    for (int j = 0; j < 3; ++j) {
      const real coefj = coef1 * (ui[j] * uit / (U + 1e-30) + an[j] * U);
      for (int i = 0; i < 3; ++i) addfj(&M_(bj, mi + i, mi + j), coefj * an[i]);
    }
    for (int i = 0; i < 3; ++i) {
      real coefx = 0.5 * U * uit;
      addfj(&M_(bj, mi + i, 0), coef2 * (coefx * drdt) * an[i]);
      addfj(&M_(bj, mi + i, 4), coef2 * (1.0 + coefx * drdp) * an[i]);
    }
  }
  if (p->viscous) {
    real mu_f, k_f, uf[3], vf[4];
    viscous_face(g, p, q, qf, grad, f, &mu_f, &k_f, uf, vf);
    real dv[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
    visc_jac(bj, -1.0, an, sada, dot3(an, dv), mu_f, k_f, uf, vf);
  }
}

/* Assemble the linear system of one steady Newton step:
 *   rhs = src                       src/solverTimestep.loci:337-338
 *   cflpdt, dtcfl, dtau             src/solverSetup.loci:107-119, src/solverTimestep.loci:290-329
 *   fjp/fjm (Roe + viscous)         src/fluxInviscid.loci:1190-1221, src/fluxViscous.loci:315-403
 *   srcJ, fluid_D                   src/solverLinearSystem.loci:106-120,176-224
 *   B = rhs rounded to real_fj      src/solverLinearSystem.loci:45-54
 *   LU of D with pivots             src/linearSolverSGS.loci:43-50 */
void og_assemble(const og_grid *g, const og_params *p, const double *q, double *src,
                 double *dtau_out, void *fjp_v, void *fjm_v, void *D_v, int *pivot, void *B_v,
                 og_norms *nrm) {
  og_assemble2(g, p, q, src, dtau_out, fjp_v, fjm_v, D_v, pivot, B_v, nrm, NULL);
}

void og_assemble2(const og_grid *g, const og_params *p, const double *q, double *src,
                  double *dtau_out, void *fjp_v, void *fjm_v, void *D_v, int *pivot, void *B_v,
                  og_norms *nrm, void *Draw_v) {
  og_assemble_t(g, p, q, NULL, NULL, src, dtau_out, fjp_v, fjm_v, D_v, pivot, B_v, nrm, Draw_v, NULL);
}

/* KEC's level n+1 from the iterate q* (src/solverTimestep.loci:366-385,
 * src/solverNewton.loci:127-157): harmonic-mean temperature, biased
 * pressure, density-weighted velocity; out = {T, u, v, w, pg, rho} */
static void kec_np1(const og_params *p, real Ts, real pgs, const real *us, real Tn, real pgn, const real *un,
                    real peps, real *out) {
  const real Tnp1 = Tn * Ts / (2. * Tn - Ts);
  const real pgnp1 = (2. * pgs - (1. - peps) * pgn) / (1 + peps);
  const real rhon = (pgn + p->Pambient) / (p->Rt * Tn);
  const real rhonp = maxr((pgnp1 + p->Pambient) / (p->Rt * Tnp1), 1e-15);
  const real srhonp = sqrt(rhonp), srhon = sqrt(rhon);
  const real w1 = (srhonp + srhon) / srhonp, w2 = srhon / srhonp;
  out[0] = Tnp1;
  for (int d = 0; d < 3; ++d) out[1 + d] = w1 * us[d] - w2 * un[d];
  out[4] = pgnp1;
  out[5] = rhonp;
}

/* localPBias, src/solverTimestep.loci:346-351 */
static real kec_pbias(const og_params *p, real cflpdt) {
  const real cfl = cflpdt * p->dtmax;
  const real factor = minr(1.0, maxr(1e-40, 0.1 * (cfl - 1.)));
  return maxr(p->pbias_control, pow(factor, 0.7));
}

/* KEC scalar level n+1 (and X_np1) from the iterate,
 * src/solverScalarTransport.loci:367-395 */
static real kec_xnp1(real X, real Xn) {
  const real Xstar = maxr(X, 1e-15);
  real dXstar = maxr((Xstar - Xn) / Xstar, -1.95);
  if (dXstar < 0.0) dXstar += 0.25 * dXstar * dXstar;
  return Xstar * (1. + dXstar);
}

/* density, momentum and total energy per volume of a primitive state */
static void cons_of(const og_params *p, real T, real pg, const real *u, real *Q) {
  const real rho = (pg + p->Pambient) / (p->Rt * T);
  Q[0] = rho;
  for (int d = 0; d < 3; ++d) Q[1 + d] = rho * u[d];
  Q[4] = rho * (p->Rt * T / (p->gamma - 1.) + .5 * dot3(u, u));
}

void og_assemble_t(const og_grid *g, const og_params *p, const double *q, const double *qn, const double *qm1,
                   double *src, double *dtau_out, void *fjp_v, void *fjm_v, void *D_v, int *pivot, void *B_v,
                   og_norms *nrm, void *Draw_v, double *pbias_out) {
  const int bdf2 = p->unsteady == 1 && qn && qm1, kec = p->unsteady == 2 && qn;
  real_fj *Draw = Draw_v;
  int nc = g->nc, nf = g->nf;
  real_fj *fjp = fjp_v, *fjm = fjm_v, *D = D_v, *B = B_v;
  double *qf = calloc((size_t)7 * nf, sizeof(double));
  double *grad = malloc(sizeof(double) * 21 * nc), *phi = malloc(sizeof(double) * 7 * nc);
  double *iflux = malloc(sizeof(double) * 5 * nf), *vflux = malloc(sizeof(double) * 4 * nf);
  double *maxev = malloc(sizeof(double) * nf);
  og_face_states(g, p, q, qf);
  og_gradients(g, p, q, qf, grad, phi);
  og_fluxes(g, p, q, qf, grad, phi, iflux, vflux, maxev);
  og_adj a;
  build_adj(g, &a);
  gather_src(g, p, &a, iflux, vflux, src);
  double *rhs = bdf2 || kec ? malloc(sizeof(double) * 5 * (size_t)nc) : NULL; /* norms of the rhs */
  if (nrm && !rhs) residual_norms(g, src, nrm);

  /* face Jacobians */
  const real us_n = 0.0;
  real *eta_p = malloc(sizeof(real) * nc);
  eta_impl(g, p, q, grad, eta_p);
  real_fj *bblk = calloc((size_t)nf * 25, sizeof(real_fj));
  OG_PAR
  for (int f = 0; f < nf; ++f) {
    real_fj *P = fjp + 25 * (size_t)f, *Mm = fjm + 25 * (size_t)f;
    for (int i = 0; i < 25; ++i) P[i] = Mm[i] = 0;
    int t = OG_BT(g, f);
    if (t == OG_INTERIOR || t == OG_PERIODIC) {
      const int rc = t == OG_INTERIOR ? g->cr[f] : g->cl[g->pface[f]];
      const real eta = maxr(eta_p[g->cl[f]], eta_p[rc]); /* :1196-1220 */
      real an[3], xf[3], xl[3], xr[3], qL[5], qR[5];
      fnormal(g, f, an);
      fcenter(g, f, xf);
      int c = g->cl[f];
      ccenter(g, c, xl);
      int o = rcell(g, f, xr);
      recon(g, p, q, grad, phi, c, xl, xf, qL);
      recon(g, p, q, grad, phi, o, xr, xf, qR);
      real UL[3] = {qL[1], qL[2], qL[3]}, UR[3] = {qR[1], qR[2], qR[3]};
      roe_jacobian(P, qL[4], qL[0], UL, qR[4], qR[0], UR, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n, eta, 1);
      roe_jacobian(Mm, qL[4], qL[0], UL, qR[4], qR[0], UR, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n, eta, -1);
      if (p->viscous) {
        real mu_f, k_f, uf[3], vf[4];
        viscous_face(g, p, q, qf, grad, f, &mu_f, &k_f, uf, vf);
        real dv[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
        const real dist = dot3(an, dv);
        visc_jac(Mm, -1.0, an, g->area[f], dist, mu_f, k_f, uf, vf);
        visc_jac(P, 1.0, an, g->area[f], dist, mu_f, k_f, uf, vf);
      }
      if (p->temporal_damping > 0) { /* src/fluxInviscid.loci:1230-1251 */
        damping_jac(P, 1.0, qL[4], qL[0], UL, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n,
                    p->temporal_damping);
        damping_jac(Mm, -1.0, qR[4], qR[0], UR, an, g->area[f], p->Pambient, p->Rt, p->gamma, us_n,
                    p->temporal_damping);
      }
    } else {
      boundary_block(g, p, q, qf, grad, phi, f, eta_p[g->cl[f]], bblk + 25 * (size_t)f);
    }
  }

  /* cell loop: cflpdt, dtau, srcJ, D, B, LU */
  OG_PAR
  for (int c = 0; c < nc; ++c) {
    /* cflpdt: lower, upper, boundary */
    real meva = 0.0;
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) meva += maxev[a.lo[j]] * g->area[a.lo[j]];
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) meva += maxev[a.up[j]] * g->area[a.up[j]];
    for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) meva += maxev[a.bd[j]] * g->area[a.bd[j]];
    const real cflpdt = 0.5 * meva / g->vol[c];
    real dtcfl = p->dtmax;
    if (p->cflmax > 0) dtcfl = minr(dtcfl, p->cflmax / cflpdt);
    /* dtau */
    real dt = dtcfl;
    const real rvol = 1. / g->vol[c];
    const real Pamb = p->Pambient;
    const real pg = q[4 * nc + c];
    const real pp = pg + Pamb;
    const real T = q[c];
    const real Rt = p->Rt;
    const real rP = 1. / (Rt * T);
    const real rho = (pg + Pamb) * rP;
    const real rT = -rho / T;
    const real gm = p->gamma;
    real U[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
    const real qq = 0.5 * dot3(U, U);
    const real fT = rT * qq;
    const real fP = (1. / (gm - 1) + rP * qq);
    const real reta = 1. / eta_p[c]; /* src/solverLinearSystem.loci:195 */
    const real a2 = gm * Rt * T;
    const real pterm = (1. - reta) * (gm - 1.) / a2;
    const real cterm = fP * rT * reta - fT * rP * reta;
    const real q2 = 2. * qq;
    real s[5];
    for (int k = 0; k < 5; ++k) s[k] = src[k * nc + c];
    const real dsu = s[1] * U[0] + s[2] * U[1] + s[3] * U[2];
    const real dtemp = rvol * (s[0] * (fP * rT * reta + fT * pterm) + (dsu - s[4] - q2 * s[0]) * (rT * (pterm + rP * reta))) / (cterm * rT);
    dt = minr(dt, p->urelax * T / (fabs(dtemp) + 1e-30));
    const real dp = rvol * (rT * (q2 * s[0] - dsu + s[4]) - fT * s[0]);
    dt = minr(dt, p->urelax * pp / (fabs(dp) + 1e-30));
    if (dtau_out) dtau_out[c] = dt;

    real_fj sJ[25];
    for (int i = 0; i < 25; ++i) sJ[i] = 0;
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
      const real_fj *P = fjp + 25 * (size_t)a.up[j];
      for (int i = 0; i < 25; ++i) sJ[i] -= P[i];
    }
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
      const real_fj *Mm = fjm + 25 * (size_t)a.lo[j];
      for (int i = 0; i < 25; ++i) sJ[i] += Mm[i];
    }
    for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) {
      const real_fj *bb = bblk + 25 * (size_t)a.bd[j];
      for (int i = 0; i < 25; ++i) sJ[i] += bb[i];
    }
    real_fj *Dc = D + 25 * (size_t)c;
    for (int i = 0; i < 25; ++i) Dc[i] = -sJ[i];
    real coef = g->vol[c] / dt;
    real rta = reta, ptm = pterm;
    if (bdf2) { /* BDF2 fluid_D, src/solverLinearSystem.loci:384-420 */
      coef = 1.5 * g->vol[c] / dt;
      const real pseudo_dt = p->pseudo_cflmax / cflpdt;
      const real w = 1. / (1. + pseudo_dt / dt);
      rta = 1. / (1. + w * (eta_p[c] - 1.));
      ptm = (1. - rta) * (gm - 1.) / a2;
    }
    const real pb = kec ? kec_pbias(p, cflpdt) : 0.0;
    if (kec && pbias_out) pbias_out[c] = pb;
    if (kec) { /* KEC fluid_D: (V/dtau) dQ/dq dq/dq*, src/solverLinearSystem.loci:266-378 */
      real dQ[5][5], dqs[5][5];
      memset(dQ, 0, sizeof(dQ));
      memset(dqs, 0, sizeof(dqs));
      dQ[0][0] = rT; dQ[1][0] = rT * U[0]; dQ[2][0] = rT * U[1]; dQ[3][0] = rT * U[2]; dQ[4][0] = fT;
      dQ[0][4] = rP; dQ[1][4] = rP * U[0]; dQ[2][4] = rP * U[1]; dQ[3][4] = rP * U[2]; dQ[4][4] = fP;
      dQ[1][1] = rho; dQ[4][1] = rho * U[0];
      dQ[2][2] = rho; dQ[4][2] = rho * U[1];
      dQ[3][3] = rho; dQ[4][3] = rho * U[2];
      const real Tn = qn[c], pgn = qn[4 * nc + c];
      const real un[3] = {qn[nc + c], qn[2 * nc + c], qn[3 * nc + c]};
      real np1[6];
      kec_np1(p, T, pg, U, Tn, pgn, un, pb, np1);
      const real rhon = (pgn + Pamb) / (Rt * Tn);
      const real Pcoef = 2.0 / (1 + pb);
      const real srhonp1 = sqrt(np1[5]), srhon = sqrt(rhon);
      const real ucoef = (srhonp1 + srhon) / srhonp1;
      const real tfactor = 1. / (2 * Tn - T);
      const real dTdTs = 2. * Tn * Tn * tfactor * tfactor;
      const real drdt = -dTdTs * rho / T;
      const real drdp = Pcoef * rho / (pg + Pamb);
      dqs[0][0] = dTdTs;
      dqs[1][1] = ucoef; dqs[2][2] = ucoef; dqs[3][3] = ucoef;
      dqs[4][4] = Pcoef;
      const real ucoef2 = 0.5 * srhon / (srhonp1 * srhonp1 * srhonp1);
      for (int d = 0; d < 3; ++d) {
        dqs[1 + d][0] += ucoef2 * drdt * (un[d] - U[d]);
        dqs[1 + d][4] += ucoef2 * drdp * (un[d] - U[d]);
      }
      const real kcoef = g->vol[c] / dt;
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) dQ[i][j] *= kcoef;
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
          for (int k = 0; k < 5; ++k) addfj(&M_(Dc, i, j), dQ[i][k] * dqs[k][j]);
      /* KEC rhs, src/solverTimestep.loci:366-396 */
      const real e0n = p->Rt * Tn / (p->gamma - 1.) + .5 * dot3(un, un);
      const real e0np = p->Rt * np1[0] / (p->gamma - 1.) + .5 * dot3(np1 + 1, np1 + 1);
      const real rdtvol = g->vol[c] / p->dtmax;
      const real rhonp = np1[5];
      s[0] = s[0] - (rhonp - rhon) * rdtvol;
      for (int d = 0; d < 3; ++d) s[1 + d] = s[1 + d] - (rhonp * np1[1 + d] - rhon * un[d]) * rdtvol;
      s[4] = s[4] - (rhonp * e0np - rhon * e0n) * rdtvol;
      for (int k = 0; k < 5; ++k) rhs[k * (size_t)nc + c] = s[k];
    } else {
    const real reta = rta, pterm = ptm;
    addfj(&M_(Dc, 0, 0), coef * rT);
    addfj(&M_(Dc, 1, 0), coef * rT * U[0]);
    addfj(&M_(Dc, 2, 0), coef * rT * U[1]);
    addfj(&M_(Dc, 3, 0), coef * rT * U[2]);
    addfj(&M_(Dc, 4, 0), coef * fT);
    addfj(&M_(Dc, 0, 4), coef * (pterm + rP * reta));
    addfj(&M_(Dc, 1, 4), coef * (pterm + rP * reta) * U[0]);
    addfj(&M_(Dc, 2, 4), coef * (pterm + rP * reta) * U[1]);
    addfj(&M_(Dc, 3, 4), coef * (pterm + rP * reta) * U[2]);
    addfj(&M_(Dc, 4, 4), coef * (fT * pterm / rT + fP * reta));
    addfj(&M_(Dc, 1, 1), coef * rho);
    addfj(&M_(Dc, 4, 1), coef * rho * U[0]);
    addfj(&M_(Dc, 2, 2), coef * rho);
    addfj(&M_(Dc, 4, 2), coef * rho * U[1]);
    addfj(&M_(Dc, 3, 3), coef * rho);
    addfj(&M_(Dc, 4, 3), coef * rho * U[2]);
    }
    if (bdf2) { /* BDF2 rhs, src/solverTimestep.loci:400-428 (w = 1/2) */
      real un[3] = {qn[nc + c], qn[2 * nc + c], qn[3 * nc + c]};
      real um[3] = {qm1[nc + c], qm1[2 * nc + c], qm1[3 * nc + c]};
      real Qp[5], Qn[5], Qm[5];
      cons_of(p, T, pg, U, Qp);
      cons_of(p, qn[c], qn[4 * nc + c], un, Qn);
      cons_of(p, qm1[c], qm1[4 * nc + c], um, Qm);
      const real rdtvol = g->vol[c] / p->dtmax;
      const real w = 0.5;
      const real w1 = (1.0 + w) * rdtvol, w2 = w * rdtvol;
      for (int k = 0; k < 5; ++k) s[k] = (s[k] - w1 * (Qp[k] - Qn[k]) + w2 * (Qn[k] - Qm[k]));
      for (int k = 0; k < 5; ++k) rhs[k * (size_t)nc + c] = s[k];
    }
    for (int k = 0; k < 5; ++k) B[5 * (size_t)c + k] = (real_fj)s[k];
    if (Draw) memcpy(Draw + 25 * (size_t)c, Dc, 25 * sizeof(real_fj));
    lu_decompose(Dc, pivot + 5 * (size_t)c, 5);
  }
  if (nrm && rhs) residual_norms(g, rhs, nrm); /* src/solverResidual.loci:113-150 */
  free(rhs);
  free(bblk);
  free(eta_p);
  free_adj(&a);
  free(qf); free(grad); free(phi); free(iflux); free(vflux); free(maxev);
}

/* Block SGS, src/linearSolverSGS.loci:57-69 (Jacobi start), :124-127 and
 * :215-219 (periodic ghost copies), :186-208 (forward), :277-300 (backward).
 * Cells are visited in index order, a topological order of the cl -> cr
 * face graph since every interior face has cl < cr. */
void og_sgs(const og_grid *g, int nsweeps, const void *fjp_v, const void *fjm_v, const void *D_v,
            const int *pivot, const void *B_v, double *X) {
  int nc = g->nc, nf = g->nf;
  const real_fj *fjp = fjp_v, *fjm = fjm_v, *D = D_v, *B = B_v;
  og_adj a;
  build_adj(g, &a);
  double *Xg = calloc((size_t)5 * nf, sizeof(double)); /* ghost copies per periodic face */
  for (int c = 0; c < nc; ++c) {
    real b[5];
    for (int k = 0; k < 5; ++k) b[k] = B[5 * (size_t)c + k];
    lu_solve(D + 25 * (size_t)c, pivot + 5 * (size_t)c, b, X + 5 * (size_t)c, 5);
  }
  for (int it = 0; it < nsweeps; ++it) {
    for (int pass = 0; pass < 2; ++pass) {
      for (int f = 0; f < nf; ++f)
        if (OG_BT(g, f) == OG_PERIODIC)
          for (int k = 0; k < 5; ++k) Xg[5 * (size_t)f + k] = X[5 * (size_t)g->cl[g->pface[f]] + k];
      for (int ii = 0; ii < nc; ++ii) {
        int c = pass == 0 ? ii : nc - 1 - ii;
        real t[5], v[5];
        for (int k = 0; k < 5; ++k) t[k] = -(real)B[5 * (size_t)c + k];
        for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
          int f = a.up[j];
          const real *xn = OG_BT(g, f) == OG_PERIODIC ? Xg + 5 * (size_t)f : X + 5 * (size_t)g->cr[f];
          dotprod_accum(fjm + 25 * (size_t)f, xn, t); /* fluid_U = fjm */
        }
        for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
          int f = a.lo[j];
          real_fj L[25];
          const real_fj *P = fjp + 25 * (size_t)f;
          for (int i = 0; i < 25; ++i) L[i] = -P[i]; /* fluid_L = -fjp */
          dotprod_accum(L, X + 5 * (size_t)g->cl[f], t);
        }
        for (int k = 0; k < 5; ++k) v[k] = -t[k];
        lu_solve(D + 25 * (size_t)c, pivot + 5 * (size_t)c, v, X + 5 * (size_t)c, 5);
      }
    }
  }
  free(Xg);
  free_adj(&a);
}

/* Steady Newton steps, src/solverNewton.loci:62-78 (damped update). */
/* nsteps Newton iterations at q; with time levels qn, qm1 (BDF2) the
 * systems carry the time terms and the update limit is the BDF2 one */
static int newton_run(const og_grid *g, const og_params *p, double *q, const double *qn, const double *qm1,
                      int nsteps, double *hist) {
  const int bdf2 = p->unsteady == 1 && qn && qm1, kec = p->unsteady == 2 && qn;
  double *pbias = kec ? malloc(sizeof(double) * g->nc) : NULL;
  int nc = g->nc, nf = g->nf;
  double *src = malloc(sizeof(double) * 5 * nc);
  real_fj *fjp = malloc(sizeof(real_fj) * 25 * (size_t)nf), *fjm = malloc(sizeof(real_fj) * 25 * (size_t)nf);
  real_fj *D = malloc(sizeof(real_fj) * 25 * (size_t)nc), *B = malloc(sizeof(real_fj) * 5 * (size_t)nc);
  int *piv = malloc(sizeof(int) * 5 * (size_t)nc);
  double *X = malloc(sizeof(double) * 5 * (size_t)nc);
  real_fj *Draw = p->linear_solver != 0 ? malloc(sizeof(real_fj) * 25 * (size_t)nc) : NULL;
  double *ssrc = NULL, *sfp = NULL, *sfm = NULL, *sD = NULL, *sB = NULL, *sX = NULL;
  const int ns = p->turb == 2 ? 2 : p->turb ? 1 : 0; /* turbulence scalars */
  if (ns) {
    ssrc = malloc(sizeof(double) * ns * nc); sfp = malloc(sizeof(double) * ns * nf);
    sfm = malloc(sizeof(double) * ns * nf); sD = malloc(sizeof(double) * ns * nc);
    sB = malloc(sizeof(double) * ns * nc); sX = malloc(sizeof(double) * ns * nc);
  }
  for (int s = 0; s < nsteps; ++s) {
    /* the scalar systems from the state at the start of the step, each
     * solved on its own (scalarTransportP(nuTilde | k | w, ...)) */
    if (ns) {
      if (p->turb == 2) og_sst_assemble_t(g, p, q, qn, qm1, ssrc, sfp, sfm, sD, sB, NULL, NULL);
      else og_sa_assemble_t(g, p, q, qn, qm1, ssrc, sfp, sfm, sD, sB, NULL);
      for (int j = 0; j < ns; ++j) {
        const size_t oc = (size_t)j * nc, of = (size_t)j * nf;
        if (p->linear_solver == 1) /* LSGSScalarSolve, src/solverScalarTransport.loci:633-637 */
          og_sa_lsgs(g, p->lsgs_maxiter, p->lsgs_abstol, p->lsgs_reltol, p->lsgs_relax, sfp + of, sfm + of, sD + oc,
                     sB + oc, sX + oc, NULL, NULL);
        else if (p->linear_solver == 2) /* FSGSScalarSolve, :639-643 */
          og_sa_fsgs(g, p->sgs_iter, sfp + of, sfm + of, sD + oc, sB + oc, sX + oc);
        else /* SGSScalarSolve, :627-631 */
          og_sa_sgs(g, p->sgs_iter, sfp + of, sfm + of, sD + oc, sB + oc, sX + oc);
      }
    }
    og_norms nrm;
    og_assemble_t(g, p, q, qn, qm1, src, NULL, fjp, fjm, D, piv, B, &nrm, Draw, pbias);
    if (hist) {
      hist[6 * s + 0] = nrm.rtrms; hist[6 * s + 1] = nrm.etrms; hist[6 * s + 2] = nrm.mtrms;
      hist[6 * s + 3] = nrm.rtmax; hist[6 * s + 4] = nrm.etmax; hist[6 * s + 5] = nrm.mtmax;
    }
    if (p->linear_solver == 1)  /* fluidLinearSolver lsgs, src/solverNewton.loci:256-259 */
      og_lsgs(g, p->lsgs_maxiter, p->lsgs_abstol, p->lsgs_reltol, p->lsgs_relax, fjp, fjm, Draw, B, X, NULL, NULL);
    else if (p->linear_solver == 2)  /* fluidLinearSolver fsgs */
      og_fsgs(g, p->sgs_iter, fjp, fjm, Draw, B, X);
    else
      og_sgs(g, p->sgs_iter, fjp, fjm, D, piv, B, X);
    OG_PAR
    for (int c = 0; c < nc; ++c) {
      const real *dq = X + 5 * (size_t)c;
      const real dT = dq[0], dp = dq[4];
      real alpha = 1.0;
      if (bdf2) { /* src/solverNewton.loci:160-180 */
        const real relax = 0.75;
        const real pnit = q[4 * nc + c] + p->Pambient;
        alpha = minr(alpha, relax * pnit / maxr(fabs(dp), 1e-30));
        alpha = minr(alpha, relax * q[c] / maxr(fabs(dT), 1e-30));
      } else if (kec) { /* :93-125 */
        const real relax = 0.8;
        const real Tstar = q[c], Tn = qn[c];
        const real dTlim = (relax * (2 * Tn - Tstar)) / 2.0;
        alpha = minr(alpha, dTlim / maxr(fabs(dT), 1e-30));
        const real pgstar = q[4 * nc + c], peps = pbias[c];
        const real pgnp1 = (2. * pgstar - (1. - peps) * qn[4 * nc + c]) / (1 + peps);
        const real dplim = (1 + peps) * relax * (pgnp1 + p->Pambient) / 2.0;
        alpha = minr(alpha, dplim / maxr(fabs(dp), 1e-30));
      } else {
        alpha = minr(alpha, 0.9 * q[c] / (fabs(dT) + 1e-30));
        alpha = minr(alpha, 0.9 * (q[4 * nc + c] + p->Pambient) / (fabs(dp) + 1e-30));
      }
      q[c] = q[c] + alpha * dT;
      q[4 * nc + c] = q[4 * nc + c] + alpha * dp;
      for (int k = 1; k < 4; ++k) q[k * nc + c] = q[k * nc + c] + alpha * dq[k];
    }
    for (int j = 0; j < ns; ++j) /* X_next, src/solverScalarTransport.loci:400-408 */
      for (int c = 0; c < nc; ++c) {
        const real Xc = q[(5 + j) * nc + c], dX = sX[(size_t)j * nc + c];
        real tscale = 1.0;
        tscale = minr(tscale, p->urelax * Xc / (fabs(dX) + 1e-15));
        real Xn = Xc + tscale * dX;
        q[(5 + j) * nc + c] = maxr(Xn, 1e-15);
      }
  }
  if (kec && nsteps > 0) { /* q{n+1} from the last iterate (src/solverNewton.loci:127-157 with
                            * localPBias_next = the last iteration's localPBias; scalars
                            * src/solverScalarTransport.loci:367-377) */
    for (int c = 0; c < nc; ++c) {
      const real us[3] = {q[nc + c], q[2 * nc + c], q[3 * nc + c]};
      const real un[3] = {qn[nc + c], qn[2 * nc + c], qn[3 * nc + c]};
      real np1[6];
      kec_np1(p, q[c], q[4 * nc + c], us, qn[c], qn[4 * nc + c], un, pbias[c], np1);
      for (int k = 0; k < 5; ++k) q[k * nc + c] = np1[k];
      for (int j = 0; j < ns; ++j) q[(5 + j) * nc + c] = kec_xnp1(q[(5 + j) * nc + c], qn[(5 + j) * nc + c]);
    }
  }
  free(pbias);
  free(src); free(fjp); free(fjm); free(D); free(B); free(piv); free(X); free(Draw);
  free(ssrc); free(sfp); free(sfm); free(sD); free(sB); free(sX);
  return 0;
}

int og_newton(const og_grid *g, const og_params *p, double *q, int nsteps, double *hist) {
  return newton_run(g, p, q, NULL, NULL, nsteps, hist);
}

/* BDF2 dual time stepping: q{n,0} = q{n} (src/solverNewton.loci), newton_iter
 * iterations, q{n+1} = the last iterate (:182-189), then the levels shift */
int og_unsteady(const og_grid *g, const og_params *p, double *q, double *qm1, int nsteps, double *hist) {
  if (!p->unsteady || p->newton_iter < 1) return -1;
  const int nq = 5 + (p->turb == 2 ? 2 : p->turb ? 1 : 0);
  const size_t n = (size_t)nq * g->nc;
  double *qn = malloc(sizeof(double) * n);
  for (int s = 0; s < nsteps; ++s) {
    memcpy(qn, q, sizeof(double) * n);
    newton_run(g, p, q, qn, qm1, p->newton_iter, hist ? hist + 6 * (size_t)s * p->newton_iter : NULL);
    memcpy(qm1, qn, sizeof(double) * n);
  }
  free(qn);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Spalart-Allmaras nuTilde equation (baseline model, standard          */
/* production), segregated as in the reference                         */
/* ------------------------------------------------------------------ */

/* scalar reconstruction lefts/rights(X) of variable k at face centre xf */
static real recon_k(const og_grid *g, const double *q, const double *grad, const double *phi, int k, int c,
                    const real *xcell, const real *xf) {
  const int nc = g->nc;
  real d[3] = {xf[0] - xcell[0], xf[1] - xcell[1], xf[2] - xcell[2]};
  real gd = grad[(3 * k) * nc + c] * d[0] + grad[(3 * k + 1) * nc + c] * d[1] + grad[(3 * k + 2) * nc + c] * d[2];
  return q[k * nc + c] + phi[k * nc + c] * gd;
}

/* Per-cell SA terms (turbulence/SA/SA_turb.loci:93-165, 258-315, 347-351):
 * production - destruction into *srcv, the destruction Jacobian into *srcJ. */
static void sa_core(const og_params *p, real T, real pg, real nut, real vortMag, real y, real vol, real *srcv,
                    real *srcJ);
static void sa_cell_terms(const og_grid *g, const og_params *p, const double *q, const double *grad, int c,
                          real *srcv, real *srcJ) {
  const int nc = g->nc;
  /* vortMag from the cell velocity gradient: d u_i / d x_j at grad[3(1+i)+j]
   * (src/solverTurbulence.loci:44-48,77-79) */
#define GU(i, j) grad[(3 * (1 + (i)) + (j)) * nc + c]
  const real wx = GU(2, 1) - GU(1, 2), wy = GU(0, 2) - GU(2, 0), wz = GU(1, 0) - GU(0, 1);
#undef GU
  const real vortMag = sqrt(wx * wx + wy * wy + wz * wz);
  sa_core(p, q[c], q[4 * nc + c], q[5 * nc + c], vortMag, g->dist[c], g->vol[c], srcv, srcJ);
}

/* SA terms of one cell from its state, vorticity magnitude, wall distance
 * and volume */
static void sa_core(const og_params *p, real T, real pg, real nut, real vortMag, real y, real vol, real *srcv,
                    real *srcJ) {
  const real rho = (pg + p->Pambient) / (p->Rt * T);
  const real mu = suth_mu(p, T);
  const real xcoeff = rho * nut / mu;
  const real fv1 = sa_fv1(p, xcoeff);
  const real fv2 = 1.0 - (xcoeff / (1.0 + xcoeff * fv1));
  const real k = p->sa_k;
  const real denom = k * k * y * y;
  real S_tilda = 0.0;
  S_tilda += vortMag;
  S_tilda += nut * fv2 / denom;
  const real S_t = maxr(S_tilda, 0.3 * vortMag);
  const real rclip = 10;
  const real r = (S_t < nut / (rclip * denom)) ? rclip : nut / (S_t * denom);
  const real r2 = r * r, r6 = r2 * r2 * r2;
  const real gc = r + p->sa_cw2 * (r6 - r);
  const real cw36 = pow(p->sa_cw3, 6.0);
  const real g2 = gc * gc, g6 = g2 * g2 * g2;
  const real fw = gc * pow((1.0 + cw36) / (g6 + cw36), 1.0 / 6.0);
  const real cw1 = (p->sa_cb1) / (k * k) + (1.0 + (p->sa_cb2)) / (p->sa_sigma);
  real sv = 0.0;
  sv += rho * nut * S_t * p->sa_cb1 * vol;
  const real t = nut / y;
  sv -= rho * fw * cw1 * t * t * vol;
  /* destruction Jacobian */
  real rp;
  const real rclip2 = 9e6;
  if (S_t < nut / (rclip2 * denom)) {
    rp = 0;
  } else {
    const real chi = xcoeff, chiP = 1. / mu;
    const real chi3 = chi * chi * chi, chi4 = chi3 * chi, chi6 = chi3 * chi3;
    const real cv1 = p->sa_cv1, cv13 = cv1 * cv1 * cv1;
    const real fv2p = chi * (4 * (chi3 * chiP) / (chi3 + cv13) - 3 * (chi6 * chiP) / pow(chi3 + cv13, 2.0)) /
                          pow(chi4 / (chi3 + cv13) + 1, 2.0) -
                      chiP / (chi4 / (chi3 + cv13) + 1);
    const real Stp = (fv2 / rho + nut * fv2p) / (denom);
    rp = 1. / (rho * denom * S_t) - nut * Stp / (denom * S_t * S_t);
  }
  const real gp = rp + p->sa_cw2 * (6. * pow(r, 5.0) * rp - rp);
  const real term = pow((cw36 + 1.) / (cw36 + g6), 1. / 6.);
  const real term2 = term * term, term5 = term2 * term2 * term;
  const real rterm5 = 1. / term5, rsum = 1. / (g6 + cw36);
  const real fwp = gp * term - gp * ((cw36 + 1) * g6) * rsum * rsum * rterm5;
  *srcv = sv;
  *srcJ = -(vol * cw1 * rho * nut * (rho * nut * fwp + 2. * fw) / (y * y));
}

/* rho and laminar mu of a face: volume-weighted interior averages, the face
 * state on boundaries (src/fluxViscous.loci:121-133,210-215, solverEoS.loci:53) */
static void face_rho_mu(const og_grid *g, const og_params *p, const double *q, const double *qf, int f,
                        real *rho_f, real *mu_f) {
  const int nc = g->nc, nf = g->nf, t = OG_BT(g, f), c = g->cl[f];
  if (t == OG_INTERIOR || t == OG_PERIODIC) {
    real xr[3];
    const int o = rcell(g, f, xr);
    const real vl = g->vol[c], vr = g->vol[o];
    const real rl = (q[4 * nc + c] + p->Pambient) / (p->Rt * q[c]);
    const real rr = (q[4 * nc + o] + p->Pambient) / (p->Rt * q[o]);
    const real factor = 1. / (vl + vr);
    *rho_f = factor * (vl * rr + vr * rl);
    *mu_f = (vl * suth_mu(p, q[o]) + vr * suth_mu(p, q[c])) / (vl + vr);
  } else {
    *rho_f = (qf[4 * nf + f] + p->Pambient) / (p->Rt * qf[f]);
    *mu_f = suth_mu(p, qf[f]);
  }
}

/* One steady Newton step of the nuTilde equation:
 *   convection mdot = iflux[0] upwind, floored at 0, with the continuity
 *   stabilisation -X rhodot (src/solverScalarTransport.loci:50-135, 478-535);
 *   SA sources and destruction Jacobian; baseline diffusion diff1 and the
 *   cb2 term with their Jacobians (turbulence/SA/SA_turb.loci:354-482,
 *   737-754); flux Jacobians into srcJ (solverScalarTransport.loci:666-678);
 *   D = -srcJ + rho V / dtau, B = src (:681-690, 592-599).
 * Join order per cell (Loci's is unpinned): the convective face terms
 * (lower, upper, boundary faces), stabilisation, SA sources, diffusion face
 * terms (lower, upper, boundary), the cb2 term. */
void og_sa_point(const og_params *p, double T, double pg, double nut, double vortMag, double y, double vol,
                 double *out3) {
  real sv, sj;
  sa_core(p, T, pg, nut, vortMag, y, vol, &sv, &sj);
  const real rho = (pg + p->Pambient) / (p->Rt * T);
  const real x = rho * nut / suth_mu(p, T);
  out3[0] = nut * sa_fv1(p, x) * rho;
  out3[1] = sv;
  out3[2] = sj;
}

void og_sa_assemble(const og_grid *g, const og_params *p, const double *q, double *src, double *fjp,
                    double *fjm, double *D, double *B, double *tmu) {
  og_sa_assemble_t(g, p, q, NULL, NULL, src, fjp, fjm, D, B, tmu);
}

/* scalar time terms (src/solverScalarTransport.loci): steady X_rhs/X_D
 * (:584-590, :680-690), BDF2 (:602-612, :702-716) with the stabilising
 * Jacobian -1.5 rhodot (:534-539) */
static void scalar_time(const og_grid *g, const og_params *p, const double *q, const double *qn, const double *qm1,
                        int c, int kv, real dtau, real sv, real sJ, double *Dout, double *Bout) {
  const int nc = g->nc;
  const real rho = (q[4 * nc + c] + p->Pambient) / (p->Rt * q[c]);
  if (p->unsteady == 2 && qn) { /* KEC, :592-600 and :691-700 */
    const real rn = (qn[4 * nc + c] + p->Pambient) / (p->Rt * qn[c]);
    const real factor = g->vol[c] / p->dtmax;
    const real Xnp1 = kec_xnp1(q[kv * nc + c], qn[kv * nc + c]);
    const real rhs = factor * ((Xnp1 - qn[kv * nc + c]) * rn) - sv;
    *Bout = -rhs;
    const real sd = (2. * rn) * (g->vol[c] / dtau);
    *Dout = -sJ + sd;
    return;
  }
  if (!(p->unsteady == 1 && qn && qm1)) {
    const real sd = rho * (g->vol[c] / dtau);
    *Dout = -sJ + sd;
    const real rhs = 0.0 - sv;
    *Bout = -rhs;
    return;
  }
  const real rn = (qn[4 * nc + c] + p->Pambient) / (p->Rt * qn[c]);
  const real rnm1 = (qm1[4 * nc + c] + p->Pambient) / (p->Rt * qm1[c]);
  const real X = q[kv * nc + c], Xn = qn[kv * nc + c], Xm = qm1[kv * nc + c];
  const real factor = g->vol[c] / p->dtmax;
  const real w = 0.5;
  const real w1 = (1. + w) * factor, w2 = -w * factor;
  const real rhs = w1 * (X - Xn) * rn + w2 * (Xn * rn - Xm * rnm1) - w2 * X * (rn - rnm1) - sv;
  *Bout = -rhs;
  const real fd = g->vol[c] / dtau;
  const real sd = (1. + w) * fd * rn + w * fd * (rn - rnm1);
  *Dout = -sJ + sd;
}

void og_sa_assemble_t(const og_grid *g, const og_params *p, const double *q, const double *qn, const double *qm1,
                      double *src, double *fjp, double *fjm, double *D, double *B, double *tmu) {
  const int bdf2 = p->unsteady == 1 && qn && qm1, kec = p->unsteady == 2 && qn;
  const int nc = g->nc, nf = g->nf;
  double *qf = calloc((size_t)7 * nf, sizeof(double));
  double *grad = malloc(sizeof(double) * 21 * nc), *phi = malloc(sizeof(double) * 7 * nc);
  double *iflux = malloc(sizeof(double) * 5 * nf), *vflux = malloc(sizeof(double) * 4 * nf);
  double *maxev = malloc(sizeof(double) * nf), *dtau = malloc(sizeof(double) * nc);
  double *fsrc = malloc(sizeof(double) * 5 * nc);
  og_face_states(g, p, q, qf);
  og_gradients(g, p, q, qf, grad, phi);
  og_fluxes(g, p, q, qf, grad, phi, iflux, vflux, maxev);
  /* dtau of the fluid system (same cell time step) */
  {
    real_fj *a = malloc(sizeof(real_fj) * 25 * (size_t)nf), *b = malloc(sizeof(real_fj) * 25 * (size_t)nf);
    real_fj *d = malloc(sizeof(real_fj) * 25 * (size_t)nc), *bb = malloc(sizeof(real_fj) * 5 * (size_t)nc);
    int *pv = malloc(sizeof(int) * 5 * (size_t)nc);
    og_assemble2(g, p, q, fsrc, dtau, a, b, d, pv, bb, NULL, NULL);
    free(a); free(b); free(d); free(bb); free(pv);
  }
  og_adj a;
  build_adj(g, &a);
  const real sig = p->sa_sigma, cb2 = p->sa_cb2;
  const real factorx = 2, factory = .5;
  /* per-face quantities */
  double *fconv = calloc(nf, sizeof(double)), *fdiff = calloc(nf, sizeof(double)), *fdiv = calloc(nf, sizeof(double));
  double *sJL = calloc(nf, sizeof(double)), *sJU = calloc(nf, sizeof(double));
  double *bJ = calloc(nf, sizeof(double)); /* boundary srcJ contribution to ci */
  double *cJl = calloc(nf, sizeof(double)), *cJr = calloc(nf, sizeof(double)); /* cb2 diagonal parts */
  for (int f = 0; f < nf; ++f) {
    const int t = OG_BT(g, f), c = g->cl[f];
    real an[3], xf[3], xl[3];
    fnormal(g, f, an);
    fcenter(g, f, xf);
    ccenter(g, c, xl);
    const real area = g->area[f];
    const real mdot = iflux[f];
    real gfx[3];
    face_grad(g, q, qf, grad, f, 5, gfx);
    real rho_f, mu_f;
    face_rho_mu(g, p, q, qf, f, &rho_f, &mu_f);
    const real rhol = (q[4 * nc + c] + p->Pambient) / (p->Rt * q[c]);
    const real nul = q[5 * nc + c];
    fjp[f] = fjm[f] = 0.0;
    if (t == OG_INTERIOR || t == OG_PERIODIC) {
      real xr[3];
      const int o = rcell(g, f, xr);
      const real XL = recon_k(g, q, grad, phi, 5, c, xl, xf), XR = recon_k(g, q, grad, phi, 5, o, xr, xf);
      const real Xval = (mdot < 0.0) ? XR : XL;
      fconv[f] = mdot * maxr(Xval, 0.0);
      const real vl = g->vol[c], vr = g->vol[o];
      const real nur = q[5 * nc + o];
      const real vrvl = 1.0 / (vr + vl);
      const real nuf = (vr * nul + vl * nur) * vrvl;
      real d1 = dot3(gfx, an) * area;
      fdiff[f] = (d1 * (mu_f / rho_f + (1. + cb2) * nuf)) / (sig);
      fdiv[f] = dot3(gfx, an) * area;
      /* Jacobians: convection, diffusion, cb2 */
      fjp[f] += maxr(mdot, 0.0);
      fjm[f] += minr(mdot, 0.0);
      const real rhor = (q[4 * nc + o] + p->Pambient) / (p->Rt * q[o]);
      const real rhom = .5 * (rhol + rhor);
      const real C = (mu_f + (1. + cb2) * rhom * nuf) * area;
      real dv[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
      const real dd = dot3(dv, an);
      fjp[f] -= -C / (dd * sig);
      fjm[f] += -C / (dd * sig);
      const real C2 = cb2 * area;
      const real d2 = sig * dd;
      sJU[f] = -factorx * (C2 / d2) * rhol * nul;
      sJL[f] = -factorx * (C2 / d2) * rhor * nur;
      cJl[f] = factorx * factory * (C2 / d2) * rhol * nul;
      cJr[f] = factorx * factory * (C2 / d2) * rhor * nur;
      fjp[f] += sJL[f];
      fjm[f] -= sJU[f];
    } else {
      const real XL = recon_k(g, q, grad, phi, 5, c, xl, xf);
      const real Xf = qf[5 * nf + f];
      const real Xval = (mdot < 0.0) ? Xf : XL;
      fconv[f] = mdot * maxr(Xval, 0.0);
      real d1 = dot3(gfx, an) * area;
      fdiff[f] = (d1 * (mu_f / rho_f + (1. + cb2) * Xf)) / (sig);
      fdiv[f] = dot3(gfx, an) * area;
      real bj = 0.0;
      bj -= maxr(mdot, 0.0);
      if (t != OG_SYMMETRY) {
        const real C = (mu_f + (1. + cb2) * rho_f * Xf) * area;
        real dv[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
        const real dd = dot3(dv, an);
        bj += -C / (dd * sig);
        const real C2 = cb2 * area;
        const real d2 = sig * dd;
        bj += factorx * factory * (C2 / d2) * rhol * nul;
      }
      bJ[f] = bj;
    }
  }
  for (int c = 0; c < nc; ++c) {
    const real rho = (q[4 * nc + c] + p->Pambient) / (p->Rt * q[c]);
    const real X = q[5 * nc + c];
    real sv = 0.0, sJ = 0.0, rhodot = 0.0;
    /* convection and continuity */
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
      const int f = a.lo[j];
      sv += fconv[f];
      rhodot += iflux[f];
    }
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
      const int f = a.up[j];
      sv -= fconv[f];
      rhodot -= iflux[f];
    }
    for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) {
      const int f = a.bd[j];
      sv -= fconv[f];
      rhodot -= iflux[f];
    }
    /* stabilisation (src/solverScalarTransport.loci:508-546) */
    sv -= (kec ? kec_xnp1(X, qn[5 * nc + c]) : X) * rhodot;
    sJ += kec ? -2. * rhodot : bdf2 ? -1.5 * rhodot : -rhodot;
    /* SA sources */
    real ssrc, sdJ;
    sa_cell_terms(g, p, q, grad, c, &ssrc, &sdJ);
    sv += ssrc;
    sJ += sdJ;
    /* diffusion and its divergence */
    real div = 0.0;
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
      const int f = a.lo[j];
      sv -= rho * fdiff[f];
      div -= fdiv[f];
    }
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
      const int f = a.up[j];
      sv += rho * fdiff[f];
      div += fdiv[f];
    }
    for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) {
      const int f = a.bd[j];
      sv += rho * fdiff[f];
      div += fdiv[f];
    }
    sv += -rho * X * div * cb2 / sig;
    /* srcJ: diagonal cb2 parts, boundary terms, flux Jacobians (upper -fjp,
     * lower +fjm), then the L/U cb2 parts taken back out of the diagonal */
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) sJ += cJr[a.lo[j]];
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) sJ += cJl[a.up[j]];
    for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) sJ += bJ[a.bd[j]];
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) sJ -= fjp[a.up[j]];
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) sJ += fjm[a.lo[j]];
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) sJ += sJL[a.up[j]];
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) sJ += sJU[a.lo[j]];
    src[c] = sv;
    scalar_time(g, p, q, qn, qm1, c, 5, dtau[c], sv, sJ, &D[c], &B[c]);
    if (tmu) tmu[c] = tmu_cell(g, p, q, grad, c);
  }
  /* L = -fjp, U = fjm are stored by the caller's convention: fjp/fjm */
  free(fconv); free(fdiff); free(fdiv); free(sJL); free(sJU); free(bJ); free(cJl); free(cJr);
  free_adj(&a);
  free(qf); free(grad); free(phi); free(iflux); free(vflux); free(maxev); free(dtau); free(fsrc);
}

/* Scalar SGS, src/linearSolverSGS.loci:333-500: Dinv = 1/D, X = Dinv B, then
 * symmetric sweeps t = -B + sum_upper X[cr] U + sum_lower X[cl] L,
 * X = Dinv (-t), periodic ghosts refreshed before each half-sweep. */
/* ------------------------------------------------------------------ */
/* Menter SST k-omega (turbulence/KOmega/sst.loci) with the default       */
/* options: turbulence_model SST, kprod_func SOmega,                     */
/* compressibilityCorrection wilcox, dissipationCorrection none, smooth  */
/* walls, Kprod_clip 1e100, tmu_clip 1e5; k and omega segregated, each    */
/* through the generic scalar transport of src/solverScalarTransport.loci */
/* ------------------------------------------------------------------ */

/* sst1 (k-omega set) and sst2 (k-epsilon set), sst.loci:319-347; sigma_k of
 * the SST model 0.85 (:682-684) */
#define SST_SIGK1 0.85
#define SST_SIGW1 0.5
#define SST_BETA1 0.075
#define SST_BETAS 0.09
#define SST_KAPPA 0.41
#define SST_A1 0.31
#define SST_SIGK2 1.0
#define SST_SIGW2 0.856
#define SST_BETA2 0.0828

/* strainRate_t, divu_t (sst.loci:65-104) and vortMag_t
 * (src/solverTurbulence.loci:45-49,77-79) of gu[3i+j] = d u_i / d x_j */
static void sst_kinematics(const real *gu, real *S, real *divu, real *Om) {
  const real dudx = gu[0], dudy = gu[1], dudz = gu[2];
  const real dvdx = gu[3], dvdy = gu[4], dvdz = gu[5];
  const real dwdx = gu[6], dwdy = gu[7], dwdz = gu[8];
  *divu = (dudx + dvdy + dwdz);
  const real S2 = 2. * (dudx * dudx + dvdy * dvdy + dwdz * dwdz) + (dudy + dvdx) * (dudy + dvdx) +
                  (dudz + dwdx) * (dudz + dwdx) + (dvdz + dwdy) * (dvdz + dwdy);
  *S = sqrt(S2);
  const real wx = dwdy - dvdz, wy = dudz - dwdx, wz = dvdx - dudy;
  *Om = sqrt(wx * wx + wy * wy + wz * wz);
}

/* bF2 (sst.loci:514-521) and kw_tmu with bF3 = 1 (:768-771) */
static real sst_F2(real k, real w, real y, real nu) {
  const real arg2 = minr(maxr(2.0 * sqrt(k) / (0.09 * w * y), 500.0 * nu / (w * y * y)), 1000.0);
  return tanh(arg2 * arg2);
}
static real sst_tmu(real rho, real k, real w, real S, real F2) {
  return minr(rho * SST_A1 * k / maxr(1e-30, maxr(SST_A1 * w, S * F2 * 1.0)), 1e5);
}

static void cell_gu(const og_grid *g, const double *grad, int c, real *gu) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gu[3 * i + j] = grad[(3 * (1 + i) + j) * g->nc + c];
}

static real sst_tmu_cell(const og_grid *g, const og_params *p, const double *q, const double *grad, int c) {
  const int nc = g->nc;
  const real T = q[c], pg = q[4 * nc + c], k = q[5 * nc + c], w = q[6 * nc + c];
  const real rho = (pg + p->Pambient) / (p->Rt * T);
  const real nu = suth_mu(p, T) / rho;
  real gu[9], S, divu, Om;
  cell_gu(g, grad, c, gu);
  sst_kinematics(gu, &S, &divu, &Om);
  return sst_tmu(rho, k, w, S, sst_F2(k, w, g->dist[c], nu));
}

typedef struct {
  real F1, F2, tmu, ksrc, wsrc, kJ, wJ, Pk, CD, sigk, sigw;
} sst_cell_t;

/* One cell: crossDiffusionTerm (:464-468) from grad k . grad omega, bF1
 * (:487-499), bF2, kw_tmu, the blended coefficients (:702-711), the SOmega
 * stress (:790-796), P_k (:851-854), the k and omega sources and
 * destruction Jacobians (:870-904), cross diffusion (:910-913) and the
 * Wilcox compressibility correction (:283-287, :960-1006).  Each source
 * keeps the file's join order. */
static sst_cell_t sst_core(const og_params *p, real T, real pg, real k, real w, const real *gu, real y, real gkgw,
                           real vol) {
  sst_cell_t o;
  const real rho = (pg + p->Pambient) / (p->Rt * T);
  const real mu = suth_mu(p, T);
  const real nu = mu / rho;
  real S, divu, Om;
  sst_kinematics(gu, &S, &divu, &Om);
  const real crossterm = maxr(SST_SIGW2 * gkgw, 0.0);
  const real CD = 2. * rho * crossterm / maxr(w, 1e-30);
  const real CDkw = maxr(1.0e-10, CD);
  const real ls = maxr(sqrt(k) / (0.09 * w * y), 500.0 * nu / (w * y * y));
  const real arg1 = minr(minr(ls, 4. * rho * SST_SIGW2 * k / (CDkw * y * y)), 32.0);
  const real arg1_2 = arg1 * arg1, arg1_4 = arg1_2 * arg1_2;
  const real F1 = tanh(arg1_4);
  const real F2 = sst_F2(k, w, y, nu);
  const real tmu = sst_tmu(rho, k, w, S, F2);
  const real sigk = F1 * SST_SIGK1 + (1.0 - F1) * SST_SIGK2;
  const real sigw = F1 * SST_SIGW1 + (1.0 - F1) * SST_SIGW2;
  const real beta = (F1 * SST_BETA1 + (1.0 - F1) * SST_BETA2);
  const real gama1 = SST_BETA1 / SST_BETAS - SST_SIGW1 * pow(SST_KAPPA, 2.0) / sqrt(SST_BETAS);
  const real gama2 = SST_BETA2 / SST_BETAS - SST_SIGW2 * pow(SST_KAPPA, 2.0) / sqrt(SST_BETAS);
  const real gama = F1 * gama1 + (1.0 - F1) * gama2;
  const real stress = maxr(tmu * (S * Om - 2. / 3. * divu * divu) - 2. / 3. * rho * k * divu, 0.0);
  const real BetaS = SST_BETAS * 1.0;
  const real Pk = minr(stress, (1e100) * rho * BetaS * k * w);
  const real a = sqrt(p->gamma * p->Rt * T);
  const real Mt2 = 2.0 * k / (a * a);
  const real mt0 = 0.25;
  const real fmt = maxr(Mt2 - mt0 * mt0, 0.0);
  const real BetaF = SST_BETAS * 1.0 * 1.5 * fmt;
  const real nut = tmu / rho;
  const real Dw = (gama / (nut + 1e-30)) * stress;
  real ks = 0.0, ws = 0.0;
  ks += Pk * vol;
  ws += Dw * vol;
  ks -= BetaS * rho * w * k * vol;
  ws -= beta * rho * w * w * vol;
  ws += (1.0 - F1) * CD * vol;
  ks -= BetaF * rho * w * k * vol;
  ws -= -BetaF * rho * w * w * vol * 1.0;
  o.F1 = F1; o.F2 = F2; o.tmu = tmu; o.ksrc = ks; o.wsrc = ws;
  o.kJ = -SST_BETAS * 1.0 * rho * w * vol;
  o.wJ = -2.0 * beta * rho * w * vol;
  o.Pk = Pk; o.CD = CD; o.sigk = sigk; o.sigw = sigw;
  return o;
}

void og_sst_point(const og_params *p, double T, double pg, double k, double w, const double *gu, double y,
                  double gkgw, double vol, double *out) {
  real g9[9];
  for (int i = 0; i < 9; ++i) g9[i] = gu[i];
  const sst_cell_t o = sst_core(p, T, pg, k, w, g9, y, gkgw, vol);
  out[0] = o.F1; out[1] = o.F2; out[2] = o.tmu; out[3] = o.ksrc; out[4] = o.wsrc;
  out[5] = o.kJ; out[6] = o.wJ; out[7] = o.Pk; out[8] = o.CD;
}

/* The k and omega systems of one steady Newton step.  Per scalar X the
 * generic rules of src/solverScalarTransport.loci: upwind convection of
 * max(X, 0) (:121-135, 493-503), the continuity stabilisation -X rhodot
 * and its Jacobian (:66-103, 505-546), convective Jacobians (:548-567),
 * diffusion X_nu_f grad_f(X).n area with its Jacobians C/d (:779-826),
 * X_srcJ from upper fjp / lower fjm (:666-678), D = -srcJ + rho V/dtau
 * (:680-690), B = src.  X_nu_f = mu_f + tmu_f sigma_f with sigma_f the
 * volume-weighted blend of the two cells' sigma (sst.loci:725-737,
 * 816-836).  grad k . grad omega = (div(w grad k) - w lap k)/vol with w_f =
 * max(0, mean of the reconstructions) on interior faces and the face value
 * on boundary faces (:427-462).
 * Join order per cell (Loci's is unpinned): convection (lower, upper,
 * boundary), stabilisation, cell sources, diffusion (lower, upper,
 * boundary); Jacobian: stabilisation, cell, boundary faces, upper fjp,
 * lower fjm.  Periodic faces act as interior faces. */
void og_sst_assemble(const og_grid *g, const og_params *p, const double *q, double *src, double *fjp,
                     double *fjm, double *D, double *B, double *tmu, double *aux) {
  og_sst_assemble_t(g, p, q, NULL, NULL, src, fjp, fjm, D, B, tmu, aux);
}

void og_sst_assemble_t(const og_grid *g, const og_params *p, const double *q, const double *qn,
                       const double *qm1, double *src, double *fjp, double *fjm, double *D, double *B, double *tmu,
                       double *aux) {
  const int bdf2 = p->unsteady == 1 && qn && qm1, kec = p->unsteady == 2 && qn;
  const int nc = g->nc, nf = g->nf;
  double *qf = calloc((size_t)7 * nf, sizeof(double));
  double *grad = malloc(sizeof(double) * 21 * nc), *phi = malloc(sizeof(double) * 7 * nc);
  double *iflux = malloc(sizeof(double) * 5 * nf), *vflux = malloc(sizeof(double) * 4 * nf);
  double *maxev = malloc(sizeof(double) * nf), *dtau = malloc(sizeof(double) * nc);
  double *fsrc = malloc(sizeof(double) * 5 * nc);
  og_face_states(g, p, q, qf);
  og_gradients(g, p, q, qf, grad, phi);
  og_fluxes(g, p, q, qf, grad, phi, iflux, vflux, maxev);
  {
    real_fj *a = malloc(sizeof(real_fj) * 25 * (size_t)nf), *b = malloc(sizeof(real_fj) * 25 * (size_t)nf);
    real_fj *d = malloc(sizeof(real_fj) * 25 * (size_t)nc), *bb = malloc(sizeof(real_fj) * 5 * (size_t)nc);
    int *pv = malloc(sizeof(int) * 5 * (size_t)nc);
    og_assemble2(g, p, q, fsrc, dtau, a, b, d, pv, bb, NULL, NULL);
    free(a); free(b); free(d); free(bb); free(pv);
  }
  og_adj a;
  build_adj(g, &a);
  /* face pass 1: area grad_f(k).n and its omega-weighted form */
  double *gkn = calloc(nf, sizeof(double)), *wgkn = calloc(nf, sizeof(double));
  for (int f = 0; f < nf; ++f) {
    const int t = OG_BT(g, f), c = g->cl[f];
    real an[3], gk[3];
    fnormal(g, f, an);
    face_grad(g, q, qf, grad, f, 5, gk);
    const real gradkdotn = g->area[f] * dot3(gk, an);
    gkn[f] = gradkdotn;
    if (t == OG_INTERIOR || t == OG_PERIODIC) {
      real xf[3], xl[3], xr[3];
      fcenter(g, f, xf);
      ccenter(g, c, xl);
      const int o = rcell(g, f, xr);
      const real wL = recon_k(g, q, grad, phi, 6, c, xl, xf), wR = recon_k(g, q, grad, phi, 6, o, xr, xf);
      const real w_f = maxr(0.0, 0.5 * (wL + wR));
      wgkn[f] = w_f * gradkdotn;
    } else {
      wgkn[f] = qf[6 * nf + f] * gradkdotn;
    }
  }
  /* cell pass 1: cross diffusion, blending, eddy viscosity, sources */
  sst_cell_t *cs = malloc(sizeof(sst_cell_t) * nc);
  for (int c = 0; c < nc; ++c) {
    real lap = 0.0, dwg = 0.0;
    for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
      const int f = a.lo[j];
      lap += -gkn[f];
      dwg += -wgkn[f];
    }
    for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
      const int f = a.up[j];
      lap += gkn[f];
      dwg += wgkn[f];
    }
    for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) {
      const int f = a.bd[j];
      lap += gkn[f];
      dwg += wgkn[f];
    }
    const real w = q[6 * nc + c];
    const real gkgw = (dwg - w * lap) / g->vol[c];
    real gu[9];
    cell_gu(g, grad, c, gu);
    cs[c] = sst_core(p, q[c], q[4 * nc + c], q[5 * nc + c], w, gu, g->dist[c], gkgw, g->vol[c]);
    if (tmu) tmu[c] = cs[c].tmu;
    if (aux) {
      aux[c] = cs[c].F1; aux[nc + c] = cs[c].F2; aux[2 * nc + c] = gkgw; aux[3 * nc + c] = cs[c].CD;
    }
  }
  /* per scalar: face terms, then the cell gather */
  double *fconv = malloc(sizeof(double) * nf), *fdiff = malloc(sizeof(double) * nf), *bJ = calloc(nf, sizeof(double));
  for (int sc = 0; sc < 2; ++sc) {
    const int kv = 5 + sc;
    double *Fp = fjp + (size_t)sc * nf, *Fm = fjm + (size_t)sc * nf;
    for (int f = 0; f < nf; ++f) {
      const int t = OG_BT(g, f), c = g->cl[f];
      real an[3], xf[3], xl[3], gx[3];
      fnormal(g, f, an);
      fcenter(g, f, xf);
      ccenter(g, c, xl);
      const real area = g->area[f];
      const real mdot = iflux[f];
      face_grad(g, q, qf, grad, f, kv, gx);
      real rho_f, mu_f;
      face_rho_mu(g, p, q, qf, f, &rho_f, &mu_f);
      const real tmf = tmu_face(g, p, q, grad, f);
      const real sl = sc == 0 ? cs[c].sigk : cs[c].sigw;
      Fp[f] = Fm[f] = 0.0;
      bJ[f] = 0.0;
      if (t == OG_INTERIOR || t == OG_PERIODIC) {
        real xr[3];
        const int o = rcell(g, f, xr);
        const real XL = recon_k(g, q, grad, phi, kv, c, xl, xf), XR = recon_k(g, q, grad, phi, kv, o, xr, xf);
        const real Xval = (mdot < 0.0) ? XR : XL;
        fconv[f] = mdot * maxr(Xval, 0.0);
        const real vl = g->vol[c], vr = g->vol[o];
        const real sr = sc == 0 ? cs[o].sigk : cs[o].sigw;
        const real rvlvr = 1. / (vl + vr);
        const real sig_f = (vr * sl + vl * sr) * rvlvr;
        const real nu_f = (mu_f + tmf * sig_f);
        fdiff[f] = nu_f * dot3(gx, an) * area;
        const real C = nu_f * area;
        real dv[3] = {xr[0] - xl[0], xr[1] - xl[1], xr[2] - xl[2]};
        const real d = dot3(dv, an);
        Fp[f] += maxr(mdot, 0.0);
        Fm[f] += minr(mdot, 0.0);
        Fp[f] += C / d;
        Fm[f] += -C / d;
      } else {
        const real XL = recon_k(g, q, grad, phi, kv, c, xl, xf);
        const real Xf = qf[kv * nf + f];
        const real Xval = (mdot < 0.0) ? Xf : XL;
        fconv[f] = mdot * maxr(Xval, 0.0);
        const real nu_f = (mu_f + tmf * sl);
        fdiff[f] = nu_f * dot3(gx, an) * area;
        const real C = nu_f * area;
        real dv[3] = {xf[0] - xl[0], xf[1] - xl[1], xf[2] - xl[2]};
        const real d = dot3(dv, an);
        bJ[f] -= maxr(mdot, 0.0);
        bJ[f] += -C / d;
      }
    }
    for (int c = 0; c < nc; ++c) {
      const real X = q[kv * nc + c];
      real sv = 0.0, sJ = 0.0, rhodot = 0.0;
      for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
        const int f = a.lo[j];
        sv += fconv[f];
        rhodot += iflux[f];
      }
      for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
        const int f = a.up[j];
        sv -= fconv[f];
        rhodot -= iflux[f];
      }
      for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) {
        const int f = a.bd[j];
        sv -= fconv[f];
        rhodot -= iflux[f];
      }
      sv -= (kec ? kec_xnp1(X, qn[kv * nc + c]) : X) * rhodot;
      sJ += kec ? -2. * rhodot : bdf2 ? -1.5 * rhodot : -rhodot;
      sv += sc == 0 ? cs[c].ksrc : cs[c].wsrc;
      sJ += sc == 0 ? cs[c].kJ : cs[c].wJ;
      for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) sv -= fdiff[a.lo[j]];
      for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) sv += fdiff[a.up[j]];
      for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) sv += fdiff[a.bd[j]];
      for (int j = a.bd_off[c]; j < a.bd_off[c + 1]; ++j) sJ += bJ[a.bd[j]];
      for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) sJ -= Fp[a.up[j]];
      for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) sJ += Fm[a.lo[j]];
      const size_t i = (size_t)sc * nc + c;
      src[i] = sv;
      scalar_time(g, p, q, qn, qm1, c, kv, dtau[c], sv, sJ, &D[i], &B[i]);
    }
  }
  free(gkn); free(wgkn); free(cs); free(fconv); free(fdiff); free(bJ);
  free_adj(&a);
  free(qf); free(grad); free(phi); free(iflux); free(vflux); free(maxev); free(dtau); free(fsrc);
}

void og_sa_sgs(const og_grid *g, int nsweeps, const double *fjp, const double *fjm, const double *D,
               const double *B, double *X) {
  const int nc = g->nc, nf = g->nf;
  og_adj a;
  build_adj(g, &a);
  double *Dinv = malloc(sizeof(double) * nc), *Xg = calloc(nf, sizeof(double));
  for (int c = 0; c < nc; ++c) Dinv[c] = 1. / D[c];
  for (int c = 0; c < nc; ++c) X[c] = Dinv[c] * B[c];
  for (int it = 0; it < nsweeps; ++it)
    for (int pass = 0; pass < 2; ++pass) {
      for (int f = 0; f < nf; ++f)
        if (OG_BT(g, f) == OG_PERIODIC) Xg[f] = X[g->cl[g->pface[f]]];
      for (int ii = 0; ii < nc; ++ii) {
        const int c = pass == 0 ? ii : nc - 1 - ii;
        real t = -B[c];
        for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
          const int f = a.up[j];
          const real xn = OG_BT(g, f) == OG_PERIODIC ? Xg[f] : X[g->cr[f]];
          t += xn * fjm[f];
        }
        for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
          const int f = a.lo[j];
          t += X[g->cl[f]] * (-fjp[f]);
        }
        const real v = -t;
        X[c] = Dinv[c] * v;
      }
    }
  free(Dinv); free(Xg);
  free_adj(&a);
}

/* Scalar LSGS (src/linearSolverLSGS.loci:1112-1404) with one-cell lines in
 * the breadth-first order of og_lsgs_order: inverse diagonal 1/D (:1139),
 * line-Jacobi start X = B/D (:1165-1188), then per iteration the periodic
 * copies, a forward sweep that keeps each cell's value before it in xt
 * (:1240-1279), a backward sweep that does not (:1282-1317), the relaxation
 * X = w X + (1-w) xt with the update norm, relaxation control and stopping
 * tests (:1320-1346).  The row is B minus lower then upper couplings. */
int og_sa_lsgs(const og_grid *g, int maxiter, double abstol, double reltol, double w0, const double *fjp,
               const double *fjm, const double *D, const double *B, double *X, double *w_out,
               double *resid_out) {
  const int nc = g->nc, nf = g->nf;
  og_adj a;
  build_adj(g, &a);
  int *order = malloc(sizeof(int) * nc);
  og_lsgs_order(g, order);
  double *Dinv = malloc(sizeof(double) * nc), *Xg = calloc(nf, sizeof(double)), *xt = malloc(sizeof(double) * nc);
  for (int c = 0; c < nc; ++c) Dinv[c] = 1. / D[c];
  for (int c = 0; c < nc; ++c) X[c] = Dinv[c] * B[c];
  double w = w0, zero = 1.0, resid = 0.0;
  int iter = 0;
  for (iter = 0; iter < maxiter; ++iter) {
    for (int f = 0; f < nf; ++f)
      if (OG_BT(g, f) == OG_PERIODIC) Xg[f] = X[g->cl[g->pface[f]]];
    for (int pass = 0; pass < 2; ++pass)
      for (int ii = 0; ii < nc; ++ii) {
        const int c = pass == 0 ? order[ii] : order[nc - 1 - ii];
        double r = B[c];
        for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
          const int f = a.lo[j];
          r -= X[g->cl[f]] * (-fjp[f]);
        }
        for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
          const int f = a.up[j];
          const double xn = OG_BT(g, f) == OG_PERIODIC ? Xg[f] : X[g->cr[f]];
          r -= xn * fjm[f];
        }
        if (pass == 0) xt[c] = X[c];
        X[c] = Dinv[c] * r;
      }
    double residl = 0;
    for (int c = 0; c < nc; ++c) {
      X[c] = w * X[c] + (1. - w) * xt[c];
      const double r = X[c] - xt[c];
      residl += r * r;
    }
    resid = sqrt(residl);
    if (iter == 0) zero = resid > 1e-16 ? resid : 1e-16;
    const double rel = resid / zero;
    if (rel > 1.0) w *= 1. / rel;
    if (resid < abstol) { ++iter; break; }
    if (rel < reltol) { ++iter; break; }
  }
  if (w_out) *w_out = w;
  if (resid_out) *resid_out = resid;
  free(order); free(Dinv); free(Xg); free(xt);
  free_adj(&a);
  return iter;
}

/* FSGSScalarSolve (src/linearSolverFSGS.loci:998-1200) with one-cell lines:
 * 1/D (:1019-1031), X starting at zero (:1049), nsweeps symmetric sweeps in
 * the breadth-first order with periodic copies (:1066-1106), X = (B - lower
 * - upper couplings)/D; no relaxation, no stopping test. */
void og_sa_fsgs(const og_grid *g, int nsweeps, const double *fjp, const double *fjm, const double *D,
                const double *B, double *X) {
  const int nc = g->nc, nf = g->nf;
  og_adj a;
  build_adj(g, &a);
  int *order = malloc(sizeof(int) * nc);
  og_lsgs_order(g, order);
  double *Dinv = malloc(sizeof(double) * nc), *Xg = calloc(nf, sizeof(double));
  for (int c = 0; c < nc; ++c) Dinv[c] = 1. / D[c];
  for (int c = 0; c < nc; ++c) X[c] = 0.0;
  for (int it = 0; it < nsweeps; ++it) {
    for (int f = 0; f < nf; ++f)
      if (OG_BT(g, f) == OG_PERIODIC) Xg[f] = X[g->cl[g->pface[f]]];
    for (int pass = 0; pass < 2; ++pass)
      for (int ii = 0; ii < nc; ++ii) {
        const int c = pass == 0 ? order[ii] : order[nc - 1 - ii];
        double r = B[c];
        for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
          const int f = a.lo[j];
          r -= X[g->cl[f]] * (-fjp[f]);
        }
        for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
          const int f = a.up[j];
          const double xn = OG_BT(g, f) == OG_PERIODIC ? Xg[f] : X[g->cr[f]];
          r -= xn * fjm[f];
        }
        X[c] = Dinv[c] * r;
      }
  }
  free(order); free(Dinv); free(Xg);
  free_adj(&a);
}

/* ------------------------------------------------------------------ */
/* line SGS (src/linearSolverLSGS.loci)                                */
/* ------------------------------------------------------------------ */

/* Equation order of the LSGS topology (:361-1007).  In this revision the
 * line search only links cells through duplicate couples, which are removed
 * just before it (:625-632), so every line is a single cell and the order is
 * the breadth-first order of the lines (:902-937): start at cell 0, visit the
 * interior-face neighbours layer by layer, each layer in ascending order. */
int og_lsgs_order(const og_grid *g, int *order) {
  int nc = g->nc;
  og_adj a;
  build_adj(g, &a);
  char *vis = calloc(nc, 1);
  int *cur = malloc(sizeof(int) * nc), *nxt = malloc(sizeof(int) * nc);
  int ncur = 0, nout = 0, sp = 0;
  while (nout < nc) {
    if (ncur == 0) {
      while (sp < nc && vis[sp]) sp++;
      if (sp == nc) break;
      cur[ncur++] = sp;
      vis[sp] = 1;
    }
    int nn = 0;
    for (int t = 0; t < ncur; ++t) {
      int c = cur[t];
      for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
        int o = g->cl[a.lo[j]];
        if (!vis[o]) { vis[o] = 1; nxt[nn++] = o; }
      }
      for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
        int f = a.up[j];
        if (OG_BT(g, f) != OG_INTERIOR) continue; /* ghosts do not join the search */
        int o = g->cr[f];
        if (!vis[o]) { vis[o] = 1; nxt[nn++] = o; }
      }
      order[nout++] = c;
    }
    /* the next layer is a std::set: ascending */
    for (int i = 1; i < nn; ++i) {
      int v = nxt[i], k = i - 1;
      while (k >= 0 && nxt[k] > v) { nxt[k + 1] = nxt[k]; --k; }
      nxt[k + 1] = v;
    }
    memcpy(cur, nxt, sizeof(int) * nn);
    ncur = nn;
  }
  free(vis); free(cur); free(nxt);
  free_adj(&a);
  return nout;
}

/* unpivoted Doolittle LU in place (Loci Mat decompose_lu) and its solve */
static void lu_nopiv(real *a, int nb) {
  for (int k = 0; k < nb; ++k)
    for (int i = k + 1; i < nb; ++i) {
      real l = a[i + nb * k] / a[k + nb * k];
      a[i + nb * k] = l;
      for (int j = k + 1; j < nb; ++j) a[i + nb * j] = a[i + nb * j] - l * a[k + nb * j];
    }
}
static void lu_nopiv_solve(const real *a, const real *b, real *x, int nb) {
  for (int i = 0; i < nb; ++i) {
    real tmp = b[i];
    for (int j = 0; j < i; ++j) tmp -= a[i + nb * j] * x[j];
    x[i] = tmp;
  }
  for (int i = nb - 1; i >= 0; --i) {
    real tmp = x[i];
    for (int j = i + 1; j < nb; ++j) tmp -= a[i + nb * j] * x[j];
    x[i] = tmp / a[i + nb * i];
  }
}
/* the same factorisation and solve with the factors in real_fj (FSGS keeps
 * them in the Jacobian precision, src/linearSolverFSGS.loci:1299-1308 and
 * solve_lu_N<N, real_fj, real, real> :73-113) */
static void lu_nopiv_fj(real_fj *a, int nb) {
  for (int k = 0; k < nb; ++k)
    for (int i = k + 1; i < nb; ++i) {
      real_fj l = a[i + nb * k] / a[k + nb * k];
      a[i + nb * k] = l;
      for (int j = k + 1; j < nb; ++j) a[i + nb * j] = a[i + nb * j] - l * a[k + nb * j];
    }
}
static void lu_nopiv_solve_fj(const real_fj *a, const real *b, real *x, int nb) {
  for (int i = 0; i < nb; ++i) {
    real tmp = b[i];
    for (int j = 0; j < i; ++j) tmp -= a[i + nb * j] * x[j];
    x[i] = tmp;
  }
  for (int i = nb - 1; i >= 0; --i) {
    real tmp = x[i];
    for (int j = i + 1; j < nb; ++j) tmp -= a[i + nb * j] * x[j];
    x[i] = tmp / a[i + nb * i];
  }
}

/* t[i] += sum_j M(i,j) v[j] for a double block */
static inline void dotprod_accum_d(const real *M, const real *v, real *t) {
  for (int j = 0; j < NB; ++j) {
    const real in = v[j];
    for (int i = 0; i < NB; ++i) t[i] += M[i + NB * j] * in;
  }
}

/* LSGSBlockedSSolve (:1420-1495 matrix and diagonal, :1692-1954 solve):
 * D copied to double and LU-factored without pivots, line-Jacobi start,
 * then up to maxiter relaxed symmetric sweeps in line order; per cell
 * t = -B + sum_lower L X[cl] + sum_upper U X[cr] (lower faces first), the
 * periodic ghost values refreshed once per iteration; after both sweeps
 * X = w X + (1-w) X_forward, stopping on the update norm.  Returns the
 * iterations run; *w_out and *resid_out receive the final relaxation and
 * update norm. */
int og_lsgs(const og_grid *g, int maxiter, double abstol, double reltol, double w0, const void *fjp_v,
            const void *fjm_v, const void *Draw_v, const void *B_v, double *X, double *w_out,
            double *resid_out) {
  int nc = g->nc, nf = g->nf;
  const real_fj *fjp = fjp_v, *fjm = fjm_v, *Draw = Draw_v, *B = B_v;
  og_adj a;
  build_adj(g, &a);
  int *order = malloc(sizeof(int) * nc);
  og_lsgs_order(g, order);
  real *Dl = malloc(sizeof(real) * 25 * (size_t)nc);
  for (size_t i = 0; i < 25 * (size_t)nc; ++i) Dl[i] = (real)Draw[i];
  for (int c = 0; c < nc; ++c) lu_nopiv(Dl + 25 * (size_t)c, 5);
  double *Xg = calloc((size_t)5 * nf, sizeof(double));
  double *xt = malloc(sizeof(double) * 5 * (size_t)nc);
  /* line-Jacobi start */
  for (int c = 0; c < nc; ++c) {
    real b[5];
    for (int k = 0; k < 5; ++k) b[k] = B[5 * (size_t)c + k];
    lu_nopiv_solve(Dl + 25 * (size_t)c, b, X + 5 * (size_t)c, 5);
  }
  real w = w0, zero = 1.0, resid = 0.0;
  int iter = 0;
  for (iter = 0; iter < maxiter; ++iter) {
    for (int f = 0; f < nf; ++f)
      if (OG_BT(g, f) == OG_PERIODIC)
        for (int k = 0; k < 5; ++k) Xg[5 * (size_t)f + k] = X[5 * (size_t)g->cl[g->pface[f]] + k];
    for (int pass = 0; pass < 2; ++pass) {
      for (int ii = 0; ii < nc; ++ii) {
        int c = pass == 0 ? order[ii] : order[nc - 1 - ii];
        real t[5], v[5];
        for (int k = 0; k < 5; ++k) t[k] = -(real)B[5 * (size_t)c + k];
        for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
          int f = a.lo[j];
          real_fj L[25];
          for (int i = 0; i < 25; ++i) L[i] = -fjp[25 * (size_t)f + i];
          dotprod_accum(L, X + 5 * (size_t)g->cl[f], t);
        }
        for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
          int f = a.up[j];
          const real *xn = OG_BT(g, f) == OG_PERIODIC ? Xg + 5 * (size_t)f : X + 5 * (size_t)g->cr[f];
          dotprod_accum(fjm + 25 * (size_t)f, xn, t);
        }
        for (int k = 0; k < 5; ++k) {
          xt[5 * (size_t)c + k] = X[5 * (size_t)c + k];
          v[k] = -t[k];
        }
        lu_nopiv_solve(Dl + 25 * (size_t)c, v, X + 5 * (size_t)c, 5);
      }
    }
    double residl = 0;
    for (size_t i = 0; i < 5 * (size_t)nc; ++i) {
      X[i] = w * X[i] + (1. - w) * xt[i];
      double r = X[i] - xt[i];
      residl += r * r;
    }
    resid = sqrt(residl);
    if (iter == 0) zero = resid > 1e-16 ? resid : 1e-16;
    double rel = resid / zero;
    if (rel > 1.0) w *= 1. / rel;
    if (resid < abstol) { ++iter; break; }
    if (rel < reltol) { ++iter; break; }
  }
  if (w_out) *w_out = w;
  if (resid_out) *resid_out = resid;
  free(order); free(Dl); free(Xg); free(xt);
  free_adj(&a);
  return iter;
}

/* FSGSBlockedSSolve (src/linearSolverFSGS.loci:1228-1708): the equation order
 * of the LSGS topology (:800-845 builds the same one-cell lines in the same
 * breadth-first order), D factored without pivots in real_fj
 * (matrixHelpersNoPivotN::factorDiagonal :1299-1308), X starting at zero
 * (:1608), then nsweeps symmetric sweeps (FSGSMaxIter = gauss_seidel_iter,
 * :413-440) with periodic copies before each (:1620-1660); per cell
 * t = -B + sum of lower then upper couplings, X = D^-1 (-t) (sweep :1310-1356).
 * No relaxation and no stopping test. */
void og_fsgs(const og_grid *g, int nsweeps, const void *fjp_v, const void *fjm_v, const void *Draw_v,
             const void *B_v, double *X) {
  int nc = g->nc, nf = g->nf;
  const real_fj *fjp = fjp_v, *fjm = fjm_v, *Draw = Draw_v, *B = B_v;
  og_adj a;
  build_adj(g, &a);
  int *order = malloc(sizeof(int) * nc);
  og_lsgs_order(g, order);
  real_fj *Dl = malloc(sizeof(real_fj) * 25 * (size_t)nc);
  for (size_t i = 0; i < 25 * (size_t)nc; ++i) Dl[i] = Draw[i];
  for (int c = 0; c < nc; ++c) lu_nopiv_fj(Dl + 25 * (size_t)c, 5);
  double *Xg = calloc((size_t)5 * nf, sizeof(double));
  for (size_t i = 0; i < 5 * (size_t)nc; ++i) X[i] = 0.0;
  for (int it = 0; it < nsweeps; ++it) {
    for (int f = 0; f < nf; ++f)
      if (OG_BT(g, f) == OG_PERIODIC)
        for (int k = 0; k < 5; ++k) Xg[5 * (size_t)f + k] = X[5 * (size_t)g->cl[g->pface[f]] + k];
    for (int pass = 0; pass < 2; ++pass)
      for (int ii = 0; ii < nc; ++ii) {
        int c = pass == 0 ? order[ii] : order[nc - 1 - ii];
        real t[5], v[5];
        for (int k = 0; k < 5; ++k) t[k] = -(real)B[5 * (size_t)c + k];
        for (int j = a.lo_off[c]; j < a.lo_off[c + 1]; ++j) {
          int f = a.lo[j];
          real_fj L[25];
          for (int i = 0; i < 25; ++i) L[i] = -fjp[25 * (size_t)f + i];
          dotprod_accum(L, X + 5 * (size_t)g->cl[f], t);
        }
        for (int j = a.up_off[c]; j < a.up_off[c + 1]; ++j) {
          int f = a.up[j];
          const real *xn = OG_BT(g, f) == OG_PERIODIC ? Xg + 5 * (size_t)f : X + 5 * (size_t)g->cr[f];
          dotprod_accum(fjm + 25 * (size_t)f, xn, t);
        }
        for (int k = 0; k < 5; ++k) v[k] = -t[k];
        lu_nopiv_solve_fj(Dl + 25 * (size_t)c, v, X + 5 * (size_t)c, 5);
      }
  }
  free(order); free(Dl); free(Xg);
  free_adj(&a);
}
