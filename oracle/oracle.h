/*
 * CPU oracle for the flowPsi implicit finite-volume Newton step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or as the timed CPU baseline).  The product library
 * (libfpsi_gpu.so) never links, loads or calls it.
 *
 * It is a plain-C restatement of the flowPsi rule bodies on the Newton-step
 * path (citations are per function in oracle.c).  It is compiled twice:
 *   liboracle_f64.so : real_fj = double  ("plain FP64" Jacobian path)
 *   liboracle_f32.so : real_fj = float   (flowPsi's default, include/flowTypes.h:43-49)
 *
 * Data layout (all arrays structure-of-arrays, reference cell/face order):
 *   q    [5][nc]  primitive state (T, u, v, w, pg); [6][nc] with nuTilde under SA,
 *                 [7][nc] with k and omega under SST
 *   qf   [5][nf]  boundary-face state (T_f, u_f, pg_f); unused on interior faces
 *   grad [15][nc] gradient of q: grad[3*k+d] = d q_k / d x_d
 *   phi  [5][nc]  limiter per variable
 *   flux [5][nf]  iflux;  vflux [4][nf]
 *   blocks are 5x5 column-major (element (i,j) at i + 5*j), as Loci::Mat.
 */
#ifndef FPSI_ORACLE_H
#define FPSI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OG_INTERIOR = 0, OG_FARFIELD = 1, OG_WALL = 2, OG_SYMMETRY = 3, OG_PERIODIC = 4,
       OG_REFLECTING = 5, OG_INFLOW = 6, OG_SUPERSONIC_INFLOW = 7, OG_OUTFLOW = 8,
       OG_SUPERSONIC_OUTFLOW = 9, OG_EXTRAPOLATE = 10, OG_ISENTROPIC_INFLOW = 11, OG_FIXED_MASS = 12 };
/* a boundary face's btype: its type in bits 0-7, its parameter set
 * (og_params.bcs) in bits 8-15 */

typedef struct {
  int nc, nf;
  const int *cl, *cr;     /* cr >= 0 interior; -1 on every boundary face (periodic included) */
  const int *btype;       /* OG_* per face */
  const int *pface;       /* periodic partner face, -1 otherwise */
  const double *n;        /* [3][nf] unit normal, cl -> cr (out of the domain on boundaries) */
  const double *area;     /* [nf] */
  const double *fc;       /* [3][nf] face centre */
  const double *vol;      /* [nc] */
  const double *cc;       /* [3][nc] cell centre */
  const double *pshift;   /* [3][nf] periodic ghost centre = cc[cl[pface]] + pshift */
  const double *dist;     /* [nc] wall distance dist_noslip (turbulent runs only) */
} og_grid;

typedef struct {
  double gamma, Rt, Pambient;
  int flux;               /* 0 = HLLC, 1 = Roe (roeCT_flux) */
  int limiter;            /* 0 = none, 1 = venkatakrishnan, 2 = zero (first order), 3 = barth */
  double Kl;
  int viscous;            /* 0 = inviscid, 1 = laminar */
  double suth[6];         /* a1 a2 a3 k1 k2 k3 */
  double Twall;
  double pg0, T0, u0[3];  /* farfield reference state */
  double cflmax, dtmax, urelax;
  int sgs_iter;
  double minPg;           /* floor of the reconstructed gage pressure */
  int linear_solver;      /* 0 = block SGS, 1 = line SGS, 2 = FSGS */
  int lsgs_maxiter;
  double lsgs_reltol, lsgs_abstol, lsgs_relax;
  double Minf, etaT, PLimPC;  /* low-Mach preconditioning (Minf >= 0.5: eta = 1) */
  int turb;               /* 0 laminar, 1 Spalart-Allmaras (q gains a 6th row nuTilde),
                             2 Menter SST (q gains rows k and omega) */
  double sa_cb1, sa_k, sa_cv1, sa_cb2, sa_cw2, sa_cw3, sa_sigma;
  double Prt;             /* turbulent Prandtl number */
  double nut_bc;          /* nuTilde of inflow faces */
  double k_bc, w_bc;      /* SST: k and omega of inflow faces */
  int unsteady;           /* 0 steady; time-accurate with time step dtmax: 1 BDF2, 2 KEC */
  int newton_iter;        /* newtonMaxIter: Newton iterations per time step */
  double pseudo_cflmax;   /* pseudo-time CFL of the BDF2 preconditioning blend */
  double pbias_control;   /* timePressureBiasControl (KEC) */
  double temporal_damping; /* temporalDamping beta (src/fluxInviscid.loci:1224-1251); <= 0: not set */
  /* boundary parameter sets (src/bcInputs.loci:188-220): T, pg (p - Pambient),
   * u[3], p0, T0, mdot (fixedMass: massFlux = mdot / the set's face area) */
  double bcs[8][8];
} og_params;

typedef struct {
  double rtrms, etrms, mtrms, rtmax, etmax, mtmax;
  double ccmaxr[3], ccmaxe[3];
} og_norms;

/* point functions (one face) */
void og_inviscid_flux(double *f5, double pg, double T, const double *U, const double *an,
                      double area, double pamb, double Rt, double gamma, double us_n);
void og_hllc_flux(double *f5, double pgl, double Tl, const double *Ul, double pgr, double Tr,
                  const double *Ur, const double *an, double area, double pamb, double Rt,
                  double gamma, double us_n, double eta);
void og_roe_flux(double *f5, double pgl, double Tl, const double *Ul, double pgr, double Tr,
                 const double *Ur, const double *an, double area, double pamb, double Rt,
                 double gamma, double us_n, double eta);
/* 5x5 column-major block, real_fj precision of the build; block is accumulated (+=) */
void og_roe_jacobian(void *fj25, double pgl, double Tl, const double *Ul, double pgr, double Tr,
                     const double *Ur, const double *an, double area, double pamb, double Rt,
                     double gamma, double us_n, double eta, int dir);
void og_roe_jacobian_ev(void *fj25, double ev1, double ev2, double ev3, const double *U, double un,
                        double a, double beta, double sigma, double gamma, double Rt, double rho,
                        double T, double h0, const double *n, double area);
void og_cv_inflow(double *pgb, double *Tb, double *ub, double pgi, double Ti, const double *ui,
                  double pg0, double T0, const double *u0, const double *n, double us_n,
                  double pamb, double gamma, double Rt, double beta);
void og_cv_outflow(double *pgb, double *Tb, double *ub, double pgi, double Ti, const double *ui,
                   double pg0, double T0, const double *u0, const double *n, double us_n,
                   double pamb, double gamma, double Rt, double beta);
int  og_lu_decompose(void *a25, int *pivot, int nb);
void og_lu_solve(const void *a25, const int *pivot, const double *b, double *x, int nb);
int  og_real_fj_bytes(void);

/* preconditioning parameter per cell (Eta_p) and per boundary face (Eta_pf) */
void og_eta(const og_grid *g, const og_params *p, const double *q, double *eta_p);
double og_eta_face(const og_params *p, double T, const double *U);

/* pipeline stages (whole grid) */
void og_face_states(const og_grid *g, const og_params *p, const double *q, double *qf);
void og_gradients(const og_grid *g, const og_params *p, const double *q, const double *qf,
                  double *grad, double *phi);
void og_fluxes(const og_grid *g, const og_params *p, const double *q, const double *qf,
               const double *grad, const double *phi, double *iflux, double *vflux,
               double *maxev);
void og_residual(const og_grid *g, const og_params *p, const double *q, double *src,
                 double *iflux, double *vflux, og_norms *nrm);
/* full assembly: dtau, D (LU-factored in place) with pivots, B, and the
 * off-diagonal blocks fjp/fjm per face ([nf][25] each, real_fj). */
void og_assemble(const og_grid *g, const og_params *p, const double *q, double *src,
                 double *dtau, void *fjp, void *fjm, void *D, int *pivot, void *B,
                 og_norms *nrm);
/* as og_assemble, also returning D before its LU (Draw, [nc][25] real_fj) */
void og_assemble2(const og_grid *g, const og_params *p, const double *q, double *src,
                  double *dtau, void *fjp, void *fjm, void *D, int *pivot, void *B,
                  og_norms *nrm, void *Draw);
/* Spalart-Allmaras: the nuTilde equation of one steady Newton step.
 * src[nc] (X_src), rhs norm, scalar L = -fjp, U = fjm per face ([nf], zero
 * on boundary faces), D[nc] and B[nc]; tmu[nc] the eddy viscosity. */
void og_sa_assemble(const og_grid *g, const og_params *p, const double *q, double *src, double *fjp,
                    double *fjm, double *D, double *B, double *tmu);
/* one cell: out3 = tmu, production - destruction, destruction Jacobian */
void og_sa_point(const og_params *p, double T, double pg, double nut, double vortMag, double y, double vol,
                 double *out3);
/* Menter SST: the k and omega equations of one steady Newton step, each in
 * the layout of og_sa_assemble: src[2][nc], fjp/fjm[2][nf], D[2][nc],
 * B[2][nc] (k first); tmu[nc]; aux[4][nc] (may be NULL): F1, F2,
 * grad k . grad omega, the cross-diffusion term. */
void og_sst_assemble(const og_grid *g, const og_params *p, const double *q, double *src, double *fjp,
                     double *fjm, double *D, double *B, double *tmu, double *aux);
/* one cell of SST from its state, velocity gradient gu[9] (d u_i/d x_j at
 * 3i+j), wall distance, grad k . grad omega and volume: out[9] = F1, F2,
 * tmu, k source, omega source, k Jacobian, omega Jacobian, P_k, CD */
void og_sst_point(const og_params *p, double T, double pg, double k, double w, const double *gu, double y,
                  double gkgw, double vol, double *out);
/* scalar SGS on that system (Jacobi start + nsweeps symmetric sweeps) */
void og_sa_sgs(const og_grid *g, int nsweeps, const double *fjp, const double *fjm, const double *D,
               const double *B, double *X);
/* scalar line SGS on that system (src/linearSolverLSGS.loci:1112-1404):
 * returns iterations run, final relaxation and update norm */
int  og_sa_lsgs(const og_grid *g, int maxiter, double abstol, double reltol, double w0, const double *fjp,
                const double *fjm, const double *D, const double *B, double *X, double *w_out,
                double *resid_out);
/* line SGS: equation order and the relaxed solve (returns iterations run) */
int  og_lsgs_order(const og_grid *g, int *order);
int  og_lsgs(const og_grid *g, int maxiter, double abstol, double reltol, double w0, const void *fjp,
             const void *fjm, const void *Draw, const void *B, double *X, double *w_out,
             double *resid_out);
/* FSGS: X from zero, nsweeps symmetric sweeps in the line-SGS order with the
 * diagonal factored without pivots in real_fj (block; scalar with 1/D) */
void og_fsgs(const og_grid *g, int nsweeps, const void *fjp, const void *fjm, const void *Draw, const void *B,
             double *X);
void og_sa_fsgs(const og_grid *g, int nsweeps, const double *fjp, const double *fjm, const double *D,
                const double *B, double *X);
/* block SGS (Jacobi start + sweeps in cell-index order) on an assembled system */
void og_sgs(const og_grid *g, int nsweeps, const void *fjp, const void *fjm, const void *D,
            const int *pivot, const void *B, double *X);
/* Time-accurate steps.  BDF2 (src/solverTimestep.loci:400-428 rhs,
 * src/solverLinearSystem.loci:384-434 fluid_D, src/solverNewton.loci:160-198
 * update, src/solverScalarTransport.loci:602-716 scalars) or KEC
 * (src/solverTimestep.loci:341-398, src/solverLinearSystem.loci:226-378,
 * src/solverNewton.loci:93-157, src/solverScalarTransport.loci:367-395,
 * 514-539, 592-600, 691-700): q holds the state at level n on entry and
 * n + nsteps on exit, qm1 the level before it (both [nq][nc], updated; KEC
 * does not read it); newton_iter Newton iterations per step;
 * hist[6*(step*newton_iter + it) ..] as og_newton. */
int  og_unsteady(const og_grid *g, const og_params *p, double *q, double *qm1, int nsteps, double *hist);
/* the time-accurate systems of one Newton iteration at q with time levels
 * qn, qm1 (KEC: qm1 unused); pbias (nullable) receives KEC's localPBias */
void og_assemble_t(const og_grid *g, const og_params *p, const double *q, const double *qn, const double *qm1,
                   double *src, double *dtau, void *fjp, void *fjm, void *D, int *pivot, void *B,
                   og_norms *nrm, void *Draw, double *pbias);
void og_sa_assemble_t(const og_grid *g, const og_params *p, const double *q, const double *qn, const double *qm1,
                      double *src, double *fjp, double *fjm, double *D, double *B, double *tmu);
void og_sst_assemble_t(const og_grid *g, const og_params *p, const double *q, const double *qn,
                       const double *qm1, double *src, double *fjp, double *fjm, double *D, double *B, double *tmu,
                       double *aux);
/* one steady Newton step per call; hist[6*k..] = rtrms etrms mtrms (squared sums) per step */
int  og_newton(const og_grid *g, const og_params *p, double *q, int nsteps, double *hist);

#ifdef __cplusplus
}
#endif
#endif
