/*
 * fpsi.h -- C ABI of libfpsi_gpu.so, the MI355X (gfx950) implementation of
 * flowPsi's implicit finite-volume Newton step.
 *
 * Each entry point stands in for a group of flowPsi Loci rules; the rule
 * bodies it replaces are cited per function.  A Loci rule class binds them
 * by forwarding its compute(sequence) to the matching call (see the
 * adapter listing that ships with this header's documentation).
 *
 * Conventions
 *   - Return 0 on success, a negative FPSI_ERR_* code on failure; the message
 *     is available from fpsi_last_error().  Nothing throws across the ABI.
 *   - Host arrays are owned by the caller and copied in/out; the context owns
 *     all device memory.  Calls are ordered on the context's HIP stream and
 *     are synchronous at return unless stated otherwise.
 *   - Arrays are structure-of-arrays in the caller's (reference) cell and
 *     face order: state q[5][nc] = (T, u, v, w, pg); blocks are 5x5,
 *     column-major (element (i,j) at i + 5*j), as Loci::Mat stores them.
 *   - One context per host thread.
 */
#ifndef FPSI_H
#define FPSI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fpsi_ctx fpsi_ctx;

enum {
  FPSI_OK = 0,
  FPSI_ERR_ARG = -1,     /* bad argument / inconsistent grid */
  FPSI_ERR_HIP = -2,     /* a HIP runtime call failed */
  FPSI_ERR_NAN = -3,     /* NaN residual (src/solverResidual.loci:202-208) */
  FPSI_ERR_STATE = -4,   /* call out of order (no state, no assembly, ...) */
  FPSI_ERR_LU = -5       /* singular diagonal block */
};

/* boundary types of a face (btype; bits 8-15 of a boundary face's btype
 * select its parameter set fpsi_params.bc[], see below) */
enum { FPSI_BC_INTERIOR = 0, FPSI_BC_FARFIELD = 1, FPSI_BC_WALL = 2,
       FPSI_BC_SYMMETRY = 3, FPSI_BC_PERIODIC = 4,
       FPSI_BC_REFLECTING = 5,          /* src/bcImpermeable.loci:102-181 */
       FPSI_BC_INFLOW = 6,              /* src/bcInflow.loci:377-430 */
       FPSI_BC_SUPERSONIC_INFLOW = 7,   /* src/bcInflow.loci:270-308 */
       FPSI_BC_OUTFLOW = 8,             /* src/bcOutflow.loci:174-275 (option p) */
       FPSI_BC_SUPERSONIC_OUTFLOW = 9,  /* src/bcOutflow.loci:127-171 */
       FPSI_BC_EXTRAPOLATE = 10,        /* src/bcOutflow.loci:84-125 */
       FPSI_BC_ISENTROPIC_INFLOW = 11,  /* src/bcIsentropic.loci:66-257 */
       FPSI_BC_FIXED_MASS = 12 };       /* src/bcFixedMass.loci:85-237 */
#define FPSI_BC_TYPE(btype) ((btype) & 0xFF)
#define FPSI_BC_SET(btype) (((btype) >> 8) & 0xFF)
#define FPSI_MAX_BC_SETS 8
enum { FPSI_FLUX_HLLC = 0, FPSI_FLUX_ROE = 1 };
enum { FPSI_LIM_NONE = 0, FPSI_LIM_VENKATAKRISHNAN = 1, FPSI_LIM_ZERO = 2, FPSI_LIM_BARTH = 3 };
enum { FPSI_PREC_FP64 = 0, FPSI_PREC_FP32 = 1 };
/* fluidLinearSolver (src/solverNewton.loci:229-259): block SGS, line SGS
 * (src/linearSolverLSGS.loci), or FSGS (src/linearSolverFSGS.loci) */
enum { FPSI_SOLVER_SGS = 0, FPSI_SOLVER_LSGS = 1, FPSI_SOLVER_FSGS = 2 };
/* turbulence model, solved segregated like the reference: none,
 * Spalart-Allmaras (turbulence/SA/SA_turb.loci; baseline, standard
 * production; the state carries nuTilde as its 6th row), or Menter SST
 * (turbulence/KOmega/sst.loci with its default options; the state carries
 * k and omega as rows 6 and 7) */
enum { FPSI_TURB_NONE = 0, FPSI_TURB_SA = 1, FPSI_TURB_SST = 2 };
/* timeStepMode steady, or unsteady with timeStepScheme BDF2 or KEC */
enum { FPSI_TIME_STEADY = 0, FPSI_TIME_BDF2 = 1, FPSI_TIME_KEC = 2 };
/* How a Gauss-Seidel half-sweep runs (same result bit for bit):
 * PER_LEVEL  one launch per dependency level (default);
 * PERSISTENT one cooperative launch per half-sweep run as a dataflow: each
 *            chunk of <= 12 cells waits only for the cells it depends on
 *            (the fluid alone: by polling their values in a second X
 *            buffer that starts as sentinels; with turbulence scalars in
 *            the launch: per-chunk flags);
 * TILED      one cooperative launch over tiles of <= 32 cells whose levels run
 *            in LDS (needs a grid whose cell boxes form an acyclic tile graph;
 *            fpsi_info.sweep_tiles > 0). */
enum { FPSI_SWEEP_PER_LEVEL = 0, FPSI_SWEEP_PERSISTENT = 1, FPSI_SWEEP_TILED = 2,
       /* PERSISTENT_FLAGS: the dataflow with per-chunk flags for every system
        * (the flag-free form polls the neighbours' values themselves) */
       FPSI_SWEEP_PERSISTENT_FLAGS = 3 };

/* Grid: the Loci FVM facts cl, cr, area{n,sada}, facecenter, vol,
 * cellcenter, pmap (used at src/fluxInviscid.loci:1170-1188,
 * src/solverLinearSystem.loci:106-120, src/solverPeriodic.loci:25-87).
 * Interior faces must have cl < cr (cell order is then a topological
 * order of the Gauss-Seidel dependency graph). */
typedef struct {
  int nc, nf;
  const int *cl;        /* [nf] */
  const int *cr;        /* [nf] >= 0 interior, -1 boundary (periodic included) */
  const int *btype;     /* [nf] FPSI_BC_* */
  const int *pface;     /* [nf] periodic partner face, -1 otherwise */
  const double *n;      /* [3][nf] unit normal, cl -> cr / outward */
  const double *area;   /* [nf] */
  const double *fc;     /* [3][nf] face centre */
  const double *vol;    /* [nc] */
  const double *cc;     /* [3][nc] cell centre */
  const double *pshift; /* [3][nf] periodic ghost centre = cc[cl[pface]] + pshift */
} fpsi_grid_desc;

/* Reference values of one boundary (the BC options of
 * src/bcInputs.loci:188-220 after their unit conversion): temperatureRef_BC,
 * gagePressureRef_BC (p - Pambient) and uRef_BC for inflow, supersonic
 * inflow and (pg only) outflow; p0 and T0 for isentropic inflow; mdot and
 * T0 for fixed mass (massFlux = mdot / the summed area of the set's
 * fixed-mass faces, :230-232: give each such boundary its own set). */
typedef struct {
  double T, pg, u[3];
  double p0, T0;
  double mdot;
} fpsi_bc_set;

/* Parameters, with flowPsi's option names and code defaults:
 * gamma/molecularMass/p0 src/solverEoS.loci:27-47; flux include/inviscidFlux.h:82-90;
 * limiter/Kl (Loci FVM); sutherlandParameters src/fluxViscous.loci:59-110;
 * Twall src/bcViscousWall.loci:152-154; farfield reference state
 * src/bcInflow.loci:311-340; cflmax/dtmax/urelax src/solverTimestep.loci:281-297;
 * gauss_seidel_iter src/linearSolverSGS.loci:39-41; LSGSMaxIter/RelTol/AbsTol/
 * Relaxation src/linearSolverLSGS.loci:96-107 (10, 0.1, 1e-10, 0.25). */
typedef struct {
  double gamma, Rtilde, Pambient;
  int flux;             /* FPSI_FLUX_* */
  int limiter;          /* FPSI_LIM_* */
  double Kl;
  int viscous;          /* 0 inviscid, 1 laminar */
  double suth[6];       /* a1 a2 a3 k1 k2 k3 */
  double Twall;
  double pg0, T0, u0[3];
  double cflmax;        /* <= 0: not set */
  double dtmax, urelax;
  int linear_solver;    /* FPSI_SOLVER_* */
  int sgs_iter;         /* symmetric sweeps */
  int precision;        /* FPSI_PREC_*: storage/arithmetic of real_fj (Jacobians, D, LU, B) */
  double minPg;         /* floor of the reconstructed gage pressure */
  int sweep_mode;       /* FPSI_SWEEP_*: how a Gauss-Seidel half-sweep is launched (SGS) */
  int lsgs_maxiter;     /* line SGS: at most this many relaxed symmetric sweeps */
  double lsgs_reltol;   /* stop when the update norm falls below reltol x the first one */
  double lsgs_abstol;   /* ... or below abstol */
  double lsgs_relax;    /* relaxation w: X = w X_sweep + (1-w) X_forward */
  /* low-Mach preconditioning (src/fluxInviscid.loci:39-155); a value <= 0
   * takes flowPsi's default: Minf 1 (no preconditioning), etaT 0.25, PLimPC 10 */
  double Minf, etaT, PLimPC;
  int turb;             /* FPSI_TURB_* */
  double Prt;           /* turbulentPrandtlNumber (<= 0: 0.9) */
  double nut_bc;        /* nuTilde of inflow faces (<= 0: 5e-9, the BC default) */
  double k_bc, w_bc;    /* SST: k and omega of inflow faces (<= 0: the BC defaults from |u0|,
                           turbulence/KOmega/sst_viscous.loci:63-118) */
  /* timeStepMode (src/solverTimestep.loci:85-140): FPSI_TIME_STEADY, or
   * unsteady with timeStepScheme BDF2 (FPSI_TIME_BDF2) or KEC (FPSI_TIME_KEC)
   * with the time step dtmax, newtonMaxIter Newton iterations per time step
   * (src/solverNewton.loci:197-199, <= 0: 6), pseudo_cflmax of BDF2's
   * preconditioning blend (src/solverLinearSystem.loci:380-382, <= 0: 10)
   * and KEC's timePressureBiasControl (src/solverTimestep.loci:340-344) */
  int time_mode;
  int newton_iter;
  double pseudo_cflmax;
  double pbias_control;
  /* polls of one dependency wait of a dataflow sweep before it gives up and
   * the call returns FPSI_ERR_HIP (<= 0: 1 << 22, far longer than any real
   * wait); after such an error the system must be assembled again */
  int sweep_spin_limit;
  /* temporalDamping beta (src/fluxInviscid.loci:1224-1251): when > 0 the
   * face Jacobians fjp / fjm of interior and periodic faces gain the
   * damping_fjp / damping_fjm terms (:794-866); <= 0: not set */
  double temporal_damping;
  /* co-resident waves per compute unit of a line-SGS / FSGS dataflow
   * launch (<= 0: every wave that fits) */
  int sweep_waves_per_cu;
  /* the parameter sets of the boundaries of types REFLECTING ..
   * FIXED_MASS: a face reads bc[FPSI_BC_SET(btype)] (farfield keeps pg0,
   * T0, u0 above).  These types are inviscid or laminar only. */
  fpsi_bc_set bc[FPSI_MAX_BC_SETS];
} fpsi_params;

/* src/solverResidual.loci:31-112 (struct residual in include/flowTypes.h:517-523);
 * rms fields are the volume-weighted sums before the square root. */
typedef struct {
  double rtrms, etrms, mtrms, rtmax, etmax, mtmax;
  double ccmaxr[3], ccmaxe[3];
} fpsi_norms;

typedef struct {
  int nlevels;          /* Gauss-Seidel dependency levels */
  int max_level_width;
  int nnz_blocks;       /* off-diagonal blocks stored */
  double device_bytes;  /* device memory held by the context */
  int lsgs_nlevels;     /* dependency levels of the line-SGS order */
  int lsgs_contiguous;  /* 1 when every line-SGS level is a contiguous cell range */
  int sweep_tiles;      /* tiles of the tiled sweep (0: the grid admits none) */
  int lsgs_dataflow;    /* 1 when line SGS / FSGS run as the dataflow sweep under
                           FPSI_SWEEP_PERSISTENT (their levels are the SGS levels) */
} fpsi_info;

/* Build the context: renumbers cells into level order, builds the cell->face
 * gather lists and the block-row layout, copies the grid to the device.
 * Replaces Loci's setup of the FVM maps (createLowerUpper,
 * src/gridReader.loci:226-229) for this path. */
int fpsi_create(fpsi_ctx **ctx, const fpsi_grid_desc *grid);
void fpsi_destroy(fpsi_ctx *ctx);
const char *fpsi_last_error(const fpsi_ctx *ctx);
int fpsi_get_info(fpsi_ctx *ctx, fpsi_info *info);

/* Options facts (the $rule default(...) values listed in fpsi_params). */
int fpsi_set_params(fpsi_ctx *ctx, const fpsi_params *params);

/* temperature, u, gagePressure stores (include/flowPsi.lh:53-65):
 * q[5][nc] (T, u, v, w, pg); q[6][nc] with nuTilde under FPSI_TURB_SA;
 * q[7][nc] with k and omega under FPSI_TURB_SST; fpsi_set_params first. */
int fpsi_set_state(fpsi_ctx *ctx, const double *q);
int fpsi_get_state(fpsi_ctx *ctx, double *q);

/* dist_noslip[nc] (src/solverTurbulence.loci:133-142, Loci's nearest
 * no-slip face): needed by the turbulence model. */
int fpsi_set_wall_distance(fpsi_ctx *ctx, const double *dist);

/* Residual path: boundary face states (src/bcInflow.loci:311-340,
 * src/bcViscousWall.loci:205-214, src/bcImpermeable.loci:48-64), gradients
 * and limiters (Loci FVM), iflux (src/fluxInviscid.loci:1170-1188 and the
 * boundary iflux rules), vflux (src/fluxViscous.loci:217-292), the src join
 * (src/fluxInviscid.loci:1253-1268, src/fluxViscous.loci:299-310) and the
 * residual norms (src/solverResidual.loci:74-112).  Any output may be NULL.
 * src[5][nc], iflux[5][nf], vflux[4][nf]. */
int fpsi_residual(fpsi_ctx *ctx, double *src, double *iflux, double *vflux, fpsi_norms *norms);

/* Gradient [15][nc] (d q_k / d x_d at 3k+d) and limiter [5][nc] of the current state. */
int fpsi_gradients(fpsi_ctx *ctx, double *grad, double *phi);

/* Linear system of one steady Newton step: dtau (src/solverTimestep.loci:290-329),
 * fjp/fjm (src/fluxInviscid.loci:1190-1221, src/fluxViscous.loci:315-403),
 * boundary srcJ (src/bcInflow.loci:357-375, src/bcViscousWall.loci:230-240,
 * src/bcImpermeable.loci:75-100, src/fluxViscous.loci:405-448), srcJ and
 * fluid_D/L/U/B (src/solverLinearSystem.loci:27-224), LU of D
 * (src/linearSolverSGS.loci:43-50) and the Jacobi start (:57-69). */
int fpsi_assemble(fpsi_ctx *ctx);

/* Copy the assembled system out in the reference layout: fjp/fjm [nf][25]
 * for interior and periodic faces (zero elsewhere), D [nc][25] (LU factors
 * under FPSI_SOLVER_SGS; the unfactored fluid_D under FPSI_SOLVER_LSGS, whose
 * factors are kept in FP64), pivot [nc][5] (identity under LSGS), B [nc][5]
 * (real_fj of the current precision: 8 or 4 bytes per element), dtau [nc].
 * Any output may be NULL. */
int fpsi_get_system(fpsi_ctx *ctx, void *fjp, void *fjm, void *D, int *pivot, void *B, double *dtau);

/* Linear solve of the assembled system; dq[nc][5] may be NULL (the result
 * stays on the device as dfluid).
 * FPSI_SOLVER_SGS: block SGS (src/linearSolverSGS.loci:57-331), Jacobi start
 *   then nsweeps symmetric level-scheduled sweeps.
 * FPSI_SOLVER_LSGS: LSGSBlockedSSolve (src/linearSolverLSGS.loci:1692-1954),
 *   line-Jacobi start then at most nsweeps relaxed symmetric sweeps in the
 *   breadth-first line order, stopping on lsgs_abstol / lsgs_reltol.
 * FPSI_SOLVER_FSGS: FSGSBlockedSSolve (src/linearSolverFSGS.loci:1598-1708),
 *   X from zero, nsweeps symmetric sweeps in the same breadth-first order
 *   with the diagonal factored without pivots in the Jacobian precision. */
int fpsi_linear_solve(fpsi_ctx *ctx, int nsweeps, double *dq);

/* nuTilde system of the assembled step (FPSI_TURB_SA): src[nc] (X_src),
 * fjp/fjm[nf] (zero on boundary faces; L = -fjp, U = fjm), D[nc], B[nc],
 * tmu[nc]; any output may be NULL (the X_* rules of
 * src/solverScalarTransport.loci:140-615 for X = nuTilde,
 * turbulence/SA/SA_turb.loci:755-756).  fpsi_sa_solve solves it from the
 * Jacobi start and returns dX[nc]:
 * FPSI_SOLVER_SGS: SGSScalarSolve (src/linearSolverSGS.loci:333-500),
 *   nsweeps symmetric level-scheduled sweeps;
 * FPSI_SOLVER_LSGS: LSGSScalarSolve (src/linearSolverLSGS.loci:1112-1404),
 *   at most nsweeps relaxed sweeps in the breadth-first order with the
 *   lsgs_* stopping tests; fpsi_sa_lsgs_stats reports its outcome;
 * FPSI_SOLVER_FSGS: FSGSScalarSolve (src/linearSolverFSGS.loci:1039-1200),
 *   X from zero, nsweeps symmetric sweeps in that order. */
int fpsi_sa_get_system(fpsi_ctx *ctx, double *src, double *fjp, double *fjm, double *D, double *B, double *tmu);
int fpsi_sa_solve(fpsi_ctx *ctx, int nsweeps, double *dX);
int fpsi_sa_lsgs_stats(fpsi_ctx *ctx, int *iters, double *relax, double *resid);

/* The same for any turbulence scalar: which = 0 nuTilde (SA) or k (SST),
 * 1 omega (SST).  Under SST the k and omega systems are the
 * scalarTransportP(k | w, SST_BSL_KW) instances (turbulence/KOmega/
 * sst.loci:856-860) with the sources, Jacobians and diffusion coefficients
 * of sst.loci:425-1006 and the generic scalar rules of
 * src/solverScalarTransport.loci. */
int fpsi_turb_get_system(fpsi_ctx *ctx, int which, double *src, double *fjp, double *fjm, double *D, double *B,
                         double *tmu);
int fpsi_turb_solve(fpsi_ctx *ctx, int which, int nsweeps, double *dX);
int fpsi_turb_lsgs_stats(fpsi_ctx *ctx, int which, int *iters, double *relax, double *resid);

/* Outcome of the last line-SGS solve: sweeps run, final relaxation w and
 * update norm (the diagnostics of src/linearSolverLSGS.loci:1934-1940). */
int fpsi_lsgs_stats(fpsi_ctx *ctx, int *iters, double *relax, double *resid);

/* Damped update q += alpha dq (src/solverNewton.loci:62-78; BDF2 :160-180)
 * with the current dfluid. */
int fpsi_update(fpsi_ctx *ctx);

/* nsteps fused steady Newton steps, device resident: residual, assembly,
 * linear solve (sgs_iter sweeps, or lsgs_maxiter under LSGS), update.  hist[nsteps] (nullable) receives the norms of the
 * residual evaluated at the start of each step. */
int fpsi_newton_step(fpsi_ctx *ctx, int nsteps, fpsi_norms *hist);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, serving loops that forbid hipMalloc):
 * fpsi_newton_step with the per-step norm history held in a fixed device
 * array of FPSI_NOALLOC_MAX_STEPS steps (12 doubles each, 384 KiB per GPU,
 * held until the process ends) instead of a per-call allocation; the call
 * allocates no device or host memory on success.  nsteps above the maximum
 * is refused (FPSI_ERR_ARG), and so is a context with profiling on (its
 * events are created on demand).  The array is one per GPU, shared by every
 * context there and keeping the last call's contents: calls that overlap on
 * one GPU (several host threads) race. */
#define FPSI_NOALLOC_MAX_STEPS 4096
int fpsi_newton_step_noalloc(fpsi_ctx *ctx, int nsteps, fpsi_norms *hist);

/* Time-accurate steps (FPSI_TIME_BDF2, FPSI_TIME_KEC).  The context holds the
 * state q at time level n and the level before it, q{n-1} (read by BDF2
 * only); fpsi_set_state makes q{n-1} = q (the reference's start "constant in
 * time", src/solverTimestep.loci:156-170).  Within a time step, the systems
 * fpsi_assemble builds are those of a Newton iteration with time levels q{n}
 * and q{n-1}: BDF2 rhs src/solverTimestep.loci:400-428, fluid_D
 * src/solverLinearSystem.loci:384-434, scalar rows
 * src/solverScalarTransport.loci:534-539, 602-616, 702-716; KEC (the iterate
 * is the mid-step state q*) localPBias and rhs src/solverTimestep.loci:346-396,
 * fluid_D src/solverLinearSystem.loci:226-378, scalar rows
 * src/solverScalarTransport.loci:388-395, 514-539, 592-600, 691-700.  The
 * norms are those of the right-hand side (src/solverResidual.loci:113-150).
 *
 * fpsi_set_time_levels sets q{n} and q{n-1} (a restart's icfile levels,
 * src/solverIC.loci:562-570, 612-660); qn NULL takes the current state.
 * fpsi_get_time_levels returns them ([nvar][nc] each, nullable). */
int fpsi_set_time_levels(fpsi_ctx *ctx, const double *qn, const double *qnm1);
int fpsi_get_time_levels(fpsi_ctx *ctx, double *qn, double *qnm1);

/* nsteps time steps, device resident: q{n,0} = q{n}; newton_iter
 * iterations of residual + time terms, assembly, linear solve (as
 * fpsi_newton_step), update (fluid and turbulence scalars, src/solverNewton.loci:93-125
 * and :160-180); then fpsi_finish_time_step.  hist[nsteps * newton_iter]
 * (nullable): the norms of each iteration's right-hand side. */
int fpsi_unsteady_step(fpsi_ctx *ctx, int nsteps, fpsi_norms *hist);

/* Ends a time step driven by fpsi_assemble / fpsi_linear_solve /
 * fpsi_update: q{n+1} = the last iterate (BDF2, src/solverNewton.loci:182-189)
 * or its KEC reconstruction (:127-157, scalars
 * src/solverScalarTransport.loci:367-377), then q{n-1} = q{n}. */
int fpsi_finish_time_step(fpsi_ctx *ctx);

/* Kernel groups timed with HIP events on the context's stream while
 * profiling is on, and their algorithmic bytes per launch (compulsory
 * traffic, each array counted once per launch, computed from the actual
 * buffer sizes). */
enum { FPSI_K_RESIDUAL = 0,   /* boundary states, gradients+limiters, face fluxes, gather+norms */
       FPSI_K_FACE_JAC = 1,   /* face Jacobian blocks */
       FPSI_K_ASSEMBLE = 2,   /* cell assembly + LU + Jacobi start */
       FPSI_K_SGS_SWEEP = 3,  /* one Gauss-Seidel half-sweep (all of its level launches) */
       FPSI_K_UPDATE = 4,     /* Newton update */
       FPSI_K_NGROUPS = 5 };
typedef struct {
  double ms[FPSI_K_NGROUPS];        /* summed event time */
  long launches[FPSI_K_NGROUPS];    /* timed launches */
  double bytes[FPSI_K_NGROUPS];     /* algorithmic bytes of one launch */
} fpsi_profile;
int fpsi_profile_enable(fpsi_ctx *ctx, int on);   /* clears the accumulators */
int fpsi_profile_read(fpsi_ctx *ctx, fpsi_profile *out);

/* ---- linear-solve tolerance and the learned initial guess ------------- */

/* Linear residual ||B - A X|| of the assembled system for the current X
 * (dfluid), each equation scaled by the RMS of its right-hand side.  The
 * reference's SGS runs a fixed sweep count and has no such test
 * (src/linearSolverSGS.loci:315-328); it is used to run the solve to a
 * tolerance and to judge a learned guess. */
int fpsi_linear_residual(fpsi_ctx *ctx, double *norm);

typedef struct {
  double r0;       /* residual of the plain Jacobi start X = D^-1 B */
  double rguess;   /* residual of the learned guess (0 if not tried) */
  double rfinal;   /* residual after the sweeps */
  int sweeps;      /* symmetric sweeps run */
  int used_guess;  /* 1 if the guess beat the Jacobi start and was kept */
  double alpha;    /* step taken along the learned correction (X0 + alpha (guess - X0)) */
} fpsi_lin_stats;

/* Block SGS (or, with linear_solver = FPSI_SOLVER_LSGS and the dataflow
 * sweep mode, relaxed line-SGS sweeps) from the Jacobi start, or from the
 * learned guess when use_guess is set and the guess has a smaller residual
 * than the Jacobi start (otherwise the plain start is kept), sweeping until
 * ||r|| <= rtol * r0 or max_sweeps.  The residual costs about a quarter of a
 * symmetric sweep, so it is first evaluated one sweep before the count the
 * previous tolerance solve since fpsi_set_state needed. */
int fpsi_linear_solve_tol(fpsi_ctx *ctx, double rtol, int max_sweeps, int use_guess, fpsi_lin_stats *st);

/* Steady Newton steps whose fluid linear solves run to a tolerance (see
 * above); the turbulence scalars keep their own solves (gauss_seidel_iter
 * block-SGS sweeps, riding in the fluid's first dataflow launches, or
 * their line-SGS loops in the fluid's line-SGS launches).  hist[nsteps]
 * and lin[nsteps] are nullable. */
int fpsi_newton_adaptive(fpsi_ctx *ctx, int nsteps, double rtol, int max_sweeps, int use_guess,
                         fpsi_norms *hist, fpsi_lin_stats *lin);

/* Learned initial guess: a two-hop message-passing MLP of width 64 (bf16
 * GEMM operands on MFMA, fp32 accumulation) mapping per-cell features (the
 * Jacobi start X0, the previous Newton update, the diagonal factors' pivots,
 * the residual r0 = B - A X0 of the start and its Jacobi correction D^-1 r0,
 * face geometry) to a correction of the Jacobi start, under block SGS and
 * under line SGS.  No counterpart in the reference. */
int fpsi_mlp_num_params(void);
int fpsi_mlp_init(fpsi_ctx *ctx, unsigned long long seed);
int fpsi_mlp_get_params(fpsi_ctx *ctx, float *p);
int fpsi_mlp_set_params(fpsi_ctx *ctx, const float *p);
/* On an assembled system: record features and, as the target, the SGS
 * solution after target_sweeps sweeps (left in dfluid); *slot receives the
 * sample slot (the oldest of 16 is overwritten). */
int fpsi_mlp_harvest(fpsi_ctx *ctx, int target_sweeps, int *slot);
/* Adam over all recorded samples, epochs passes, mean loss of the last pass. */
int fpsi_mlp_train(fpsi_ctx *ctx, int epochs, float lr, double *loss);
/* Forward (and backward) on one sample: out[nc][16], grad[nparams], loss. */
int fpsi_mlp_eval(fpsi_ctx *ctx, int slot, float *out, float *grad, double *loss);
/* One sample in caller cell order: x[nc][32], y[nc][8]; the neighbour lists
 * nb_off[nc+1], nb_idx[nb_off[nc]] the network aggregates over. */
int fpsi_mlp_get_sample(fpsi_ctx *ctx, int slot, float *x, float *y, int *nb_off, int *nb_idx);

/* Wait for the context's stream. */
int fpsi_synchronize(fpsi_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
