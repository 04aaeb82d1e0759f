// Launchers of the Newton-step kernels (kernels.hip).
#pragma once
#include "layout.h"
#include "physics.h"

namespace fpsi {
void launch_bface(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
void launch_grad(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
void launch_eta(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
void launch_face_flux(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
int norm_parts(const DevGrid &g);
void launch_gather(const DevGrid &g, const Prm &p, double vtot, DevState &s, hipStream_t st);
// norms of the BDF2 right-hand side in DevState::rhs into DevState::nout
void launch_rhs_norms(const DevGrid &g, double vtot, DevState &s, hipStream_t st);
void launch_face_jac(const DevGrid &g, const Prm &p, int fp32, DevState &s, hipStream_t st);
// fluid_D, its factors, B and the Jacobi start; with Prm::unsteady the BDF2
// system of time levels DevState::qn, qm1 (rhs also into DevState::rhs)
void launch_cell_assemble(const DevGrid &g, const Prm &p, int fp32, int lsgs, DevState &s, hipStream_t st);
void launch_sgs_level(const DevGrid &g, int fp32, int lo, int hi, DevState &s, hipStream_t st);
// dataflow half-sweep in one cooperative launch (co-resident waves: sgs_sweep_waves);
// ns > 0 sweeps the turbulence scalar systems in the same launch
int sgs_sweep_waves(int fp32, int ncu);
hipError_t launch_sgs_sweep(const DevGrid &g, int fp32, int backward, int nchunks, const int *chunk_lo,
                            const int *dep_off, const int *dep, int *flag, int epoch, int spin_limit, int *err,
                            int nwaves, int ns, long nnz, DevState &s, hipStream_t st);
// the fluid-only dataflow half-sweep without flags: reads Xs, writes Xt,
// which launch_sweep_prep has filled with sentinels (ghosts in both buffers)
hipError_t launch_sgs_flow(const DevGrid &g, int fp32, int backward, int nchunks, const int *chunk_lo, int spin_limit,
                           int *err, int nwaves, DevState &s, double *Xs, double *Xt, hipStream_t st);
// fill = 1 before the first half-sweep of a solve; the sweep leaves the
// sentinels in its source buffer for the next one
void launch_sweep_prep(const DevGrid &g, int fill, double *Xs, double *Xt, hipStream_t st);
// line SGS (fsgs = 0) or FSGS (fsgs = 1) half-sweep as one cooperative
// dataflow launch over the SGS chunks (valid when the breadth-first levels
// are the SGS levels)
int lsgs_sweep_waves(int fp32, int ncu);
// line SGS half-sweep as a flag-free dataflow (X and X2, sX and sX2
// double-buffered; see k_lsgs_flow); before the first forward half-sweep of
// a solve launch_lsgs_flow_fill puts sentinels into the forward targets,
// and launch_lsgs_relax_flow / launch_sa_lsgs_relax_flow relax into X (sX)
// and refill them
int lsgs_flow_waves(int fp32, int ncu);
hipError_t launch_lsgs_flow(const DevGrid &g, int fp32, int backward, int nchunks, const int *chunk_lo, int spin_limit,
                            int *err, int nwaves, int fluid, int ns, int smask, long nnz, DevState &s,
                            hipStream_t st);
void launch_lsgs_flow_fill(const DevGrid &g, DevState &s, int fluid, int ns, hipStream_t st);
void launch_lsgs_relax_flow(const DevGrid &g, int iter, double abstol, double reltol, DevState &s, hipStream_t st);
void launch_sa_lsgs_relax_flow(const DevGrid &g, int sc, int iter, double abstol, double reltol, DevState &s,
                               hipStream_t st, int slot);
// fluid = 0 skips the block rows; bit sc of smask sweeps turbulence scalar sc
// (line SGS rows, xt kept before the forward sweep) in the same launch
hipError_t launch_lsgs_sweep(const DevGrid &g, int fp32, int fsgs, int backward, int nchunks, const int *chunk_lo,
                             const int *dep_off, const int *dep, int *flag, int epoch, int spin_limit, int *err,
                             int nwaves, int fluid, int ns, int smask, long nnz, DevState &s, hipStream_t st);
void launch_ghost(const DevGrid &g, DevState &s, hipStream_t st);
void launch_update(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
// KEC: the time step's level n+1 from its last Newton iterate
void launch_kec_final(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
// linear residual norm of the current X; rawd: the diagonal array holds the
// unfactored fluid_D (line SGS, FSGS) rather than pivoted LU factors; rout
// (nullable) receives r = B - A X per cell [nc][5]
void launch_linres(const DevGrid &g, int fp32, int rawd, DevState &s, double *parts, double *bsum, double *out,
                   hipStream_t st, double *rout = nullptr);
void launch_copy_d(long n, const double *a, double *b, hipStream_t st);
// turbulence models: eddy viscosity (after the gradients)
void launch_tmu(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
// the turbulence scalar systems: nuTilde (turb 1) or k and omega (turb 2)
void launch_sa_assemble(const DevGrid &g, const Prm &p, long nnz, DevState &s, hipStream_t st);
// one SGS level of scalar systems sc0 .. sc0 + ns - 1 in one launch
void launch_sa_level(const DevGrid &g, int sc0, int ns, long nnz, const int *cells, int lo, int hi, DevState &s,
                     hipStream_t st);
// scalar line SGS: one level of a sweep; relaxation + stopping test into
// DevState::lst (shares the block solver's lst/lparts)
void launch_sa_lsgs_level(const DevGrid &g, int sc, long nnz, int backward, const int *cells, int lo, int hi,
                          DevState &s, hipStream_t st);
void launch_sa_lsgs_relax(const DevGrid &g, int sc, int iter, double abstol, double reltol, DevState &s,
                          hipStream_t st, int slot = 0);
void launch_sa_update(const DevGrid &g, const Prm &p, DevState &s, hipStream_t st);
// weighted r.r, r.q, q.q of r = B - A X0, q = A D into out[0..2] (needs 3 * nblk(nc) parts)
void launch_linres_pair(const DevGrid &g, int fp32, int rawd, DevState &s, const double *X0, const double *D,
                        double *parts, double *bsum, double *out, hipStream_t st);
void launch_sub_d(long n, const double *a, const double *b, double *d, hipStream_t st);
void launch_axpy_d(long n, const double *a, double alpha, const double *d, double *x, hipStream_t st);
// periodic ghost entries V[nc + i] = V[partner cell], 5-wide rows
void launch_ghost_of(const DevGrid &g, double *V, hipStream_t st);
// the same for ns scalars per cell (V holds ns x (nc + nper + 1) entries)
void launch_ghost1(const DevGrid &g, int ns, double *V, hipStream_t st);
// FSGS: unpivoted factors of the unfactored D in lu (Jacobian precision) into
// luf, one level of a sweep in the line order, and a zero fill
void launch_fsgs_factor(const DevGrid &g, int fp32, DevState &s, hipStream_t st);
void launch_fsgs_level(const DevGrid &g, int fp32, const int *cells, int lo, int hi, DevState &s, hipStream_t st);
void launch_zero_d(long n, double *x, hipStream_t st);
// line SGS
int lsgs_parts(const DevGrid &g);
void launch_lsgs_init(DevState &s, double w0, hipStream_t st, int slot = 0);
void launch_lsgs_level(const DevGrid &g, int fp32, int backward, const int *cells, int lo, int hi, DevState &s,
                       hipStream_t st);
void launch_lsgs_relax(const DevGrid &g, int iter, double abstol, double reltol, DevState &s, hipStream_t st);
}  // namespace fpsi
