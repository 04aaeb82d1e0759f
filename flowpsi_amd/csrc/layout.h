// Device layout of one grid and its Newton-step buffers.
//
// Cells are stored in level order (Gauss-Seidel dependency level, then the
// caller's index), so every SGS level is a contiguous cell range and the
// block rows of the matrix stream sequentially through a sweep.
// Faces are stored as [interior | periodic | boundary], interior and
// periodic faces ordered by their (new) cl cell, boundary faces by type
// then cl, so a wavefront reads neighbouring cells and takes one branch.
#pragma once
#include <hip/hip_runtime.h>

namespace fpsi {

struct DevGrid {
  int nc, nf;
  int nfi;        // interior faces   [0, nfi)
  int nper;       // periodic faces   [nfi, nfi + nper)
  int nb;         // boundary faces   [nfi + nper, nf)
  int bnd_begin;  // nfi + nper
  int nlev;       // Gauss-Seidel levels
  // faces
  const double *fnx, *fny, *fnz, *farea, *fcx, *fcy, *fcz;
  const int *fcl;    // left cell
  const int *frc;    // right cell (interior), partner cell (periodic), -1 (boundary)
  const int *fbt;    // boundary type
  const double *psx, *psy, *psz;  // [nper] ghost-centre shift
  // cells
  const double *vol, *ccx, *ccy, *ccz;
  const double *lsq;  // [6][nc] inverse least-squares normal matrix
  // gather lists per cell: lower, upper, boundary faces (each ascending in the
  // caller's face index); entry = face, or ~face when the cell is its cl.
  const int *g_off, *g_ent;
  // block rows: upper faces then lower faces (each ascending); per slot the
  // neighbour (cell, or nc + periodic index) and the partner slot holding the
  // same face's other block (>= 0), or -(1 + periodic index).
  const int *m_off, *m_nbr, *m_part;
  const int *nbr_pad;  // [nc][6] row neighbours padded with the zero X entry nc + nper
  const int *slot_cl;  // [nfi + nper] slot of U_f = fjm in row cl
  const int *slot_cr;  // [nfi] slot of L_f = -fjp in row cr
  const int *old_id;   // [nc] caller's index of each cell (tie-breaks of the norms)
  // line SGS: rows read lower faces first, then upper faces.  m_nup[c] is the
  // number of upper slots of row c; lnbr_pad is nbr_pad in that order.
  const int *m_nup;    // [nc]
  const int *lnbr_pad; // [nc][6]
  const double *dist;  // [nc] wall distance (turbulent runs)
};

struct DevState {
  double *q;      // [5][nc]
  double *qf;     // [5][nb]
  double *grad;   // [15][nc]
  double *phi;    // [5][nc]
  double *iflux;  // [5][nf]
  double *vflux;  // [4][nf]
  double *maxev;  // [nf]
  double *src;    // [5][nc]
  double *dtau;   // [nc]
  void *off;      // [nnz][25] real_fj
  void *pblk;     // [nper][25] real_fj (fjp of periodic faces)
  void *bblk;     // [nb][25] real_fj (boundary contribution to srcJ)
  void *lu;       // [nc][25] real_fj
  int *piv;       // [nc] packed 3 bits per row
  void *B;        // [nc][5] real_fj
  double *X;      // [nc + nper + 1][5], last entry stays zero
  double *X2;     // second buffer of the flag-free dataflow sweep, same shape
  double *npart;  // norm partials
  double *nout;   // final norms
  int *cflag;     // [nchunks] epoch each sweep chunk last finished (dataflow sweep)
  int *serr;      // [1] set when a dataflow wait exceeded its spin limit
  double *eta;    // [nc] Eta_p (read only when Prm::pre)
  double *eta_t;  // [nc] Eta_l / Eta_ll scratch
  // Spalart-Allmaras (allocated when selected): q holds nuTilde as its 6th row
  // turbulence scalars (ns = 1 nuTilde, ns = 2 k and omega), scalar j at
  // offset j x (its per-scalar size)
  double *tmu;    // [nc] eddy viscosity
  double *saf;    // [12][nf] per-face terms of the scalar systems
  double *scell;  // [8][nc] SST per-cell terms: F1, grad k.grad w, sigma_k, sigma_w, k/w sources, k/w Jacobians
  double *soff;   // [ns][nnz] scalar row coefficients (U = fjm, L = -fjp)
  double *ssrc;   // [ns][nc] X_src
  double *sdinv;  // [ns][nc] 1 / D
  double *sB;     // [ns][nc]
  double *sX;     // [ns][nc + nper + 1] scalar solutions with ghosts and the zero entry
  double *sX2;    // second buffer of the flag-free line-SGS sweep, same shape
  double *sxt;    // [ns][nc] scalar line SGS: each cell's value before the forward sweep
  void *luf;      // [nc][25] FSGS: diagonal factored without pivots in the Jacobian precision
  // line SGS (allocated when that solver is selected)
  double *lu64;   // [nc][25] unpivoted LU of fluid_D in FP64
  double *xt;     // [nc][5] X after the forward sweep
  double *lparts; // update-norm partials
  double *lst;    // [8] w, zero_resid, resid, iterations, done
  // time-accurate BDF2 (allocated when selected)
  double *qn;     // [nvar][nc] state at time level n (start of the time step)
  double *qm1;    // [nvar][nc] state at time level n - 1
  double *rhs;    // [5][nc] BDF2 / KEC right-hand side (its norms are the residual norms)
  double *pbias;  // [nc] KEC localPBias of the current iteration
};

}  // namespace fpsi
