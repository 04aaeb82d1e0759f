// Host side of libfpsi_gpu.so: grid setup, device buffers and the C ABI of
// include/fpsi.h.  All numerics run in the kernels of kernels.hip; this file
// only builds index structures, moves data and orders launches on the
// context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fpsi.h"
#include "kernels.h"
#include "mlp.h"
#include "tiles.h"

using namespace fpsi;

// Norm history of fpsi_newton_step_noalloc: one array per GPU in the code
// object, FPSI_NOALLOC_MAX_STEPS steps of 12 doubles.  Host code never uses
// the name (a host-side stand-in): the device address of each GPU's copy is
// looked up once per device, with that device current, when the first
// context is created on it, and kept in scratch_dev.
static __device__ __align__(256) double scratch[12 * FPSI_NOALLOC_MAX_STEPS];
static constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex scratch_mu;
static double *scratch_dev[SCRATCH_MAX_DEVICES] = {};

struct fpsi_ctx {
  std::string err;
  hipStream_t stream = nullptr;
  double bc_area[FPSI_MAX_BC_SETS] = {};  // summed area of each parameter set's fixed-mass faces
  bool has_set_bc = false;                // some face is of a type REFLECTING .. FIXED_MASS
  double *bcs = nullptr;                  // device [FPSI_MAX_BC_SETS][8]: T, pg, u, p0, T0, massFlux
  double *scratch = nullptr;  // this device's copy of `scratch`
  double *hist_host = nullptr;  // pinned, 12 x 64 doubles: norm history read-back
  int nc = 0, nf = 0, nfi = 0, nper = 0, nb = 0, nnz = 0;
  std::vector<int> cperm;   // new -> old cell
  std::vector<int> icperm;  // old -> new cell
  std::vector<int> fperm;   // new -> old face
  std::vector<int> ifperm;  // old -> new face
  std::vector<int> lev_off; // level ranges in new numbering
  int nchunks = 0;          // sweep chunks (<= 12 cells, within one level)
  int sweep_waves = 0;      // co-resident waves of the cooperative sweep
  int flow_waves = 0;       // waves of the flag-free cooperative sweep
  // dataflow sweep: chunk ranges and, per chunk, the chunks holding its
  // lower (forward) / upper (backward) neighbours; off when some chunk has
  // more than 64 of either
  const int *chunk_lo = nullptr, *fdep_off = nullptr, *fdep = nullptr, *bdep_off = nullptr, *bdep = nullptr;
  bool have_flow = false;
  int flow_epoch = 0;
  // line SGS runs as the same dataflow when its breadth-first levels are the
  // SGS levels (every cell in the same level, so the same dependency graph)
  bool lsgs_flow = false;
  int lsgs_waves = 0;
  int lsgs_fwaves = 0;      // waves of the flag-free line-SGS sweep
  int ncu = 1;              // compute units of the device
  int max_width = 0;
  double vtot = 0;
  DevGrid g{};
  DevState s{};
  std::vector<void *> allocs;
  size_t bytes = 0;
  fpsi_params prm{};
  Prm kp{};
  bool have_params = false, have_state = false, assembled = false;
  int fp32 = 0;
  // profiling: event pairs per kernel group, resolved at fpsi_profile_read
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  struct TimedLaunch {
    int grp;
    hipEvent_t start, end;
    double bytes;  // algorithmic bytes of this launch (< 0: the group's fixed model)
  };
  std::vector<TimedLaunch> ev_pending;
  double prof_ms[FPSI_K_NGROUPS] = {0};
  long prof_n[FPSI_K_NGROUPS] = {0};
  double prof_b[FPSI_K_NGROUPS] = {0};   // summed per-launch bytes
  long prof_nb[FPSI_K_NGROUPS] = {0};    // launches that carried their own bytes
  // linear residual scratch
  double *lr_parts = nullptr, *lr_bsum = nullptr, *lr_out = nullptr, *Xj = nullptr, *Dg = nullptr;
  // learned guess: graph and geometry (new cell order), buffers, samples
  std::vector<int> mlp_nb_off, mlp_nb_idx;
  std::vector<float> mlp_geo;
  bool mlp_alloc = false, mlp_norm = false;
  MlpBufs mb{};
  static constexpr int MLP_SLOTS = 16;  // harvested samples kept (the oldest is overwritten)
  float *mlp_xs[MLP_SLOTS] = {}, *mlp_ys[MLP_SLOTS] = {};
  float *mlp_xcur = nullptr;
  double *mlp_xprev = nullptr;  // [nc][5] the last Newton update (zero before the first)
  int tol_sweeps_last = 0;       // sweeps the last solve to a tolerance ran
  int mlp_nsamp = 0, mlp_next = 0, mlp_step = 0;
  // line SGS: levels of the breadth-first line order (new cell numbering);
  // lsgs_cells is null when every level is a contiguous cell range
  std::vector<int> lsgs_lev_off;
  const int *lsgs_cells = nullptr;
  bool lsgs_alloc = false;
  double *lsgs_host = nullptr;  // pinned mirror of DevState::lst
  // pinned readback slots: [0, 12) step norms, [12] linear residual, [13]
  // the sweep error word (as int), so several copies share one sync
  double *rb_host = nullptr;
  int lsgs_iters = 0;
  double lsgs_w = 0, lsgs_resid = 0;
  int sa_lsgs_iters[2] = {0, 0};  // per turbulence scalar
  double sa_lsgs_w[2] = {0, 0}, sa_lsgs_resid[2] = {0, 0};
  size_t fj_bytes() const { return fp32 ? 4 : 8; }
  bool lsgs() const { return prm.linear_solver == FPSI_SOLVER_LSGS; }
  bool fsgs() const { return prm.linear_solver == FPSI_SOLVER_FSGS; }
  // line SGS and FSGS share the breadth-first order and the unfactored D
  bool line_order() const { return prm.linear_solver != FPSI_SOLVER_SGS; }
  // turbulence models: buffers allocated on first selection, wall distance set by the caller
  bool sa_alloc = false, have_dist = false;
  int nscal() const { return prm.turb == FPSI_TURB_SST ? 2 : prm.turb == FPSI_TURB_SA ? 1 : 0; }
  int nvar() const { return 5 + nscal(); }
  // tiled sweeps (none when the grid admits no acyclic tiling)
  TileGrid tg{};
  bool have_tiles = false;
  std::string tile_msg;
  int tile_groups = 0;
  int epoch = 0;
  bool tiled() const {
    return have_tiles && prm.sweep_mode == FPSI_SWEEP_TILED;
  }
  // time-accurate BDF2: buffers allocated on first selection; have_prev when
  // DevState::qm1 holds q{n-1}, in_step when DevState::qn holds the q{n} of
  // a time step in progress
  bool bdf2() const { return prm.time_mode == FPSI_TIME_BDF2 || prm.time_mode == FPSI_TIME_KEC; }
  bool bdf2_alloc = false, have_prev = false, in_step = false;
  int newton_iter = 6;
};

static int fail(fpsi_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, FPSI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A dependency wait of a dataflow sweep ran past its spin limit: the launch
// ended early and X holds a partial result.  The error word is cleared so
// that later launches start clean, and the system must be assembled again
// (which rewrites the Jacobi start; the next flag-free solve refills its
// sentinels) before the next solve.
static int sweep_failed(fpsi_ctx *c, const char *what) {
  hipMemsetAsync(c->s.serr, 0, sizeof(int), c->stream);
  hipStreamSynchronize(c->stream);
  c->assembled = false;
  return fail(c, FPSI_ERR_HIP, std::string(what) + ": a dependency wait exceeded its limit");
}

// read the dataflow error word after the sweeps enqueued so far
static int check_sweep_error(fpsi_ctx *c, const char *what) {
  int err = 0;
  if (hipMemcpyAsync(&err, c->s.serr, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(c, FPSI_ERR_HIP, std::string(what) + ": reading the sweep error word failed");
  return err ? sweep_failed(c, what) : 0;
}

// polls of one dependency wait before a dataflow sweep gives up: far longer
// than any real wait (~1e6 polls of a few hundred ns each), so only a broken
// schedule reaches it; fpsi_params.sweep_spin_limit overrides it
// both dataflow modes: flag-free (PERSISTENT) or with per-chunk flags
static bool dataflow(const fpsi_ctx *c) {
  return c->prm.sweep_mode == FPSI_SWEEP_PERSISTENT || c->prm.sweep_mode == FPSI_SWEEP_PERSISTENT_FLAGS;
}

// co-resident waves of a dataflow line-SGS launch: sweep_waves_per_cu when
// set (capped at what fits), else every co-resident wave (wide levels)
static int line_waves(const fpsi_ctx *c, int cap) {
  const int per_cu = c->prm.sweep_waves_per_cu;
  return per_cu > 0 ? std::max(1, std::min(cap, per_cu * c->ncu)) : cap;
}

static int spin_limit(const fpsi_ctx *c) { return c->prm.sweep_spin_limit > 0 ? c->prm.sweep_spin_limit : 1 << 22; }

template <typename T>
static int dalloc(fpsi_ctx *c, T **p, size_t n) {
  void *v = nullptr;
  size_t b = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc(&v, b);
  if (e != hipSuccess) return fail(c, FPSI_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(v);
  c->bytes += b;
  *p = (T *)v;
  return 0;
}

template <typename T>
static int upload(fpsi_ctx *c, const T **dst, const std::vector<T> &h) {
  T *d = nullptr;
  int rc = dalloc(c, &d, h.size());
  if (rc) return rc;
  if (!h.empty()) HIPCHK(c, hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  *dst = d;
  return 0;
}

extern "C" const char *fpsi_last_error(const fpsi_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static hipEvent_t ev_get(fpsi_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// record the start of a timed kernel group; returns the event (or null)
static hipEvent_t prof_begin(fpsi_ctx *c) {
  if (!c->prof) return nullptr;
  hipEvent_t e = ev_get(c);
  if (e) hipEventRecord(e, c->stream);
  return e;
}
static void prof_end(fpsi_ctx *c, int grp, hipEvent_t start, double bytes = -1.0) {
  if (!c->prof || !start) return;
  hipEvent_t e = ev_get(c);
  if (!e) return;
  hipEventRecord(e, c->stream);
  c->ev_pending.push_back({grp, start, e, bytes});
}

// Algorithmic bytes of one flag-free line-SGS half-sweep (each array once,
// neighbour gathers cache-perfect), from this context's buffer sizes: the
// fluid's block rows (row offset and neighbour index per slot), FP64
// diagonal factors, B, X read and written, and in the forward direction the
// sentinel written over the source word; per turbulence scalar in the
// launch its row coefficients, B, 1/D, X read and written, and forward its
// xt and sentinel.
static double sweep_bytes(const fpsi_ctx *c, bool fluid, int nscal, int backward) {
  const double nc = c->nc, nnz = c->nnz, ps = 8, pj = (double)c->fj_bytes();
  double b = nc * 8 /* m_off, m_nup */ + nc * 6 * 4 /* padded neighbour list */;
  if (fluid) b += nnz * 25 * pj + nc * (25 * 8 + 5 * pj + 2 * 5 * ps + (backward ? 0 : 5 * ps));
  b += nscal * (nnz * ps + nc * (ps + ps + 2 * ps + (backward ? 0 : 2 * ps)));
  return b;
}

extern "C" void fpsi_destroy(fpsi_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto &p : ctx->ev_pending) { hipEventDestroy(p.start); hipEventDestroy(p.end); }
  for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
  for (void *p : ctx->allocs) hipFree(p);
  if (ctx->lsgs_host) hipHostFree(ctx->lsgs_host);
  if (ctx->rb_host) hipHostFree(ctx->rb_host);
  if (ctx->hist_host) hipHostFree(ctx->hist_host);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int fpsi_create(fpsi_ctx **out, const fpsi_grid_desc *gd) {
  if (!out || !gd) return FPSI_ERR_ARG;
  *out = nullptr;
  fpsi_ctx *c = new fpsi_ctx();
  const int nc = gd->nc, nf = gd->nf;
  if (nc <= 0 || nf <= 0 || !gd->cl || !gd->cr || !gd->btype || !gd->n || !gd->area || !gd->fc ||
      !gd->vol || !gd->cc) {
    *out = c;
    return fail(c, FPSI_ERR_ARG, "grid descriptor incomplete");
  }
  c->nc = nc;
  c->nf = nf;
  // ---- validate and classify faces
  std::vector<int> kind(nf);  // 0 interior, 1 periodic, 2 boundary
  for (int f = 0; f < nf; ++f) {
    const int t = gd->btype[f], l = gd->cl[f], r = gd->cr[f];
    if (l < 0 || l >= nc) { *out = c; return fail(c, FPSI_ERR_ARG, "cl out of range at face " + std::to_string(f)); }
    if (t == FPSI_BC_INTERIOR) {
      if (r <= l || r >= nc) { *out = c; return fail(c, FPSI_ERR_ARG, "interior face needs cl < cr < nc at face " + std::to_string(f)); }
      kind[f] = 0;
    } else if (t == FPSI_BC_PERIODIC) {
      const int pf = gd->pface ? gd->pface[f] : -1;
      if (pf < 0 || pf >= nf || gd->btype[pf] != FPSI_BC_PERIODIC || gd->pface[pf] != f || !gd->pshift) {
        *out = c;
        return fail(c, FPSI_ERR_ARG, "periodic face without a matching partner at face " + std::to_string(f));
      }
      kind[f] = 1;
    } else if (FPSI_BC_TYPE(t) >= FPSI_BC_FARFIELD && FPSI_BC_TYPE(t) <= FPSI_BC_FIXED_MASS &&
               FPSI_BC_TYPE(t) != FPSI_BC_PERIODIC && FPSI_BC_SET(t) < FPSI_MAX_BC_SETS && (t >> 16) == 0) {
      if (r != -1) { *out = c; return fail(c, FPSI_ERR_ARG, "boundary face must have cr = -1"); }
      kind[f] = 2;
      if (FPSI_BC_TYPE(t) == FPSI_BC_FIXED_MASS) c->bc_area[FPSI_BC_SET(t)] += gd->area[f];
      if (FPSI_BC_TYPE(t) >= FPSI_BC_REFLECTING) c->has_set_bc = true;
    } else {
      *out = c;
      return fail(c, FPSI_ERR_ARG, "unsupported boundary type at face " + std::to_string(f));
    }
  }
  // ---- per-cell face lists (caller order), lower / upper / boundary
  std::vector<int> lo_off(nc + 1, 0), up_off(nc + 1, 0), bd_off(nc + 1, 0);
  for (int f = 0; f < nf; ++f) {
    if (kind[f] == 0) { up_off[gd->cl[f] + 1]++; lo_off[gd->cr[f] + 1]++; }
    else if (kind[f] == 1) up_off[gd->cl[f] + 1]++;
    else bd_off[gd->cl[f] + 1]++;
  }
  for (int i = 0; i < nc; ++i) { lo_off[i + 1] += lo_off[i]; up_off[i + 1] += up_off[i]; bd_off[i + 1] += bd_off[i]; }
  std::vector<int> lo(lo_off[nc]), up(up_off[nc]), bd(bd_off[nc]);
  {
    std::vector<int> a(lo_off.begin(), lo_off.end() - 1), b(up_off.begin(), up_off.end() - 1),
        d(bd_off.begin(), bd_off.end() - 1);
    for (int f = 0; f < nf; ++f) {
      if (kind[f] == 0) { up[b[gd->cl[f]]++] = f; lo[a[gd->cr[f]]++] = f; }
      else if (kind[f] == 1) up[b[gd->cl[f]]++] = f;
      else bd[d[gd->cl[f]]++] = f;
    }
  }
  // ---- Gauss-Seidel levels: longest path in the cl -> cr graph
  std::vector<int> level(nc, 0);
  int nlev = 0;
  for (int i = 0; i < nc; ++i) {
    int L = 0;
    for (int j = lo_off[i]; j < lo_off[i + 1]; ++j) L = std::max(L, level[gd->cl[lo[j]]] + 1);
    level[i] = L;
    nlev = std::max(nlev, L + 1);
  }
  c->lev_off.assign(nlev + 1, 0);
  for (int i = 0; i < nc; ++i) c->lev_off[level[i] + 1]++;
  for (int l = 0; l < nlev; ++l) {
    c->max_width = std::max(c->max_width, c->lev_off[l + 1]);
    c->lev_off[l + 1] += c->lev_off[l];
  }
  c->cperm.resize(nc);
  c->icperm.resize(nc);
  {
    std::vector<int> pos(c->lev_off.begin(), c->lev_off.end() - 1);
    for (int i = 0; i < nc; ++i) {
      const int ni = pos[level[i]]++;
      c->cperm[ni] = i;
      c->icperm[i] = ni;
    }
  }
  const std::vector<int> &ic = c->icperm;
  // ---- face order: interior by new cl, periodic by new cl, boundary by type then new cl
  std::vector<int> fi, fp, fb;
  for (int f = 0; f < nf; ++f) (kind[f] == 0 ? fi : kind[f] == 1 ? fp : fb).push_back(f);
  auto by_cl = [&](int a, int b) {
    const int ca = ic[gd->cl[a]], cb = ic[gd->cl[b]];
    return ca != cb ? ca < cb : a < b;
  };
  std::stable_sort(fi.begin(), fi.end(), by_cl);
  std::stable_sort(fp.begin(), fp.end(), by_cl);
  std::stable_sort(fb.begin(), fb.end(), [&](int a, int b) {
    if (gd->btype[a] != gd->btype[b]) return gd->btype[a] < gd->btype[b];
    return by_cl(a, b);
  });
  c->nfi = (int)fi.size();
  c->nper = (int)fp.size();
  c->nb = (int)fb.size();
  c->fperm.clear();
  c->fperm.insert(c->fperm.end(), fi.begin(), fi.end());
  c->fperm.insert(c->fperm.end(), fp.begin(), fp.end());
  c->fperm.insert(c->fperm.end(), fb.begin(), fb.end());
  c->ifperm.assign(nf, -1);
  for (int i = 0; i < nf; ++i) c->ifperm[c->fperm[i]] = i;
  const std::vector<int> &ifp = c->ifperm;
  const int nfi = c->nfi, nper = c->nper;

  // ---- face arrays in new order
  std::vector<double> fnx(nf), fny(nf), fnz(nf), farea(nf), fcx(nf), fcy(nf), fcz(nf);
  std::vector<int> fcl(nf), frc(nf), fbt(nf);
  std::vector<double> psx(nper), psy(nper), psz(nper);
  for (int i = 0; i < nf; ++i) {
    const int f = c->fperm[i];
    fnx[i] = gd->n[f]; fny[i] = gd->n[nf + f]; fnz[i] = gd->n[2 * nf + f];
    farea[i] = gd->area[f];
    fcx[i] = gd->fc[f]; fcy[i] = gd->fc[nf + f]; fcz[i] = gd->fc[2 * nf + f];
    fcl[i] = ic[gd->cl[f]];
    fbt[i] = gd->btype[f];
    if (kind[f] == 0) frc[i] = ic[gd->cr[f]];
    else if (kind[f] == 1) {
      frc[i] = ic[gd->cl[gd->pface[f]]];
      const int pi = i - nfi;
      psx[pi] = gd->pshift[f]; psy[pi] = gd->pshift[nf + f]; psz[pi] = gd->pshift[2 * nf + f];
    } else frc[i] = -1;
  }
  // ---- cell arrays in new order
  std::vector<double> vol(nc), ccx(nc), ccy(nc), ccz(nc);
  c->vtot = 0;
  for (int i = 0; i < nc; ++i) c->vtot += gd->vol[i];  // caller order, as the norm join sums
  for (int ni = 0; ni < nc; ++ni) {
    const int i = c->cperm[ni];
    vol[ni] = gd->vol[i];
    ccx[ni] = gd->cc[i]; ccy[ni] = gd->cc[nc + i]; ccz[ni] = gd->cc[2 * nc + i];
  }
  // ---- gather lists (lower, upper, boundary), least-squares inverse
  std::vector<int> g_off(nc + 1, 0), g_ent;
  g_ent.reserve(lo.size() + up.size() + bd.size());
  std::vector<double> lsq(6 * (size_t)nc);
  for (int ni = 0; ni < nc; ++ni) {
    const int i = c->cperm[ni];
    double A[6] = {0, 0, 0, 0, 0, 0};
    const double xc[3] = {gd->cc[i], gd->cc[nc + i], gd->cc[2 * nc + i]};
    auto addA = [&](const double *x) {
      const double dx = x[0] - xc[0], dy = x[1] - xc[1], dz = x[2] - xc[2];
      A[0] += dx * dx; A[1] += dx * dy; A[2] += dx * dz;
      A[3] += dy * dy; A[4] += dy * dz; A[5] += dz * dz;
    };
    for (int j = lo_off[i]; j < lo_off[i + 1]; ++j) {
      const int f = lo[j];
      g_ent.push_back(ifp[f]);
      const int o = gd->cl[f];
      const double x[3] = {gd->cc[o], gd->cc[nc + o], gd->cc[2 * nc + o]};
      addA(x);
    }
    for (int j = up_off[i]; j < up_off[i + 1]; ++j) {
      const int f = up[j];
      g_ent.push_back(~ifp[f]);
      double x[3];
      if (kind[f] == 0) {
        const int o = gd->cr[f];
        x[0] = gd->cc[o]; x[1] = gd->cc[nc + o]; x[2] = gd->cc[2 * nc + o];
      } else {
        const int o = gd->cl[gd->pface[f]];
        x[0] = gd->cc[o]; x[1] = gd->cc[nc + o]; x[2] = gd->cc[2 * nc + o];
        for (int d = 0; d < 3; ++d) x[d] += gd->pshift[d * nf + f];
      }
      addA(x);
    }
    for (int j = bd_off[i]; j < bd_off[i + 1]; ++j) {
      const int f = bd[j];
      g_ent.push_back(~ifp[f]);
      const double x[3] = {gd->fc[f], gd->fc[nf + f], gd->fc[2 * nf + f]};
      addA(x);
    }
    g_off[ni + 1] = (int)g_ent.size();
    const double c00 = A[3] * A[5] - A[4] * A[4];
    const double c01 = A[2] * A[4] - A[1] * A[5];
    const double c02 = A[1] * A[4] - A[2] * A[3];
    const double c11 = A[0] * A[5] - A[2] * A[2];
    const double c12 = A[1] * A[2] - A[0] * A[4];
    const double c22 = A[0] * A[3] - A[1] * A[1];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (!(std::fabs(det) > 0)) {
      *out = c;
      return fail(c, FPSI_ERR_ARG, "degenerate least-squares stencil at cell " + std::to_string(i));
    }
    const double rdet = 1.0 / det;
    lsq[ni] = c00 * rdet; lsq[nc + ni] = c01 * rdet; lsq[2 * nc + ni] = c02 * rdet;
    lsq[3 * nc + ni] = c11 * rdet; lsq[4 * nc + ni] = c12 * rdet; lsq[5 * nc + ni] = c22 * rdet;
  }
  // ---- block rows: upper faces then lower faces
  std::vector<int> m_off(nc + 1, 0), m_nbr, m_part;
  std::vector<int> slot_cl(nfi + nper, -1), slot_cr(nfi, -1);
  for (int ni = 0; ni < nc; ++ni) {
    const int i = c->cperm[ni];
    for (int j = up_off[i]; j < up_off[i + 1]; ++j) {
      const int f = up[j];
      const int nfid = ifp[f];
      slot_cl[nfid] = (int)m_nbr.size();
      m_nbr.push_back(kind[f] == 0 ? ic[gd->cr[f]] : nc + (nfid - nfi));
      m_part.push_back(0);
    }
    for (int j = lo_off[i]; j < lo_off[i + 1]; ++j) {
      const int f = lo[j];
      slot_cr[ifp[f]] = (int)m_nbr.size();
      m_nbr.push_back(ic[gd->cl[f]]);
      m_part.push_back(0);
    }
    m_off[ni + 1] = (int)m_nbr.size();
  }
  // partner slots: for every stored block, where the face's other block sits
  for (int nfid = 0; nfid < nfi + nper; ++nfid) {
    const int s_cl = slot_cl[nfid];
    if (nfid < nfi) {
      const int s_cr = slot_cr[nfid];
      m_part[s_cl] = s_cr;   // upper entry of cl reads L = -fjp (srcJ -= fjp)
      m_part[s_cr] = s_cl;   // lower entry of cr reads U = fjm  (srcJ += fjm)
    } else {
      m_part[s_cl] = -1 - (nfid - nfi);
    }
  }
  c->nnz = (int)m_nbr.size();
  std::vector<int> old_id(c->cperm);
  // neighbour graph of the learned guess (periodic ghosts mapped to cells)
  c->mlp_nb_off.assign(nc + 1, 0);
  c->mlp_nb_idx.clear();
  c->mlp_nb_idx.reserve(m_nbr.size());
  for (int ni = 0; ni < nc; ++ni) {
    for (int j = m_off[ni]; j < m_off[ni + 1]; ++j) {
      const int v = m_nbr[j];
      c->mlp_nb_idx.push_back(v < nc ? v : frc[nfi + (v - nc)]);
    }
    c->mlp_nb_off[ni + 1] = (int)c->mlp_nb_idx.size();
  }
  // geometry features: sum_f A |n_d| / vol^(2/3) per axis, log vol - mean
  c->mlp_geo.assign(4 * (size_t)nc, 0.f);
  {
    double lmean = 0;
    for (int i = 0; i < nc; ++i) lmean += std::log(gd->vol[i]);
    lmean /= nc;
    for (int ni = 0; ni < nc; ++ni) {
      const int i = c->cperm[ni];
      double sh[3] = {0, 0, 0};
      auto addf = [&](int f) {
        for (int d = 0; d < 3; ++d) sh[d] += gd->area[f] * std::fabs(gd->n[(size_t)d * nf + f]);
      };
      for (int j = lo_off[i]; j < lo_off[i + 1]; ++j) addf(lo[j]);
      for (int j = up_off[i]; j < up_off[i + 1]; ++j) addf(up[j]);
      for (int j = bd_off[i]; j < bd_off[i + 1]; ++j) addf(bd[j]);
      const double s23 = std::pow(gd->vol[i], 2.0 / 3.0);
      for (int d = 0; d < 3; ++d) c->mlp_geo[4 * (size_t)ni + d] = (float)(sh[d] / s23);
      c->mlp_geo[4 * (size_t)ni + 3] = (float)(std::log(gd->vol[i]) - lmean);
    }
  }
  // fixed-width neighbour rows for the sweeps; pads name the zero X entry
  const int zslot = nc + nper;
  std::vector<int> nbr_pad(6 * (size_t)nc, zslot);
  for (int ni = 0; ni < nc; ++ni)
    for (int j = m_off[ni]; j < m_off[ni + 1] && j - m_off[ni] < 6; ++j) nbr_pad[6 * (size_t)ni + (j - m_off[ni])] = m_nbr[j];

  // ---- line SGS.  Rows read lower slots first: slot j of the rotated row is
  // slot (j + nup) mod len of the stored row.
  std::vector<int> m_nup(nc), lnbr_pad(6 * (size_t)nc, zslot);
  for (int ni = 0; ni < nc; ++ni) {
    const int i = c->cperm[ni];
    const int nup = up_off[i + 1] - up_off[i], len = m_off[ni + 1] - m_off[ni];
    m_nup[ni] = nup;
    for (int j = 0; j < len && j < 6; ++j) {
      int jr = j + nup;
      if (jr >= len) jr -= len;
      lnbr_pad[6 * (size_t)ni + j] = m_nbr[m_off[ni] + jr];
    }
  }
  // Breadth-first order of the one-cell lines (src/linearSolverLSGS.loci:902-937):
  // from the lowest unvisited cell, layer by layer over interior faces, each
  // layer ascending.  Its sweep dependencies run from earlier to later cells;
  // level = longest path, cells of a level sorted by new index.
  std::vector<int> lsgs_cells;
  {
    std::vector<char> vis(nc, 0);
    std::vector<int> order, cur, nxt;
    order.reserve(nc);
    int sp = 0;
    while ((int)order.size() < nc) {
      if (cur.empty()) {
        while (sp < nc && vis[sp]) ++sp;
        if (sp == nc) break;
        cur.push_back(sp);
        vis[sp] = 1;
      }
      nxt.clear();
      for (int i : cur) {
        for (int j = lo_off[i]; j < lo_off[i + 1]; ++j) {
          const int o = gd->cl[lo[j]];
          if (!vis[o]) { vis[o] = 1; nxt.push_back(o); }
        }
        for (int j = up_off[i]; j < up_off[i + 1]; ++j) {
          const int f = up[j];
          if (kind[f] != 0) continue;
          const int o = gd->cr[f];
          if (!vis[o]) { vis[o] = 1; nxt.push_back(o); }
        }
        order.push_back(i);
      }
      std::sort(nxt.begin(), nxt.end());
      cur.swap(nxt);
    }
    std::vector<int> pos(nc), lv(nc, 0);
    for (int k = 0; k < nc; ++k) pos[order[k]] = k;
    int nl = 0;
    for (int k = 0; k < nc; ++k) {
      const int i = order[k];
      int L = 0;
      for (int j = lo_off[i]; j < lo_off[i + 1]; ++j) {
        const int o = gd->cl[lo[j]];
        if (pos[o] < k) L = std::max(L, lv[o] + 1);
      }
      for (int j = up_off[i]; j < up_off[i + 1]; ++j) {
        const int f = up[j];
        if (kind[f] != 0) continue;
        const int o = gd->cr[f];
        if (pos[o] < k) L = std::max(L, lv[o] + 1);
      }
      lv[i] = L;
      nl = std::max(nl, L + 1);
    }
    c->lsgs_lev_off.assign(nl + 1, 0);
    for (int i = 0; i < nc; ++i) c->lsgs_lev_off[lv[i] + 1]++;
    for (int l = 0; l < nl; ++l) c->lsgs_lev_off[l + 1] += c->lsgs_lev_off[l];
    lsgs_cells.assign(nc, 0);
    std::vector<int> fill(c->lsgs_lev_off.begin(), c->lsgs_lev_off.end() - 1);
    for (int ni = 0; ni < nc; ++ni) lsgs_cells[fill[lv[c->cperm[ni]]]++] = ni;
    bool ident = true;
    for (int k = 0; k < nc && ident; ++k) ident = lsgs_cells[k] == k;
    if (ident) lsgs_cells.clear();
    c->lsgs_flow = ident && c->lsgs_lev_off == c->lev_off;  // and have_flow, below
  }

  // ---- tiles of the tiled sweep
  TileHost th;
  {
    std::vector<int> lvl_new(nc);
    for (int ni = 0; ni < nc; ++ni) lvl_new[ni] = level[c->cperm[ni]];
    c->tile_msg = build_tiles(nc, ccx, ccy, ccz, lvl_new, m_off, m_nbr, nbr_pad, th);
    c->have_tiles = c->tile_msg.empty();
  }

  // ---- every index the kernels dereference must be in bounds
  {
    const int nnz = c->nnz;
    bool ok = (int)g_ent.size() == g_off[nc] && (int)m_nbr.size() == m_off[nc];
    for (int e : g_ent) ok = ok && (e < 0 ? ~e : e) < nf;
    for (int v : m_nbr) ok = ok && v >= 0 && v < nc + nper;
    for (int v : m_part) ok = ok && (v >= 0 ? v < nnz : (-1 - v) < nper);
    for (int v : slot_cl) ok = ok && v >= 0 && v < nnz;
    for (int v : slot_cr) ok = ok && v >= 0 && v < nnz;
    for (int v : nbr_pad) ok = ok && v >= 0 && v <= nc + nper;
    for (int v : lnbr_pad) ok = ok && v >= 0 && v <= nc + nper;
    for (int ni = 0; ni < nc; ++ni) ok = ok && m_nup[ni] >= 0 && m_nup[ni] <= m_off[ni + 1] - m_off[ni];
    for (int v : lsgs_cells) ok = ok && v >= 0 && v < nc;
    if (c->have_tiles) {
      ok = ok && (int)th.cells.size() == nc && th.cell_off.back() == nc && (int)th.nbr.size() == 6 * nc &&
           (int)th.row.size() == nc && (int)th.len.size() == nc && (int)th.lvi.size() == nc;
      for (int k = 0; k < th.ntiles; ++k)
        for (int x = th.cell_off[k]; x < th.cell_off[k + 1]; ++x)
          ok = ok && th.lvi[x] >= 0 && th.lvi[x] < th.lev_off[k + 1] - th.lev_off[k] - 1;
      for (int x = 0; x < nc; ++x)
        ok = ok && th.len[x] >= 0 && th.len[x] <= 6 && th.row[x] >= 0 && th.row[x] + th.len[x] <= c->nnz;
      for (int v : th.cells) ok = ok && v >= 0 && v < nc;
      for (int k = 0; k < th.ntiles; ++k) {
        const int n = th.cell_off[k + 1] - th.cell_off[k];
        ok = ok && n >= 1 && n <= TILE_MAX && th.lev[th.lev_off[k]] == 0 && th.lev[th.lev_off[k + 1] - 1] == n;
        for (int x = th.cell_off[k] * 6; x < th.cell_off[k + 1] * 6; ++x)
          ok = ok && (th.nbr[x] >= 0 ? th.nbr[x] <= nc + nper : ~th.nbr[x] < n);
      }
      for (int v : th.up) ok = ok && v >= 0 && v < th.ntiles;
      for (int v : th.dn) ok = ok && v >= 0 && v < th.ntiles;
    }
    ok = ok && c->lsgs_lev_off.back() == nc;
    for (int i = 0; i < nf; ++i) {
      ok = ok && fcl[i] >= 0 && fcl[i] < nc;
      if (i < nfi + nper) ok = ok && frc[i] >= 0 && frc[i] < nc;
    }
    if (!ok) {
      *out = c;
      return fail(c, FPSI_ERR_ARG, "internal index structure out of bounds");
    }
  }

  // ---- device upload (a failure from here on leaves the context to the
  // caller for fpsi_last_error and fpsi_destroy)
  *out = c;
  HIPCHK(c, hipStreamCreate(&c->stream));
  HIPCHK(c, hipHostMalloc((void **)&c->rb_host, 16 * sizeof(double), hipHostMallocDefault));
  HIPCHK(c, hipHostMalloc((void **)&c->hist_host, 12 * 64 * sizeof(double), hipHostMallocDefault));
  int sweep_ncu = 1, sweep_cap = 0;
  {
    int dev = 0, ncu = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
      *out = c;
      return fail(c, FPSI_ERR_ARG, "device ordinal beyond the norm-history table");
    }
    {
      std::lock_guard<std::mutex> lk(scratch_mu);
      if (!scratch_dev[dev]) {
        double *p = nullptr;
        HIPCHK(c, hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = p;
      }
      c->scratch = scratch_dev[dev];
    }
    HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    sweep_ncu = ncu;
    c->ncu = ncu;
    sweep_cap = sgs_sweep_waves(1, ncu) > 0 ? std::min(sgs_sweep_waves(0, ncu), sgs_sweep_waves(1, ncu)) : 0;
    const int lcap = lsgs_sweep_waves(1, ncu) > 0 ? std::min(lsgs_sweep_waves(0, ncu), lsgs_sweep_waves(1, ncu)) : 0;
    c->lsgs_waves = std::max(1, lcap);  // wide levels: every co-resident wave
    const int fcap = lsgs_flow_waves(1, ncu) > 0 ? std::min(lsgs_flow_waves(0, ncu), lsgs_flow_waves(1, ncu)) : 0;
    c->lsgs_fwaves = std::max(1, fcap);
  }
  DevGrid &g = c->g;
  g.nc = nc; g.nf = nf; g.nfi = nfi; g.nper = nper; g.nb = c->nb; g.bnd_begin = nfi + nper;
  g.nlev = nlev;
  std::vector<int> chunk_lo;
  for (int l = 0; l < nlev; ++l)
    for (int a = c->lev_off[l]; a < c->lev_off[l + 1]; a += 12) chunk_lo.push_back(a);  // 12 cells per wave
  c->nchunks = (int)chunk_lo.size();
  chunk_lo.push_back(nc);
  {
    // 1.75 waves per chunk of a mean level, to the nearest whole number of
    // waves per CU, at least four per CU and at most every co-resident wave:
    // narrow levels (128^3 channel, ~460 chunks) are fastest at 4 per CU,
    // the 8M-cell prism grid (~1170 chunks) at 8 (10 per CU polls more and
    // is 2.5 % slower)
    const long want = 7L * c->nchunks / (4L * std::max(1, nlev));
    const long per_cu = std::max(4L, (want + sweep_ncu / 2) / std::max(1, sweep_ncu));
    c->sweep_waves = (int)std::max(1L, std::min((long)sweep_cap, per_cu * sweep_ncu));
    // the flag-free sweep of the fluid alone polls its neighbours' words
    // rather than flags and is fastest with at least 6 waves per CU on the
    // channel (1.451 ms per half-sweep against 1.473 at 4, 1.468 at 8)
    c->flow_waves = (int)std::max(1L, std::min((long)sweep_cap, std::max(6L, per_cu) * sweep_ncu));
  }
  std::vector<int> fdep_off(c->nchunks + 1, 0), fdep, bdep_off(c->nchunks + 1, 0), bdep;
  {
    std::vector<int> chunk_of(nc);
    for (int ch = 0; ch < c->nchunks; ++ch)
      for (int a = chunk_lo[ch]; a < chunk_lo[ch + 1]; ++a) chunk_of[a] = ch;
    bool ok = true;
    std::vector<int> lo_d, up_d;
    for (int ch = 0; ch < c->nchunks; ++ch) {
      lo_d.clear();
      up_d.clear();
      for (int a = chunk_lo[ch]; a < chunk_lo[ch + 1]; ++a)
        for (int j = m_off[a]; j < m_off[a + 1]; ++j) {
          const int o = m_nbr[j];
          if (o >= nc) continue;  // periodic ghost: a lagged copy, no dependency
          if (o < chunk_lo[ch]) lo_d.push_back(chunk_of[o]);
          else if (o >= chunk_lo[ch + 1]) up_d.push_back(chunk_of[o]);
        }
      std::sort(lo_d.begin(), lo_d.end());
      lo_d.erase(std::unique(lo_d.begin(), lo_d.end()), lo_d.end());
      std::sort(up_d.begin(), up_d.end());
      up_d.erase(std::unique(up_d.begin(), up_d.end()), up_d.end());
      ok = ok && lo_d.size() <= 64 && up_d.size() <= 64;
      fdep.insert(fdep.end(), lo_d.begin(), lo_d.end());
      bdep.insert(bdep.end(), up_d.begin(), up_d.end());
      fdep_off[ch + 1] = (int)fdep.size();
      bdep_off[ch + 1] = (int)bdep.size();
    }
    c->have_flow = ok;
    c->lsgs_flow = c->lsgs_flow && ok;
    fdep.push_back(0);  // never empty
    bdep.push_back(0);
  }
  int rc = 0;
  if ((rc = upload(c, &g.fnx, fnx)) || (rc = upload(c, &g.fny, fny)) || (rc = upload(c, &g.fnz, fnz)) ||
      (rc = upload(c, &g.farea, farea)) || (rc = upload(c, &g.fcx, fcx)) || (rc = upload(c, &g.fcy, fcy)) ||
      (rc = upload(c, &g.fcz, fcz)) || (rc = upload(c, &g.fcl, fcl)) || (rc = upload(c, &g.frc, frc)) ||
      (rc = upload(c, &g.fbt, fbt)) || (rc = upload(c, &g.psx, psx)) || (rc = upload(c, &g.psy, psy)) ||
      (rc = upload(c, &g.psz, psz)) || (rc = upload(c, &g.vol, vol)) || (rc = upload(c, &g.ccx, ccx)) ||
      (rc = upload(c, &g.ccy, ccy)) || (rc = upload(c, &g.ccz, ccz)) || (rc = upload(c, &g.lsq, lsq)) ||
      (rc = upload(c, &g.g_off, g_off)) || (rc = upload(c, &g.g_ent, g_ent)) ||
      (rc = upload(c, &g.m_off, m_off)) || (rc = upload(c, &g.m_nbr, m_nbr)) ||
      (rc = upload(c, &g.m_part, m_part)) || (rc = upload(c, &g.slot_cl, slot_cl)) ||
      (rc = upload(c, &g.slot_cr, slot_cr)) || (rc = upload(c, &g.old_id, old_id)) ||
      (rc = upload(c, &g.nbr_pad, nbr_pad)) || (rc = upload(c, &g.m_nup, m_nup)) ||
      (rc = upload(c, &g.lnbr_pad, lnbr_pad)) ||
      (rc = upload(c, &c->chunk_lo, chunk_lo)) || (rc = upload(c, &c->fdep_off, fdep_off)) ||
      (rc = upload(c, &c->fdep, fdep)) || (rc = upload(c, &c->bdep_off, bdep_off)) ||
      (rc = upload(c, &c->bdep, bdep))) {
    *out = c;
    return rc;
  }
  DevState &s = c->s;
  const size_t nbb = c->nb;
  // state, face states, gradients and limiters hold up to two turbulence
  // scalars after the five fluid variables
  if ((rc = dalloc(c, &s.q, 7 * (size_t)nc)) || (rc = dalloc(c, &s.qf, 7 * nbb)) ||
      (rc = dalloc(c, &s.grad, 21 * (size_t)nc)) || (rc = dalloc(c, &s.phi, 7 * (size_t)nc)) ||
      (rc = dalloc(c, &s.iflux, 5 * (size_t)nf)) || (rc = dalloc(c, &s.vflux, 4 * (size_t)nf)) ||
      (rc = dalloc(c, &s.maxev, (size_t)nf)) || (rc = dalloc(c, &s.src, 5 * (size_t)nc)) ||
      (rc = dalloc(c, &s.dtau, (size_t)nc)) || (rc = dalloc(c, &s.piv, (size_t)nc)) ||
      (rc = dalloc(c, &s.X, 5 * ((size_t)nc + nper + 1))) ||
      (rc = dalloc(c, &s.X2, 5 * ((size_t)nc + nper + 1))) ||
      (rc = dalloc(c, &s.npart, 12 * (size_t)norm_parts(g))) || (rc = dalloc(c, &s.nout, 12)) ||
      (rc = dalloc(c, &s.cflag, (size_t)std::max(c->nchunks, 1))) || (rc = dalloc(c, &s.serr, 1)) ||
      (rc = dalloc(c, &s.eta, (size_t)nc)) ||
      (rc = dalloc(c, &s.eta_t, (size_t)nc))) {
    *out = c;
    return rc;
  }
  if ((rc = dalloc(c, &c->lr_parts, 5 * (size_t)norm_parts(g) + 5)) || (rc = dalloc(c, &c->lr_bsum, 8)) ||
      (rc = dalloc(c, &c->lr_out, 8)) || (rc = dalloc(c, &c->Xj, 5 * ((size_t)nc + nper + 1))) ||
      (rc = dalloc(c, &c->Dg, 5 * ((size_t)nc + nper + 1)))) {
    *out = c;
    return rc;
  }
  // real_fj buffers sized for FP64 so the precision can change without reallocation
  double *tmp = nullptr;
  if ((rc = dalloc(c, &tmp, 25 * (size_t)std::max(c->nnz, 1)))) { *out = c; return rc; }
  s.off = tmp;
  if ((rc = dalloc(c, &tmp, 25 * (size_t)nper))) { *out = c; return rc; }
  s.pblk = tmp;
  if ((rc = dalloc(c, &tmp, 25 * nbb))) { *out = c; return rc; }
  s.bblk = tmp;
  if ((rc = dalloc(c, &tmp, 25 * (size_t)nc))) { *out = c; return rc; }
  s.lu = tmp;
  if ((rc = dalloc(c, &tmp, 5 * (size_t)nc))) { *out = c; return rc; }
  s.B = tmp;
  HIPCHK(c, hipMemsetAsync(s.vflux, 0, 4 * (size_t)nf * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(s.X, 0, 5 * ((size_t)nc + nper + 1) * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(s.X2, 0, 5 * ((size_t)nc + nper + 1) * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(c->Dg, 0, 5 * ((size_t)nc + nper + 1) * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(s.cflag, 0, (size_t)std::max(c->nchunks, 1) * sizeof(int), c->stream));
  HIPCHK(c, hipMemsetAsync(s.serr, 0, sizeof(int), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!lsgs_cells.empty()) {
    if ((rc = upload(c, &c->lsgs_cells, lsgs_cells))) { *out = c; return rc; }
  }
  if (c->have_tiles) {
    TileGrid &t = c->tg;
    t.ntiles = th.ntiles;
    if ((rc = upload(c, &t.cell_off, th.cell_off)) || (rc = upload(c, &t.cells, th.cells)) ||
        (rc = upload(c, &t.lev_off, th.lev_off)) || (rc = upload(c, &t.lev, th.lev)) ||
        (rc = upload(c, &t.up_off, th.up_off)) || (rc = upload(c, &t.up, th.up)) ||
        (rc = upload(c, &t.dn_off, th.dn_off)) || (rc = upload(c, &t.dn, th.dn)) ||
        (rc = upload(c, &t.nbr, th.nbr)) || (rc = upload(c, &t.row, th.row)) ||
        (rc = upload(c, &t.len, th.len)) || (rc = upload(c, &t.lvi, th.lvi)) ||
        (rc = dalloc(c, &t.flag, (size_t)th.ntiles))) {
      *out = c;
      return rc;
    }
    HIPCHK(c, hipMemset(t.flag, 0, (size_t)th.ntiles * sizeof(int)));
    int dev = 0, ncu = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int cap = std::min(sgs_tiles_groups(0, ncu), sgs_tiles_groups(1, ncu));
    c->tile_groups = std::min(cap, th.ntiles);
    if (c->tile_groups <= 0) {
      c->have_tiles = false;
      c->tile_msg = "no co-resident workgroups for the tiled sweep";
    }
  }
  *out = c;
  return 0;
}

// line-SGS buffers, allocated the first time that solver is selected
static int lsgs_alloc(fpsi_ctx *c) {
  if (c->lsgs_alloc) return 0;
  DevState &s = c->s;
  int rc = 0;
  if ((rc = dalloc(c, &s.lu64, 25 * (size_t)c->nc)) || (rc = dalloc(c, &s.xt, 5 * (size_t)c->nc)) ||
      (rc = dalloc(c, &s.lparts, 3 * (size_t)lsgs_parts(c->g))) || (rc = dalloc(c, &s.lst, 24)))
    return rc;
  // slot 0 the fluid, slots 1-2 the turbulence scalars (their loops may run together)
  HIPCHK(c, hipHostMalloc((void **)&c->lsgs_host, 24 * sizeof(double), hipHostMallocDefault));
  c->lsgs_alloc = true;
  return 0;
}

extern "C" int fpsi_get_info(fpsi_ctx *c, fpsi_info *info) {
  if (!c || !info) return FPSI_ERR_ARG;
  info->nlevels = (int)c->lev_off.size() - 1;
  info->max_level_width = c->max_width;
  info->nnz_blocks = c->nnz;
  info->device_bytes = (double)c->bytes;
  info->lsgs_nlevels = (int)c->lsgs_lev_off.size() - 1;
  info->lsgs_contiguous = c->lsgs_cells == nullptr;
  info->lsgs_dataflow = c->lsgs_flow ? 1 : 0;
  info->sweep_tiles = c->have_tiles ? c->tg.ntiles : 0;
  return 0;
}

static int sa_alloc(fpsi_ctx *c);

extern "C" int fpsi_set_params(fpsi_ctx *c, const fpsi_params *p) {
  if (!c || !p) return FPSI_ERR_ARG;
  if (p->flux != FPSI_FLUX_HLLC && p->flux != FPSI_FLUX_ROE) return fail(c, FPSI_ERR_ARG, "unknown flux");
  if (p->limiter < 0 || p->limiter > 3) return fail(c, FPSI_ERR_ARG, "unknown limiter");
  if (p->precision != FPSI_PREC_FP64 && p->precision != FPSI_PREC_FP32) return fail(c, FPSI_ERR_ARG, "unknown precision");
  if (p->linear_solver != FPSI_SOLVER_SGS && p->linear_solver != FPSI_SOLVER_LSGS &&
      p->linear_solver != FPSI_SOLVER_FSGS)
    return fail(c, FPSI_ERR_ARG, "unknown linear solver");
  if (p->turb != FPSI_TURB_NONE && p->turb != FPSI_TURB_SA && p->turb != FPSI_TURB_SST)
    return fail(c, FPSI_ERR_ARG, "unknown turbulence model");
  if (p->turb != FPSI_TURB_NONE) {
    if (!p->viscous) return fail(c, FPSI_ERR_ARG, "a turbulence model needs viscous = 1");
    if (c->has_set_bc)
      return fail(c, FPSI_ERR_ARG, "turbulence scalars are not built for the boundary types REFLECTING .. FIXED_MASS");
    if (int rc = sa_alloc(c)) return rc;
  }
  if (p->linear_solver == FPSI_SOLVER_LSGS) {
    if (p->lsgs_maxiter < 0) return fail(c, FPSI_ERR_ARG, "LSGSMaxIter < 0");
    if (!(p->lsgs_relax > 0.0)) return fail(c, FPSI_ERR_ARG, "LSGSRelaxation must be > 0");
    if (int rc = lsgs_alloc(c)) return rc;
  }
  if (p->linear_solver == FPSI_SOLVER_FSGS) {
    if (int rc = lsgs_alloc(c)) return rc;  // the line-SGS assembly writes lu64
    if (!c->s.luf) {
      double *luf = nullptr;
      if (int rc = dalloc(c, &luf, 25 * (size_t)c->nc)) return rc;
      c->s.luf = luf;
    }
  }
  if (p->sgs_iter < 0) return fail(c, FPSI_ERR_ARG, "gauss_seidel_iter < 0");
  if ((p->sweep_mode == FPSI_SWEEP_PERSISTENT || p->sweep_mode == FPSI_SWEEP_PERSISTENT_FLAGS) && !c->have_flow)
    return fail(c, FPSI_ERR_ARG, "no dataflow sweep on this grid (a sweep chunk has more than 64 dependencies)");
  if (p->sweep_mode != FPSI_SWEEP_PERSISTENT && p->sweep_mode != FPSI_SWEEP_PER_LEVEL &&
      p->sweep_mode != FPSI_SWEEP_PERSISTENT_FLAGS &&
      p->sweep_mode != FPSI_SWEEP_TILED)
    return fail(c, FPSI_ERR_ARG, "unknown sweep mode");
  if (p->sweep_mode == FPSI_SWEEP_TILED && !c->have_tiles)
    return fail(c, FPSI_ERR_ARG, "no tiled sweep on this grid: " + c->tile_msg);
  if (!(p->gamma > 1.0) || !(p->Rtilde > 0.0)) return fail(c, FPSI_ERR_ARG, "bad gas constants");
  if (p->time_mode != FPSI_TIME_STEADY && p->time_mode != FPSI_TIME_BDF2 && p->time_mode != FPSI_TIME_KEC)
    return fail(c, FPSI_ERR_ARG, "unknown time mode");
  if (p->time_mode != FPSI_TIME_STEADY) {
    if (!(p->dtmax > 0.0) || !std::isfinite(p->dtmax))
      return fail(c, FPSI_ERR_ARG, "a time-accurate scheme needs a finite time step dtmax > 0");
    if (!c->bdf2_alloc) {
      const size_t n = 7 * (size_t)c->nc;  // room for every turbulence model's rows
      if (int rc = dalloc(c, &c->s.qn, n)) return rc;
      if (int rc = dalloc(c, &c->s.qm1, n)) return rc;
      if (int rc = dalloc(c, &c->s.rhs, 5 * (size_t)c->nc)) return rc;
      if (int rc = dalloc(c, &c->s.pbias, (size_t)c->nc)) return rc;
      c->bdf2_alloc = true;
    }
  }
  {  // boundary parameter sets; fixedMass's massFlux = mdot / boundary area (src/bcFixedMass.loci:230-232)
    double h[FPSI_MAX_BC_SETS * 8];
    for (int i = 0; i < FPSI_MAX_BC_SETS; ++i) {
      const fpsi_bc_set &b = p->bc[i];
      double *o = h + 8 * i;
      o[0] = b.T; o[1] = b.pg; o[2] = b.u[0]; o[3] = b.u[1]; o[4] = b.u[2];
      o[5] = b.p0; o[6] = b.T0;
      o[7] = c->bc_area[i] > 0 ? b.mdot / c->bc_area[i] : 0.0;
    }
    if (!c->bcs)
      if (int rc = dalloc(c, &c->bcs, (size_t)FPSI_MAX_BC_SETS * 8)) return rc;
    HIPCHK(c, hipMemcpyAsync(c->bcs, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->prm = *p;
  c->newton_iter = p->newton_iter > 0 ? p->newton_iter : 6;
  Prm &k = c->kp;
  k.bcs = c->bcs;
  k.gamma = p->gamma; k.Rt = p->Rtilde; k.Pamb = p->Pambient;
  k.flux = p->flux; k.limiter = p->limiter; k.Kl = p->Kl; k.viscous = p->viscous;
  k.a1 = p->suth[0]; k.a2 = p->suth[1]; k.a3 = p->suth[2];
  k.k1 = p->suth[3]; k.k2 = p->suth[4]; k.k3 = p->suth[5];
  k.Twall = p->Twall; k.pg0 = p->pg0; k.T0 = p->T0;
  k.u0x = p->u0[0]; k.u0y = p->u0[1]; k.u0z = p->u0[2];
  k.cflmax = p->cflmax; k.dtmax = p->dtmax; k.urelax = p->urelax; k.minPg = p->minPg;
  {  // Mlim2, src/fluxInviscid.loci:98-104; unset (<= 0) options take flowPsi's defaults
    const double Minf = p->Minf > 0 ? p->Minf : 1.0, etaT = p->etaT > 0 ? p->etaT : 0.25;
    const double M2 = Minf * Minf;
    double ml2 = 1.0;
    if (Minf < 0.5) ml2 = etaT * 2. * M2 / (1. - 2. * M2);
    k.Mlim2 = ml2;
    k.PLimPC = p->PLimPC > 0 ? p->PLimPC : 10.0;
    k.pre = ml2 < 1.0 ? 1 : 0;
  }
  // Spa_All_param defaults (turbulence/SA/SA_param.h:25-34), Cp
  // (src/solverEoS.loci:76-78), turbulentPrandtlNumber 0.9
  // (src/fluxViscous.loci:268-270), inflow nu_t 5e-9 (SA_turb.loci:167-174)
  k.turb = p->turb == FPSI_TURB_SST ? 2 : p->turb == FPSI_TURB_SA ? 1 : 0;
  k.sa_cb1 = 0.1355; k.sa_k = 0.41; k.sa_cv1 = 7.1; k.sa_cb2 = 0.622;
  k.sa_cw2 = 0.3; k.sa_cw3 = 2.0; k.sa_sigma = 2. / 3.;
  k.Cp = p->Rtilde * p->gamma / (p->gamma - 1.0);
  k.Prt = p->Prt > 0 ? p->Prt : 0.9;
  k.nut_bc = p->nut_bc > 0 ? p->nut_bc : 5e-9;
  {  // sst_k_bc / sst_w_bc without options (turbulence/KOmega/sst_viscous.loci:63-118)
    const double v = std::sqrt(p->u0[0] * p->u0[0] + p->u0[1] * p->u0[1] + p->u0[2] * p->u0[2]);
    const double kd = std::max(1.5 * (1e-2 * v) * (1e-2 * v), 1e-3);
    k.k_bc = p->k_bc > 0 ? p->k_bc : kd;
    k.w_bc = p->w_bc > 0 ? p->w_bc : 0.09 * kd / (1e-3 * 1e-5);
  }
  k.unsteady = p->time_mode == FPSI_TIME_KEC ? 2 : p->time_mode == FPSI_TIME_BDF2 ? 1 : 0;
  k.pseudo_cflmax = p->pseudo_cflmax > 0 ? p->pseudo_cflmax : 10.0;
  k.pbias_ctl = p->pbias_control;
  k.tdamp = p->temporal_damping;
  c->fp32 = p->precision == FPSI_PREC_FP32;
  c->have_params = true;
  c->assembled = false;
  return 0;
}

extern "C" int fpsi_set_state(fpsi_ctx *c, const double *q) {
  if (!c || !q) return FPSI_ERR_ARG;
  const int nc = c->nc;
  if (!c->have_params) return fail(c, FPSI_ERR_STATE, "fpsi_set_params must precede fpsi_set_state");
  const int nv = c->nvar();
  std::vector<double> h((size_t)nv * nc);
  for (int k = 0; k < nv; ++k)
    for (int ni = 0; ni < nc; ++ni) h[(size_t)k * nc + ni] = q[(size_t)k * nc + c->cperm[ni]];
  HIPCHK(c, hipMemcpyAsync(c->s.q, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_state = true;
  c->tol_sweeps_last = 0;  // a new run: tolerance solves check from the first sweep
  c->assembled = false;
  c->have_prev = c->in_step = false;  // q{n-1} = q until set otherwise
  return 0;
}

static int copy_cells_out(fpsi_ctx *c, const double *dsrc, int ncomp, double *out) {
  const int nc = c->nc;
  std::vector<double> h((size_t)ncomp * nc);
  HIPCHK(c, hipMemcpyAsync(h.data(), dsrc, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < ncomp; ++k)
    for (int ni = 0; ni < nc; ++ni) out[(size_t)k * nc + c->cperm[ni]] = h[(size_t)k * nc + ni];
  return 0;
}

static int copy_faces_out(fpsi_ctx *c, const double *dsrc, int ncomp, double *out) {
  const int nf = c->nf;
  std::vector<double> h((size_t)ncomp * nf);
  HIPCHK(c, hipMemcpyAsync(h.data(), dsrc, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < ncomp; ++k)
    for (int i = 0; i < nf; ++i) out[(size_t)k * nf + c->fperm[i]] = h[(size_t)k * nf + i];
  return 0;
}

extern "C" int fpsi_get_state(fpsi_ctx *c, double *q) {
  if (!c || !q) return FPSI_ERR_ARG;
  if (!c->have_state) return fail(c, FPSI_ERR_STATE, "no state set");
  return copy_cells_out(c, c->s.q, c->nvar(), q);
}

static int check_ready(fpsi_ctx *c) {
  if (!c->have_params) return fail(c, FPSI_ERR_STATE, "fpsi_set_params not called");
  if (!c->have_state) return fail(c, FPSI_ERR_STATE, "fpsi_set_state not called");
  if (c->prm.turb != FPSI_TURB_NONE && !c->have_dist)
    return fail(c, FPSI_ERR_STATE, "the turbulence model needs fpsi_set_wall_distance");
  return 0;
}

extern "C" int fpsi_set_wall_distance(fpsi_ctx *c, const double *dist) {
  if (!c || !dist) return FPSI_ERR_ARG;
  std::vector<double> h(c->nc);
  for (int ni = 0; ni < c->nc; ++ni) {
    h[ni] = dist[c->cperm[ni]];
    if (!(h[ni] > 0)) return fail(c, FPSI_ERR_ARG, "wall distance must be positive");
  }
  if (!c->g.dist) {
    double *d = nullptr;
    if (int rc = dalloc(c, &d, (size_t)c->nc)) return rc;
    c->g.dist = d;
  }
  HIPCHK(c, hipMemcpy((void *)c->g.dist, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  c->have_dist = true;
  return 0;
}

static void enqueue_residual(fpsi_ctx *c) {
  hipEvent_t e = prof_begin(c);
  launch_bface(c->g, c->kp, c->s, c->stream);
  launch_grad(c->g, c->kp, c->s, c->stream);
  launch_tmu(c->g, c->kp, c->s, c->stream);  // SST's eddy viscosity needs the velocity gradient
  launch_eta(c->g, c->kp, c->s, c->stream);
  launch_face_flux(c->g, c->kp, c->s, c->stream);
  launch_gather(c->g, c->kp, c->vtot, c->s, c->stream);
  prof_end(c, FPSI_K_RESIDUAL, e);
}

static void norms_from(const double *o, fpsi_norms *n) {
  n->rtrms = o[0]; n->etrms = o[1]; n->mtrms = o[2];
  n->rtmax = o[3]; n->etmax = o[4]; n->mtmax = o[5];
  n->ccmaxr[0] = o[6]; n->ccmaxr[1] = o[7]; n->ccmaxr[2] = o[8];
  n->ccmaxe[0] = o[9]; n->ccmaxe[1] = o[10]; n->ccmaxe[2] = o[11];
}

extern "C" int fpsi_residual(fpsi_ctx *c, double *src, double *iflux, double *vflux, fpsi_norms *norms) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  enqueue_residual(c);
  HIPCHK(c, hipGetLastError());
  if (src) if (int rc = copy_cells_out(c, c->s.src, 5, src)) return rc;
  if (iflux) if (int rc = copy_faces_out(c, c->s.iflux, 5, iflux)) return rc;
  if (vflux) if (int rc = copy_faces_out(c, c->s.vflux, 4, vflux)) return rc;
  double o[12];
  HIPCHK(c, hipMemcpyAsync(o, c->s.nout, sizeof(o), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (norms) norms_from(o, norms);
  const double tot = o[0] + o[1] + o[2];
  if (tot != tot) return fail(c, FPSI_ERR_NAN, "NAN residual");
  return 0;
}

extern "C" int fpsi_gradients(fpsi_ctx *c, double *grad, double *phi) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  launch_bface(c->g, c->kp, c->s, c->stream);
  launch_grad(c->g, c->kp, c->s, c->stream);
  HIPCHK(c, hipGetLastError());
  if (grad) if (int rc = copy_cells_out(c, c->s.grad, 15, grad)) return rc;
  if (phi) if (int rc = copy_cells_out(c, c->s.phi, 5, phi)) return rc;
  return 0;
}

static void enqueue_assemble(fpsi_ctx *c) {
  enqueue_residual(c);
  hipEvent_t e = prof_begin(c);
  launch_face_jac(c->g, c->kp, c->fp32, c->s, c->stream);
  prof_end(c, FPSI_K_FACE_JAC, e);
  e = prof_begin(c);
  launch_cell_assemble(c->g, c->kp, c->fp32, c->line_order() ? 1 : 0, c->s, c->stream);
  if (c->kp.unsteady) launch_rhs_norms(c->g, c->vtot, c->s, c->stream);
  if (c->fsgs()) launch_fsgs_factor(c->g, c->fp32, c->s, c->stream);
  if (c->kp.turb) launch_sa_assemble(c->g, c->kp, c->nnz, c->s, c->stream);
  prof_end(c, FPSI_K_ASSEMBLE, e);
}

// scalar SGS sweeps (src/linearSolverSGS.loci:333-500) of turbulence
// scalars sc0 .. sc0 + ns - 1, all of them in each level's launch
static void enqueue_sa_sweeps(fpsi_ctx *c, int sc0, int ns, int nsweeps) {
  const int nlev = (int)c->lev_off.size() - 1;
  const int nall = c->nscal();
  for (int it = 0; it < nsweeps; ++it) {
    launch_ghost1(c->g, nall, c->s.sX, c->stream);
    for (int l = 0; l < nlev; ++l)
      launch_sa_level(c->g, sc0, ns, c->nnz, nullptr, c->lev_off[l], c->lev_off[l + 1], c->s, c->stream);
    launch_ghost1(c->g, nall, c->s.sX, c->stream);
    for (int l = nlev - 1; l >= 0; --l)
      launch_sa_level(c->g, sc0, ns, c->nnz, nullptr, c->lev_off[l], c->lev_off[l + 1], c->s, c->stream);
  }
}

// LSGSScalarSolve of turbulence scalar sc after the Jacobi start made by
// the assembly (src/linearSolverLSGS.loci:1156-1404): like enqueue_lsgs,
// with the scalar rows and each sweep's stopping test read back on the host.
static int enqueue_sa_lsgs(fpsi_ctx *c, int sc, int maxiter) {
  const int nl = (int)c->lsgs_lev_off.size() - 1;
  const fpsi_params &p = c->prm;
  launch_lsgs_init(c->s, p.lsgs_relax, c->stream);
  c->sa_lsgs_iters[sc] = 0;
  c->sa_lsgs_w[sc] = p.lsgs_relax;
  c->sa_lsgs_resid[sc] = 0;
  for (int it = 0; it < maxiter; ++it) {
    launch_ghost1(c->g, c->nscal(), c->s.sX, c->stream);
    for (int l = 0; l < nl; ++l)
      launch_sa_lsgs_level(c->g, sc, c->nnz, 0, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                           c->stream);
    for (int l = nl - 1; l >= 0; --l)
      launch_sa_lsgs_level(c->g, sc, c->nnz, 1, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                           c->stream);
    launch_sa_lsgs_relax(c->g, sc, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->lsgs_host, c->s.lst, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->sa_lsgs_w[sc] = c->lsgs_host[0];
    c->sa_lsgs_resid[sc] = c->lsgs_host[2];
    c->sa_lsgs_iters[sc] = (int)c->lsgs_host[3];
    if (c->lsgs_host[4] != 0.0) break;
  }
  return 0;
}

// solve turbulence scalar `which` (-1: all of them)
// FSGSScalarSolve (src/linearSolverFSGS.loci:1039-1200) of scalar sc: X
// from zero, nsweeps symmetric sweeps in the breadth-first order
static void enqueue_sa_fsgs(fpsi_ctx *c, int sc, int nsweeps) {
  const int nl = (int)c->lsgs_lev_off.size() - 1;
  const long xs = (long)c->nc + c->nper + 1;
  launch_zero_d(c->nc + (long)c->nper, c->s.sX + sc * xs, c->stream);
  for (int it = 0; it < nsweeps; ++it) {
    launch_ghost1(c->g, c->nscal(), c->s.sX, c->stream);
    for (int l = 0; l < nl; ++l)
      launch_sa_lsgs_level(c->g, sc, c->nnz, 0, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                           c->stream);
    for (int l = nl - 1; l >= 0; --l)
      launch_sa_lsgs_level(c->g, sc, c->nnz, 1, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                           c->stream);
  }
}

static int enqueue_sa_solve(fpsi_ctx *c, int which, int nsweeps) {
  const int sc0 = which < 0 ? 0 : which, ns = which < 0 ? c->nscal() : 1;
  if (c->fsgs()) {
    for (int sc = sc0; sc < sc0 + ns; ++sc) enqueue_sa_fsgs(c, sc, nsweeps);
    return 0;
  }
  if (c->lsgs()) {
    for (int sc = sc0; sc < sc0 + ns; ++sc)
      if (int rc = enqueue_sa_lsgs(c, sc, nsweeps)) return rc;
    return 0;
  }
  enqueue_sa_sweeps(c, sc0, ns, nsweeps);
  return 0;
}

static int check_scalar(fpsi_ctx *c, int which) {
  if (!c->assembled) return fail(c, FPSI_ERR_STATE, "system not assembled");
  if (c->nscal() == 0) return fail(c, FPSI_ERR_STATE, "no turbulence model selected");
  if (which < 0 || which >= c->nscal()) return fail(c, FPSI_ERR_ARG, "no such turbulence scalar");
  return 0;
}

extern "C" int fpsi_turb_lsgs_stats(fpsi_ctx *c, int which, int *iters, double *relax, double *resid) {
  if (!c) return FPSI_ERR_ARG;
  if (which < 0 || which > 1) return fail(c, FPSI_ERR_ARG, "no such turbulence scalar");
  if (iters) *iters = c->sa_lsgs_iters[which];
  if (relax) *relax = c->sa_lsgs_w[which];
  if (resid) *resid = c->sa_lsgs_resid[which];
  return 0;
}
extern "C" int fpsi_sa_lsgs_stats(fpsi_ctx *c, int *iters, double *relax, double *resid) {
  return fpsi_turb_lsgs_stats(c, 0, iters, relax, resid);
}

// scalar arrays for two scalars (nuTilde uses the first), saf and scell
// sized for SST
static int sa_alloc(fpsi_ctx *c) {
  if (c->sa_alloc) return 0;
  DevState &s = c->s;
  const size_t nc = c->nc, nf = c->nf, nnz = (size_t)std::max(c->nnz, 1), xs = nc + c->nper + 1;
  int rc = 0;
  if ((rc = dalloc(c, &s.tmu, nc)) || (rc = dalloc(c, &s.saf, 12 * nf)) || (rc = dalloc(c, &s.scell, 8 * nc)) ||
      (rc = dalloc(c, &s.soff, 2 * nnz)) || (rc = dalloc(c, &s.ssrc, 2 * nc)) ||
      (rc = dalloc(c, &s.sdinv, 2 * nc)) || (rc = dalloc(c, &s.sB, 2 * nc)) || (rc = dalloc(c, &s.sX, 2 * xs)) ||
      (rc = dalloc(c, &s.sX2, 2 * xs)) || (rc = dalloc(c, &s.sxt, 2 * nc)))
    return rc;
  HIPCHK(c, hipMemset(s.sX, 0, 2 * xs * sizeof(double)));
  HIPCHK(c, hipMemset(s.sX2, 0, 2 * xs * sizeof(double)));
  HIPCHK(c, hipMemset(s.soff, 0, 2 * nnz * sizeof(double)));
  c->sa_alloc = true;
  return 0;
}

extern "C" int fpsi_turb_get_system(fpsi_ctx *c, int which, double *src, double *fjp, double *fjm, double *D,
                                    double *B, double *tmu) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_scalar(c, which)) return rc;
  const int nc = c->nc, nf = c->nf;
  const size_t oc = (size_t)which * nc;
  if (src) if (int rc = copy_cells_out(c, c->s.ssrc + oc, 1, src)) return rc;
  if (tmu) if (int rc = copy_cells_out(c, c->s.tmu, 1, tmu)) return rc;
  if (D || B) {
    std::vector<double> di(nc), b(nc);
    HIPCHK(c, hipMemcpy(di.data(), c->s.sdinv + oc, nc * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(b.data(), c->s.sB + oc, nc * sizeof(double), hipMemcpyDeviceToHost));
    for (int ni = 0; ni < nc; ++ni) {
      if (D) D[c->cperm[ni]] = 1. / di[ni];  // the kernel keeps 1/D
      if (B) B[c->cperm[ni]] = b[ni];
    }
  }
  if (fjp || fjm) {
    // saf rows of fjp / fjm: 3 / 4 for nuTilde, 4 / 5 (k) and 9 / 10 (omega) for SST
    const int rp = c->nscal() == 2 ? 4 + 5 * which : 3;
    std::vector<double> h(12 * (size_t)nf);
    HIPCHK(c, hipMemcpy(h.data(), c->s.saf, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < nf; ++i) {
      const int f = c->fperm[i];
      const bool inner = i < c->nfi + c->nper;
      if (fjp) fjp[f] = inner ? h[(size_t)rp * nf + i] : 0.0;
      if (fjm) fjm[f] = inner ? h[(size_t)(rp + 1) * nf + i] : 0.0;
    }
  }
  return 0;
}
extern "C" int fpsi_sa_get_system(fpsi_ctx *c, double *src, double *fjp, double *fjm, double *D, double *B,
                                  double *tmu) {
  if (!c) return FPSI_ERR_ARG;
  if (c->prm.turb != FPSI_TURB_SA) return fail(c, FPSI_ERR_STATE, "the Spalart-Allmaras model is not selected");
  return fpsi_turb_get_system(c, 0, src, fjp, fjm, D, B, tmu);
}

extern "C" int fpsi_turb_solve(fpsi_ctx *c, int which, int nsweeps, double *dX) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_scalar(c, which)) return rc;
  if (nsweeps < 0) return fail(c, FPSI_ERR_ARG, "nsweeps < 0");
  if (int rc = enqueue_sa_solve(c, which, nsweeps)) {
    hipStreamSynchronize(c->stream);
    return rc;
  }
  HIPCHK(c, hipGetLastError());
  const size_t xs = (size_t)c->nc + c->nper + 1;
  if (dX) if (int rc = copy_cells_out(c, c->s.sX + which * xs, 1, dX)) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
extern "C" int fpsi_sa_solve(fpsi_ctx *c, int nsweeps, double *dX) {
  if (!c) return FPSI_ERR_ARG;
  if (c->prm.turb != FPSI_TURB_SA) return fail(c, FPSI_ERR_STATE, "the Spalart-Allmaras model is not selected");
  return fpsi_turb_solve(c, 0, nsweeps, dX);
}

// BDF2 time levels for the systems of the step in progress: q{n-1} = q
// unless set, q{n} = q at the step's start
static void enqueue_levels(fpsi_ctx *c) {
  const long n = (long)c->nvar() * c->nc;
  if (!c->have_prev) launch_copy_d(n, c->s.q, c->s.qm1, c->stream);
  if (!c->in_step) launch_copy_d(n, c->s.q, c->s.qn, c->stream);
  c->have_prev = c->in_step = true;
}

extern "C" int fpsi_assemble(fpsi_ctx *c) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  if (c->bdf2()) enqueue_levels(c);
  enqueue_assemble(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->assembled = true;
  return 0;
}

extern "C" int fpsi_get_system(fpsi_ctx *c, void *fjp, void *fjm, void *D, int *pivot, void *B, double *dtau) {
  if (!c) return FPSI_ERR_ARG;
  if (!c->assembled) return fail(c, FPSI_ERR_STATE, "system not assembled");
  const size_t e = c->fj_bytes();
  const int nc = c->nc, nf = c->nf, nfi = c->nfi, nper = c->nper;
  if (fjp || fjm) {
    std::vector<char> off((size_t)c->nnz * 25 * e), pb((size_t)nper * 25 * e);
    HIPCHK(c, hipMemcpy(off.data(), c->s.off, off.size(), hipMemcpyDeviceToHost));
    if (nper) HIPCHK(c, hipMemcpy(pb.data(), c->s.pblk, pb.size(), hipMemcpyDeviceToHost));
    std::vector<int> scl(nfi + nper), scr(nfi);
    HIPCHK(c, hipMemcpy(scl.data(), c->g.slot_cl, scl.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (nfi) HIPCHK(c, hipMemcpy(scr.data(), c->g.slot_cr, scr.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (fjp) std::memset(fjp, 0, (size_t)nf * 25 * e);
    if (fjm) std::memset(fjm, 0, (size_t)nf * 25 * e);
    for (int i = 0; i < nfi + nper; ++i) {
      const int f = c->fperm[i];
      if (fjm) std::memcpy((char *)fjm + (size_t)f * 25 * e, off.data() + (size_t)scl[i] * 25 * e, 25 * e);
      if (fjp) {
        char *dst = (char *)fjp + (size_t)f * 25 * e;
        if (i < nfi) {
          const char *srcb = off.data() + (size_t)scr[i] * 25 * e;
          for (int k = 0; k < 25; ++k) {  // stored as L = -fjp
            if (e == 8) { double v; std::memcpy(&v, srcb + 8 * k, 8); v = -v; std::memcpy(dst + 8 * k, &v, 8); }
            else { float v; std::memcpy(&v, srcb + 4 * k, 4); v = -v; std::memcpy(dst + 4 * k, &v, 4); }
          }
        } else {
          std::memcpy(dst, pb.data() + (size_t)(i - nfi) * 25 * e, 25 * e);
        }
      }
    }
  }
  if (D) {
    std::vector<char> h((size_t)nc * 25 * e);
    HIPCHK(c, hipMemcpy(h.data(), c->s.lu, h.size(), hipMemcpyDeviceToHost));
    for (int ni = 0; ni < nc; ++ni) std::memcpy((char *)D + (size_t)c->cperm[ni] * 25 * e, h.data() + (size_t)ni * 25 * e, 25 * e);
  }
  if (pivot) {
    std::vector<int> h(nc);
    HIPCHK(c, hipMemcpy(h.data(), c->s.piv, nc * sizeof(int), hipMemcpyDeviceToHost));
    for (int ni = 0; ni < nc; ++ni)
      for (int k = 0; k < 5; ++k) pivot[(size_t)c->cperm[ni] * 5 + k] = (h[ni] >> (3 * k)) & 7;
  }
  if (B) {
    std::vector<char> h((size_t)nc * 5 * e);
    HIPCHK(c, hipMemcpy(h.data(), c->s.B, h.size(), hipMemcpyDeviceToHost));
    for (int ni = 0; ni < nc; ++ni) std::memcpy((char *)B + (size_t)c->cperm[ni] * 5 * e, h.data() + (size_t)ni * 5 * e, 5 * e);
  }
  if (dtau) if (int rc = copy_cells_out(c, c->s.dtau, 1, dtau)) return rc;
  return 0;
}

// ns > 0 (dataflow mode only): the turbulence scalar systems are swept in
// the same launches as the fluid system
// check = false leaves the dataflow error word to the caller's next readback
static int enqueue_sweeps(fpsi_ctx *c, int nsweeps, int ns = 0, bool check = true) {
  const int nlev = (int)c->lev_off.size() - 1;
  if (c->tiled()) {
    for (int it = 0; it < nsweeps; ++it)
      for (int dir = 0; dir < 2; ++dir) {
        launch_ghost(c->g, c->s, c->stream);
        hipEvent_t e = prof_begin(c);
        const hipError_t he = launch_sgs_tiles(c->g, c->tg, c->fp32, dir, ++c->epoch, c->tile_groups, c->s, c->stream);
        prof_end(c, FPSI_K_SGS_SWEEP, e);
        if (he != hipSuccess) return fail(c, FPSI_ERR_HIP, std::string("tiled sweep launch: ") + hipGetErrorString(he));
      }
    return 0;
  }
  if (dataflow(c)) {
    const int spin = spin_limit(c);
    const bool flags = c->prm.sweep_mode == FPSI_SWEEP_PERSISTENT_FLAGS;
    if (!ns && !flags) {
      // fluid only: flag-free half-sweeps alternating between X and X2; an
      // even number of them leaves the result in X
      double *cur = c->s.X, *oth = c->s.X2;
      for (int it = 0; it < nsweeps; ++it)
        for (int dir = 0; dir < 2; ++dir) {
          launch_sweep_prep(c->g, it == 0 && dir == 0, cur, oth, c->stream);
          hipEvent_t e = prof_begin(c);
          const hipError_t he = launch_sgs_flow(c->g, c->fp32, dir, c->nchunks, c->chunk_lo, spin, c->s.serr,
                                                c->flow_waves, c->s, cur, oth, c->stream);
          prof_end(c, FPSI_K_SGS_SWEEP, e);
          if (he != hipSuccess) return fail(c, FPSI_ERR_HIP, std::string("cooperative sweep launch: ") + hipGetErrorString(he));
          std::swap(cur, oth);
        }
    }
    for (int it = 0; it < (ns || flags ? nsweeps : 0); ++it)
      for (int dir = 0; dir < 2; ++dir) {
        launch_ghost(c->g, c->s, c->stream);
        if (ns) launch_ghost1(c->g, c->nscal(), c->s.sX, c->stream);
        hipEvent_t e = prof_begin(c);
        const int epoch = ++c->flow_epoch;
        const hipError_t he =
            launch_sgs_sweep(c->g, c->fp32, dir, c->nchunks, c->chunk_lo, dir ? c->bdep_off : c->fdep_off,
                             dir ? c->bdep : c->fdep, c->s.cflag, epoch, spin, c->s.serr, c->sweep_waves,
                             ns, c->nnz, c->s, c->stream);
        prof_end(c, FPSI_K_SGS_SWEEP, e);
        if (he != hipSuccess) return fail(c, FPSI_ERR_HIP, std::string("cooperative sweep launch: ") + hipGetErrorString(he));
      }
    if (nsweeps > 0 && check) return check_sweep_error(c, "dataflow sweep");
    return 0;
  }
  for (int it = 0; it < nsweeps; ++it) {
    launch_ghost(c->g, c->s, c->stream);
    hipEvent_t e = prof_begin(c);
    for (int l = 0; l < nlev; ++l)
      launch_sgs_level(c->g, c->fp32, c->lev_off[l], c->lev_off[l + 1], c->s, c->stream);
    prof_end(c, FPSI_K_SGS_SWEEP, e);
    launch_ghost(c->g, c->s, c->stream);
    e = prof_begin(c);
    for (int l = nlev - 1; l >= 0; --l)
      launch_sgs_level(c->g, c->fp32, c->lev_off[l], c->lev_off[l + 1], c->s, c->stream);
    prof_end(c, FPSI_K_SGS_SWEEP, e);
  }
  return 0;
}

// LSGSBlockedSSolve after its line-Jacobi start (made by the assembly):
// relaxed symmetric sweeps in the breadth-first line order, periodic ghosts
// refreshed before each (src/linearSolverLSGS.loci:1738-1948).  The stopping
// test needs the update norm on the host, so each sweep ends with a 40-byte
// read-back.
static int enqueue_lsgs(fpsi_ctx *c, int maxiter) {
  const int nl = (int)c->lsgs_lev_off.size() - 1;
  const fpsi_params &p = c->prm;
  launch_lsgs_init(c->s, p.lsgs_relax, c->stream);
  c->lsgs_iters = 0;
  c->lsgs_w = p.lsgs_relax;
  c->lsgs_resid = 0;
  // one cooperative flag-free dataflow launch per half-sweep where the
  // levels allow (bitwise equal to the level launches), else one launch per
  // level
  const bool flow = c->prm.sweep_mode == FPSI_SWEEP_PERSISTENT && c->lsgs_flow;
  const bool flags = c->prm.sweep_mode == FPSI_SWEEP_PERSISTENT_FLAGS && c->lsgs_flow;
  if (flow) launch_lsgs_flow_fill(c->g, c->s, 1, 0, c->stream);
  for (int it = 0; it < maxiter; ++it) {
    launch_ghost(c->g, c->s, c->stream);
    for (int dir = 0; dir < 2; ++dir) {
      hipEvent_t e = prof_begin(c);
      if (flow) {
        const hipError_t he = launch_lsgs_flow(c->g, c->fp32, dir, c->nchunks, c->chunk_lo, spin_limit(c), c->s.serr,
                                               line_waves(c, c->lsgs_fwaves), 1, 0, 0, c->nnz, c->s, c->stream);
        prof_end(c, FPSI_K_SGS_SWEEP, e, sweep_bytes(c, true, 0, dir));
        if (he != hipSuccess)
          return fail(c, FPSI_ERR_HIP, std::string("cooperative line sweep launch: ") + hipGetErrorString(he));
        continue;
      } else if (flags) {
        const hipError_t he = launch_lsgs_sweep(c->g, c->fp32, 0, dir, c->nchunks, c->chunk_lo,
                                                dir ? c->bdep_off : c->fdep_off, dir ? c->bdep : c->fdep, c->s.cflag,
                                                ++c->flow_epoch, spin_limit(c), c->s.serr,
                                                line_waves(c, c->lsgs_waves), 1, 0, 0, c->nnz, c->s, c->stream);
        prof_end(c, FPSI_K_SGS_SWEEP, e, sweep_bytes(c, true, 0, dir));
        if (he != hipSuccess)
          return fail(c, FPSI_ERR_HIP, std::string("cooperative line sweep launch: ") + hipGetErrorString(he));
        continue;
      } else if (dir == 0) {
        for (int l = 0; l < nl; ++l)
          launch_lsgs_level(c->g, c->fp32, 0, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                            c->stream);
      } else {
        for (int l = nl - 1; l >= 0; --l)
          launch_lsgs_level(c->g, c->fp32, 1, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                            c->stream);
      }
      prof_end(c, FPSI_K_SGS_SWEEP, e);
    }
    if (flow) launch_lsgs_relax_flow(c->g, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream);
    else launch_lsgs_relax(c->g, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->lsgs_host, c->s.lst, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->lsgs_w = c->lsgs_host[0];
    c->lsgs_resid = c->lsgs_host[2];
    c->lsgs_iters = (int)c->lsgs_host[3];
    if (c->lsgs_host[4] != 0.0) break;
  }
  if (flow || flags) return check_sweep_error(c, "dataflow line sweep");
  return 0;
}

// Line SGS of the fluid and of the turbulence scalars in the same dataflow
// launches.  Each system keeps its own loop exactly as when solved alone
// (ghost refresh, forward, backward, relaxation and stopping test with its
// own state slot), so the bits are those of the separate solves; a system
// leaves the launches once its test stops it.
// Stopping rule of the fluid's line-SGS loop run to a linear-residual
// tolerance (in place of its update-norm test): stop once ||B - A X|| <=
// rtol r0 or after max_sweeps relaxed symmetric sweeps.
// The residual costs about a quarter of a symmetric sweep, so it is first
// evaluated after check_from sweeps (one fewer than the previous solve of
// the context needed; a solve that meets the tolerance earlier runs those
// extra sweeps).
struct LinTol {
  double rtol, r0;
  int max_sweeps;
  int sweeps;
  double rfinal;
  int check_from;
};

static int linres(fpsi_ctx *c, double *out);

static int enqueue_lsgs_fused(fpsi_ctx *c, int maxiter, LinTol *tol = nullptr, bool scalars = true) {
  const fpsi_params &p = c->prm;
  const int ns = scalars ? c->nscal() : 0;
  for (int slot = 0; slot <= ns; ++slot) launch_lsgs_init(c->s, p.lsgs_relax, c->stream, slot);
  c->lsgs_iters = 0;
  c->lsgs_w = p.lsgs_relax;
  c->lsgs_resid = 0;
  for (int sc = 0; sc < ns; ++sc) {
    c->sa_lsgs_iters[sc] = 0;
    c->sa_lsgs_w[sc] = p.lsgs_relax;
    c->sa_lsgs_resid[sc] = 0;
  }
  bool fluid = !tol || (tol->max_sweeps > 0 && !(tol->rfinal <= tol->rtol * tol->r0));
  const int fmax = tol ? tol->max_sweeps : maxiter;
  int mask = (1 << ns) - 1;
  const bool flags = c->prm.sweep_mode == FPSI_SWEEP_PERSISTENT_FLAGS;
  if (!flags) launch_lsgs_flow_fill(c->g, c->s, 1, ns, c->stream);
  for (int it = 0; (fluid && it < fmax) || (mask && it < maxiter); ++it) {
    if (it >= maxiter) mask = 0;
    if (it >= fmax) fluid = false;
    if (fluid) launch_ghost(c->g, c->s, c->stream);
    if (mask) launch_ghost1(c->g, ns, c->s.sX, c->stream);
    int nact = 0;
    for (int sc = 0; sc < ns; ++sc) nact += mask >> sc & 1;
    for (int dir = 0; dir < 2; ++dir) {
      hipEvent_t e = prof_begin(c);
      const hipError_t he =
          flags ? launch_lsgs_sweep(c->g, c->fp32, 0, dir, c->nchunks, c->chunk_lo, dir ? c->bdep_off : c->fdep_off,
                                    dir ? c->bdep : c->fdep, c->s.cflag, ++c->flow_epoch, spin_limit(c), c->s.serr,
                                    line_waves(c, c->lsgs_waves), fluid ? 1 : 0, ns, mask, c->nnz, c->s, c->stream)
                : launch_lsgs_flow(c->g, c->fp32, dir, c->nchunks, c->chunk_lo, spin_limit(c), c->s.serr,
                                   line_waves(c, c->lsgs_fwaves), fluid ? 1 : 0, ns, mask, c->nnz, c->s, c->stream);
      prof_end(c, FPSI_K_SGS_SWEEP, e, sweep_bytes(c, fluid, nact, dir));
      if (he != hipSuccess)
        return fail(c, FPSI_ERR_HIP, std::string("cooperative line sweep launch: ") + hipGetErrorString(he));
    }
    if (fluid) {
      if (flags) launch_lsgs_relax(c->g, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream);
      else launch_lsgs_relax_flow(c->g, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream);
    }
    for (int sc = 0; sc < ns; ++sc)
      if (mask >> sc & 1) {
        if (flags) launch_sa_lsgs_relax(c->g, sc, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream, 1 + sc);
        else launch_sa_lsgs_relax_flow(c->g, sc, it, p.lsgs_abstol, p.lsgs_reltol, c->s, c->stream, 1 + sc);
      }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->lsgs_host, c->s.lst, 8 * (1 + ns) * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (fluid) {
      c->lsgs_w = c->lsgs_host[0];
      c->lsgs_resid = c->lsgs_host[2];
      c->lsgs_iters = (int)c->lsgs_host[3];
      if (tol) {
        ++tol->sweeps;
        if (tol->sweeps >= tol->check_from || tol->sweeps >= fmax) {
          if (int rc = linres(c, &tol->rfinal)) return rc;
          if (tol->rfinal <= tol->rtol * tol->r0) fluid = false;
        }
      } else if (c->lsgs_host[4] != 0.0) {
        fluid = false;
      }
    }
    for (int sc = 0; sc < ns; ++sc)
      if (mask >> sc & 1) {
        const double *h = c->lsgs_host + 8 * (1 + sc);
        c->sa_lsgs_w[sc] = h[0];
        c->sa_lsgs_resid[sc] = h[2];
        c->sa_lsgs_iters[sc] = (int)h[3];
        if (h[4] != 0.0) mask &= ~(1 << sc);
      }
  }
  return check_sweep_error(c, "dataflow line sweep");
}

// FSGSBlockedSSolve (src/linearSolverFSGS.loci:1598-1708): X from zero, then
// nsweeps symmetric sweeps in the breadth-first line order with the
// unpivoted Jacobian-precision factors, periodic ghosts refreshed before each.
static int enqueue_fsgs(fpsi_ctx *c, int nsweeps) {
  const int nl = (int)c->lsgs_lev_off.size() - 1;
  const bool flow = dataflow(c) && c->lsgs_flow;
  launch_zero_d(5 * ((long)c->nc + c->nper + 1), c->s.X, c->stream);
  for (int it = 0; it < nsweeps; ++it) {
    launch_ghost(c->g, c->s, c->stream);
    for (int dir = 0; dir < 2; ++dir) {
      hipEvent_t e = prof_begin(c);
      if (flow) {
        const hipError_t he = launch_lsgs_sweep(c->g, c->fp32, 1, dir, c->nchunks, c->chunk_lo,
                                                dir ? c->bdep_off : c->fdep_off, dir ? c->bdep : c->fdep, c->s.cflag,
                                                ++c->flow_epoch, spin_limit(c), c->s.serr,
                                                line_waves(c, c->lsgs_waves), 1, 0, 0, c->nnz,
                                                c->s, c->stream);
        if (he != hipSuccess)
          return fail(c, FPSI_ERR_HIP, std::string("cooperative FSGS sweep launch: ") + hipGetErrorString(he));
      } else if (dir == 0) {
        for (int l = 0; l < nl; ++l)
          launch_fsgs_level(c->g, c->fp32, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                            c->stream);
      } else {
        for (int l = nl - 1; l >= 0; --l)
          launch_fsgs_level(c->g, c->fp32, c->lsgs_cells, c->lsgs_lev_off[l], c->lsgs_lev_off[l + 1], c->s,
                            c->stream);
      }
      prof_end(c, FPSI_K_SGS_SWEEP, e);
    }
  }
  if (flow && nsweeps > 0) return check_sweep_error(c, "dataflow FSGS sweep");
  return 0;
}

extern "C" int fpsi_lsgs_stats(fpsi_ctx *c, int *iters, double *relax, double *resid) {
  if (!c) return FPSI_ERR_ARG;
  if (iters) *iters = c->lsgs_iters;
  if (relax) *relax = c->lsgs_w;
  if (resid) *resid = c->lsgs_resid;
  return 0;
}

static int enqueue_solve(fpsi_ctx *c, int nsweeps) {
  if (c->fsgs()) return enqueue_fsgs(c, nsweeps);
  return c->lsgs() ? enqueue_lsgs(c, nsweeps) : enqueue_sweeps(c, nsweeps);
}

static int need_sgs(fpsi_ctx *c) {
  if (c->line_order()) return fail(c, FPSI_ERR_STATE, "this call needs linear_solver = FPSI_SOLVER_SGS");
  return 0;
}

// the tolerance-driven solve and the learned guess: block SGS, or line SGS
// run as the dataflow sweep (its fused loop carries the tolerance test)
static int need_tol_solver(fpsi_ctx *c) {
  if (c->fsgs()) return fail(c, FPSI_ERR_STATE, "this call needs linear_solver = FPSI_SOLVER_SGS or FPSI_SOLVER_LSGS");
  if (c->lsgs() && !(dataflow(c) && c->lsgs_flow))
    return fail(c, FPSI_ERR_STATE, "a line-SGS solve to a tolerance needs the dataflow sweep on a grid whose "
                                   "breadth-first levels are its Gauss-Seidel levels");
  return 0;
}

extern "C" int fpsi_linear_solve(fpsi_ctx *c, int nsweeps, double *dq) {
  if (!c) return FPSI_ERR_ARG;
  if (!c->assembled) return fail(c, FPSI_ERR_STATE, "system not assembled");
  if (nsweeps < 0) return fail(c, FPSI_ERR_ARG, "nsweeps < 0");
  if (int rc = enqueue_solve(c, nsweeps)) return rc;
  HIPCHK(c, hipGetLastError());
  if (dq) {
    const int nc = c->nc;
    std::vector<double> h(5 * (size_t)nc);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->s.X, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int ni = 0; ni < nc; ++ni)
      for (int k = 0; k < 5; ++k) dq[(size_t)c->cperm[ni] * 5 + k] = h[(size_t)ni * 5 + k];
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->assembled = false;  // X now holds the solution, not the Jacobi start
  return 0;
}

extern "C" int fpsi_update(fpsi_ctx *c) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  launch_update(c->g, c->kp, c->s, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->assembled = false;
  return 0;
}

// one Newton iteration after the norms of its residual went to dh: linear
// solves of the fluid and the turbulence scalars, then the updates
static int enqueue_iteration(fpsi_ctx *c, double *dh) {
  enqueue_assemble(c);
  launch_copy_d(12, c->s.nout, dh, c->stream);
  // turbulence scalars ride in the fluid's dataflow sweep when both use SGS,
  // or line SGS on levels the dataflow covers
  const bool persistent = dataflow(c);
  const bool fuse = c->kp.turb && !c->line_order() && persistent;
  const bool lfuse = c->kp.turb && c->lsgs() && persistent && c->lsgs_flow;
  if (lfuse) {
    if (int rc = enqueue_lsgs_fused(c, c->prm.lsgs_maxiter)) return rc;
  } else if (int rc = fuse ? enqueue_sweeps(c, c->prm.sgs_iter, c->nscal())
                           : enqueue_solve(c, c->lsgs() ? c->prm.lsgs_maxiter : c->prm.sgs_iter)) {
    return rc;
  }
  if (c->kp.turb && !fuse && !lfuse) {
    hipEvent_t e2 = prof_begin(c);
    if (int rc = enqueue_sa_solve(c, -1, c->lsgs() ? c->prm.lsgs_maxiter : c->prm.sgs_iter)) return rc;
    prof_end(c, FPSI_K_SGS_SWEEP, e2);
  }
  hipEvent_t e = prof_begin(c);
  launch_update(c->g, c->kp, c->s, c->stream);
  if (c->kp.turb) launch_sa_update(c->g, c->kp, c->s, c->stream);
  prof_end(c, FPSI_K_UPDATE, e);
  return 0;
}

// end of a time step: q{n+1} = the last iterate or (KEC) its reconstruction,
// then q{n-1} = q{n}
static void enqueue_finish_step(fpsi_ctx *c) {
  if (c->kp.unsteady == 2) launch_kec_final(c->g, c->kp, c->s, c->stream);
  launch_copy_d((long)c->nvar() * c->nc, c->s.qn, c->s.qm1, c->stream);
  c->in_step = false;
}

// n Newton iterations (per time step when levels: then q{n,0} = q{n} and the
// levels shift after the step's `levels` iterations); hist[n] nullable
static int enqueue_iterations(fpsi_ctx *c, int n, int levels, double *dh) {
  for (int s = 0; s < n; ++s) {
    if (levels && s % levels == 0) {  // q{n,0} = q{n}; q{n-1} = q unless set
      c->in_step = false;
      enqueue_levels(c);
    }
    if (int rc = enqueue_iteration(c, dh + 12 * (size_t)s)) return rc;
    if (levels && (s + 1) % levels == 0) enqueue_finish_step(c);
  }
  return 0;
}

static int run_iterations(fpsi_ctx *c, int n, int levels, fpsi_norms *hist) {
  std::vector<double> h(12 * (size_t)std::max(n, 1));
  double *dh = nullptr;
  if (n > 0) HIPCHK(c, hipMallocAsync((void **)&dh, h.size() * sizeof(double), c->stream));
  if (int rc = enqueue_iterations(c, n, levels, dh)) {
    hipStreamSynchronize(c->stream);
    hipFree(dh);
    return rc;
  }
  HIPCHK(c, hipGetLastError());
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(h.data(), dh, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipFreeAsync(dh, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->assembled = false;
  for (int s = 0; s < n; ++s) {
    if (hist) norms_from(&h[12 * (size_t)s], &hist[s]);
    const double tot = h[12 * s] + h[12 * s + 1] + h[12 * s + 2];
    if (tot != tot) return fail(c, FPSI_ERR_NAN, "NAN residual at Newton iteration " + std::to_string(s));
  }
  return 0;
}

extern "C" int fpsi_newton_step(fpsi_ctx *c, int nsteps, fpsi_norms *hist) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  if (nsteps < 0) return fail(c, FPSI_ERR_ARG, "nsteps < 0");
  if (c->bdf2()) return fail(c, FPSI_ERR_STATE, "time-accurate time_mode: use fpsi_unsteady_step");
  return run_iterations(c, nsteps, 0, hist);
}

// fpsi_newton_step with the history in this device's `scratch` copy: the
// copy kernel of each iteration writes all 12 norms of its step before they
// are read, so nothing there needs clearing between calls; the history
// comes back 64 steps at a time through the pinned hist_host
extern "C" int fpsi_newton_step_noalloc(fpsi_ctx *c, int nsteps, fpsi_norms *hist) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  if (nsteps < 0) return fail(c, FPSI_ERR_ARG, "nsteps < 0");
  if (nsteps > FPSI_NOALLOC_MAX_STEPS)
    return fail(c, FPSI_ERR_ARG, "nsteps above FPSI_NOALLOC_MAX_STEPS (4096): the norm history would not fit");
  if (c->bdf2()) return fail(c, FPSI_ERR_STATE, "time-accurate time_mode: use fpsi_unsteady_step");
  if (c->prof) return fail(c, FPSI_ERR_STATE, "profiling creates events on demand: disable it for the no-allocation step");
  if (!c->scratch) return fail(c, FPSI_ERR_STATE, "norm-history array not resolved for this device");
  if (int rc = enqueue_iterations(c, nsteps, 0, c->scratch)) {
    hipStreamSynchronize(c->stream);
    return rc;
  }
  HIPCHK(c, hipGetLastError());
  c->assembled = false;
  for (int s0 = 0; s0 < nsteps; s0 += 64) {
    const int m = std::min(64, nsteps - s0);
    HIPCHK(c, hipMemcpyAsync(c->hist_host, c->scratch + 12 * (size_t)s0, 12 * (size_t)m * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < m; ++k) {
      const double *h = c->hist_host + 12 * k;
      if (hist) norms_from(h, &hist[s0 + k]);
      if (h[0] + h[1] + h[2] != h[0] + h[1] + h[2])
        return fail(c, FPSI_ERR_NAN, "NAN residual at Newton iteration " + std::to_string(s0 + k));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fpsi_unsteady_step(fpsi_ctx *c, int nsteps, fpsi_norms *hist) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  if (nsteps < 0) return fail(c, FPSI_ERR_ARG, "nsteps < 0");
  if (!c->bdf2()) return fail(c, FPSI_ERR_STATE, "fpsi_unsteady_step needs a time-accurate time_mode");
  return run_iterations(c, nsteps * c->newton_iter, c->newton_iter, hist);
}

static int check_levels(fpsi_ctx *c) {
  if (!c->have_params || !c->bdf2()) return fail(c, FPSI_ERR_STATE, "time levels need a time-accurate time_mode");
  if (!c->have_state) return fail(c, FPSI_ERR_STATE, "fpsi_set_state not called");
  return 0;
}

extern "C" int fpsi_finish_time_step(fpsi_ctx *c) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_levels(c)) return rc;
  if (!c->in_step) return fail(c, FPSI_ERR_STATE, "no time step in progress (fpsi_assemble starts one)");
  enqueue_finish_step(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->assembled = false;
  return 0;
}

extern "C" int fpsi_set_time_levels(fpsi_ctx *c, const double *qn, const double *qnm1) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_levels(c)) return rc;
  const int nc = c->nc, nv = c->nvar();
  std::vector<double> h((size_t)nv * nc);
  for (int lev = 0; lev < 2; ++lev) {
    const double *src = lev == 0 ? qn : qnm1;
    double *dst = lev == 0 ? c->s.qn : c->s.qm1;
    if (src) {
      for (int k = 0; k < nv; ++k)
        for (int ni = 0; ni < nc; ++ni) h[(size_t)k * nc + ni] = src[(size_t)k * nc + c->cperm[ni]];
      HIPCHK(c, hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    } else if (lev == 0 || !c->have_prev) {
      launch_copy_d((long)nv * nc, c->s.q, dst, c->stream);
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_prev = c->in_step = true;
  c->assembled = false;
  return 0;
}

extern "C" int fpsi_get_time_levels(fpsi_ctx *c, double *qn, double *qnm1) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_levels(c)) return rc;
  if (qn) if (int rc = copy_cells_out(c, c->in_step ? c->s.qn : c->s.q, c->nvar(), qn)) return rc;
  if (qnm1) if (int rc = copy_cells_out(c, c->have_prev ? c->s.qm1 : c->s.q, c->nvar(), qnm1)) return rc;
  return 0;
}

extern "C" int fpsi_synchronize(fpsi_ctx *c) {
  if (!c) return FPSI_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fpsi_profile_enable(fpsi_ctx *c, int on) {
  if (!c) return FPSI_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &p : c->ev_pending) { c->ev_pool.push_back(p.start); c->ev_pool.push_back(p.end); }
  c->ev_pending.clear();
  for (int k = 0; k < FPSI_K_NGROUPS; ++k) { c->prof_ms[k] = 0; c->prof_n[k] = 0; c->prof_b[k] = 0; c->prof_nb[k] = 0; }
  c->prof = on != 0;
  return 0;
}

extern "C" int fpsi_profile_read(fpsi_ctx *c, fpsi_profile *out) {
  if (!c || !out) return FPSI_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &p : c->ev_pending) {
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, p.start, p.end));
    c->prof_ms[p.grp] += ms;
    c->prof_n[p.grp] += 1;
    if (p.bytes >= 0) { c->prof_b[p.grp] += p.bytes; c->prof_nb[p.grp] += 1; }
    c->ev_pool.push_back(p.start);
    c->ev_pool.push_back(p.end);
  }
  c->ev_pending.clear();
  // algorithmic bytes of one launch of each group (each array once, gathers
  // cache-perfect), from the buffer sizes of this context
  const double nc = c->nc, nf = c->nf, nb = c->nb, nfi = c->nfi, nper = c->nper, nnz = c->nnz;
  const double ps = 8, pj = (double)c->fj_bytes();
  const double nent = nfi * 2 + nper + nb;  // gather-list entries
  const double visc = c->prm.viscous ? 1 : 0;
  const double face_geo = 64 * nf;            // n, area, fc (56) + cl, rc (8)
  const double cell_rec = (5 + 15 + 5 + 3 + 1) * ps * nc;  // q, grad, phi, cc, vol
  double b[FPSI_K_NGROUPS];
  b[FPSI_K_RESIDUAL] =
      /* bface */ nb * (5 * ps + 24 + 4 + 4) + nb * 5 * ps +
      /* grad  */ nc * (5 * ps + 3 * ps + 6 * ps + 4 + ps) + 4 * nent + 24 * nf + 5 * ps * nb + nc * 20 * ps +
      /* flux  */ face_geo + cell_rec + 5 * ps * nb + nf * (5 + 4 * visc + 1) * ps +
      /* gather*/ 4 * nc + 4 * nent + nf * (5 + 4 * visc) * ps + nc * (ps + 4) + 5 * ps * nc;
  b[FPSI_K_FACE_JAC] = face_geo + cell_rec + 5 * ps * nb + visc * 4 * ps * nf + 4 * (2 * nfi + nper) +
                       (2 * nfi + 2 * nper + nb) * 25 * pj;
  b[FPSI_K_ASSEMBLE] = 4 * nc + 4 * nent + 16 * nf + (5 + 5 + 1) * ps * nc + 4 * nc + 4 * nnz +
                       (nnz + nper + nb) * 25 * pj + nc * (ps + 25 * pj + 4 + 5 * pj + 5 * ps);
  b[FPSI_K_SGS_SWEEP] = nnz * (25 * pj + 4) + nc * (4 + 25 * pj + 4 + 5 * pj + 2 * 5 * ps);
  b[FPSI_K_UPDATE] = nc * (5 * ps + 2 * 5 * ps);
  for (int k = 0; k < FPSI_K_NGROUPS; ++k) {
    out->ms[k] = c->prof_ms[k];
    out->launches[k] = c->prof_n[k];
    // launches that carried their own byte count (the line-SGS sweeps, whose
    // systems leave the launches as their tests stop them) report its mean
    out->bytes[k] = c->prof_nb[k] == c->prof_n[k] && c->prof_n[k] > 0 ? c->prof_b[k] / c->prof_n[k] : b[k];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// linear residual, tolerance-driven solve, learned guess
// ---------------------------------------------------------------------------
// the sweep error word travels with the residual, so sweeps enqueued
// without their own check (solve_tol) are checked here
static int linres(fpsi_ctx *c, double *out) {
  launch_ghost(c->g, c->s, c->stream);  // the periodic couplings read current ghosts
  launch_linres(c->g, c->fp32, c->line_order(), c->s, c->lr_parts, c->lr_bsum, c->lr_out, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->rb_host + 12, c->lr_out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->rb_host + 13, c->s.serr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (*(const int *)(c->rb_host + 13)) return sweep_failed(c, "dataflow sweep");
  *out = std::sqrt(c->rb_host[12]);
  return 0;
}

extern "C" int fpsi_linear_residual(fpsi_ctx *c, double *norm) {
  if (!c || !norm) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  return linres(c, norm);
}

// Feature scales of the current system: RMS of the Jacobi start and of the
// right-hand side per equation, mean log pivot.  The linear solve is
// equivariant under scaling of B, so normalising per system keeps the
// network's inputs and targets in range as the Newton residual falls.
static int mlp_set_norm(fpsi_ctx *c) {
  MlpBufs m = c->mb;
  const int nc = c->nc;
  float sums[32];
  mlp_feature_stats(m, nc, m.loss, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(sums, m.loss, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float nrm[32];
  for (int k = 0; k < 32; ++k) nrm[k] = 0.f;
  for (int k = 0; k < 10; ++k) nrm[k] = std::sqrt(sums[k] / (float)nc) + 1e-30f;
  for (int k = 10; k < 15; ++k) nrm[k] = sums[k] / (float)nc;
  for (int k = 16; k < 26; ++k) nrm[k] = std::sqrt(sums[k] / (float)nc) + 1e-30f;
  HIPCHK(c, hipMemcpyAsync(m.norm, nrm, sizeof(nrm), hipMemcpyHostToDevice, c->stream));
  c->mlp_norm = true;
  return 0;
}

// raw features of the assembled system with X = the plain start: X0, the
// previous Newton update, the diagonal factors, and r0 = B - A X0 (into Dg)
// with its Jacobi correction
static int mlp_raw(fpsi_ctx *c) {
  const bool line = c->line_order();
  launch_linres(c->g, c->fp32, line, c->s, c->lr_parts, c->lr_bsum, c->lr_out, c->stream, c->Dg);
  mlp_features(c->mb, c->nc, c->fp32, c->s.X, c->mlp_xprev, line ? (const void *)c->s.lu64 : c->s.lu, line ? 1 : 0,
               line ? nullptr : c->s.piv, c->Dg, c->stream);
  HIPCHK(c, hipGetLastError());
  return 0;
}

static int mlp_guess(fpsi_ctx *c) {
  MlpBufs m = c->mb;
  const int nc = c->nc;
  if (int rc = mlp_raw(c)) return rc;
  if (int rc = mlp_set_norm(c)) return rc;
  mlp_normalise(m, nc, c->mlp_xcur, c->stream);
  m.x = c->mlp_xcur;
  mlp_infer_guess(m, nc, c->s.X, c->stream);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// scalars: also solve the turbulence scalars with their own solver (a
// Newton step).  Under line SGS their line-SGS loops run in the fluid's
// launches; under block SGS with the dataflow sweep their
// gauss_seidel_iter sweeps ride in the fluid's first sweeps (the rest, if
// the fluid stops earlier, follow on their own), otherwise they are swept
// after the fluid.  Either way the scalars' bits are those of their
// separate solve.
static int solve_tol(fpsi_ctx *c, double rtol, int max_sweeps, int use_guess, fpsi_lin_stats *st,
                     bool scalars = false) {
  fpsi_lin_stats s{};
  double rcur = 0;
  if (int rc = linres(c, &s.r0)) return rc;
  rcur = s.r0;
  if (use_guess) {
    if (!c->mlp_alloc) return fail(c, FPSI_ERR_STATE, "learned guess requested before fpsi_mlp_init");
    const long n = 5 * ((long)c->nc + c->nper + 1);
    launch_copy_d(n, c->s.X, c->Xj, c->stream);
    if (int rc = mlp_guess(c)) return rc;
    // step along the learned correction D = X_guess - X0 by the alpha that
    // minimises |B - A (X0 + alpha D)| in the residual norm: never worse
    // than the plain start
    launch_sub_d(5 * (long)c->nc, c->s.X, c->Xj, c->Dg, c->stream);
    launch_ghost_of(c->g, c->Dg, c->stream);
    launch_linres_pair(c->g, c->fp32, c->line_order(), c->s, c->Xj, c->Dg, c->lr_parts, c->lr_bsum, c->lr_out,
                       c->stream);
    double h3[3] = {0, 0, 0};
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(h3, c->lr_out, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s.alpha = h3[2] > 0 ? h3[1] / h3[2] : 0.0;
    launch_axpy_d(5 * (long)c->nc, c->Xj, s.alpha, c->Dg, c->s.X, c->stream);
    if (int rc = linres(c, &s.rguess)) return rc;
    if (s.rguess < s.r0) {
      s.used_guess = 1;
      rcur = s.rguess;
    } else {
      launch_copy_d(n, c->Xj, c->s.X, c->stream);  // fall back to the plain start
    }
  }
  const int check_from = std::max(1, c->tol_sweeps_last - 1);
  if (c->line_order()) {
    LinTol t{rtol, s.r0, max_sweeps, 0, rcur, check_from};
    if (int rc = enqueue_lsgs_fused(c, c->prm.lsgs_maxiter, &t, scalars)) return rc;
    s.sweeps = t.sweeps;
    rcur = t.rfinal;
  } else {
    const bool turb = scalars && c->kp.turb;
    const int ns = turb && dataflow(c) && !c->tiled() ? c->nscal() : 0;
    int srem = ns ? c->prm.sgs_iter : 0;  // scalar sweeps still to ride along
    while (rcur > rtol * s.r0 && s.sweeps < max_sweeps) {
      // linres checks the sweeps' error word
      const int k = std::max(1, std::min(check_from - s.sweeps, max_sweeps - s.sweeps));
      const int kf = std::min(k, srem);
      if (kf)
        if (int rc = enqueue_sweeps(c, kf, ns, false)) return rc;
      if (k > kf)
        if (int rc = enqueue_sweeps(c, k - kf, 0, false)) return rc;
      srem -= kf;
      s.sweeps += k;
      if (int rc = linres(c, &rcur)) return rc;
    }
    if (srem) enqueue_sa_sweeps(c, 0, ns, srem);
    if (turb && !ns)
      if (int rc = enqueue_sa_solve(c, -1, c->prm.sgs_iter)) return rc;
  }
  c->tol_sweeps_last = s.sweeps;
  s.rfinal = rcur;
  if (c->mlp_alloc) launch_copy_d(5 * (long)c->nc, c->s.X, c->mlp_xprev, c->stream);
  if (st) *st = s;
  return 0;
}

extern "C" int fpsi_linear_solve_tol(fpsi_ctx *c, double rtol, int max_sweeps, int use_guess, fpsi_lin_stats *st) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = need_tol_solver(c)) return rc;
  if (!c->assembled) return fail(c, FPSI_ERR_STATE, "system not assembled");
  const int rc = solve_tol(c, rtol, max_sweeps, use_guess, st);
  c->assembled = false;
  return rc;
}

extern "C" int fpsi_newton_adaptive(fpsi_ctx *c, int nsteps, double rtol, int max_sweeps, int use_guess,
                                    fpsi_norms *hist, fpsi_lin_stats *lin) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = check_ready(c)) return rc;
  if (int rc = need_tol_solver(c)) return rc;
  if (c->bdf2()) return fail(c, FPSI_ERR_STATE, "time-accurate time_mode: use fpsi_unsteady_step");
  for (int s = 0; s < nsteps; ++s) {
    enqueue_assemble(c);
    HIPCHK(c, hipGetLastError());
    // the norms arrive with the linear solve's first residual readback; a
    // NaN step is refused before its update
    HIPCHK(c, hipMemcpyAsync(c->rb_host, c->s.nout, 12 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    fpsi_lin_stats st{};
    if (int rc = solve_tol(c, rtol, max_sweeps, use_guess, &st, true)) return rc;
    const double *o = c->rb_host;
    if (hist) norms_from(o, &hist[s]);
    if (o[0] + o[1] + o[2] != o[0] + o[1] + o[2]) return fail(c, FPSI_ERR_NAN, "NAN residual at Newton step " + std::to_string(s));
    if (lin) lin[s] = st;
    launch_update(c->g, c->kp, c->s, c->stream);
    if (c->kp.turb) launch_sa_update(c->g, c->kp, c->s, c->stream);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->assembled = false;
  return 0;
}

// splitmix64 + Box-Muller: deterministic initial weights
static unsigned long long sm64(unsigned long long &s) {
  unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double gauss(unsigned long long &s) {
  const double u1 = ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

extern "C" int fpsi_mlp_num_params(void) { return mlp_layout().n; }

extern "C" int fpsi_mlp_init(fpsi_ctx *c, unsigned long long seed) {
  if (!c) return FPSI_ERR_ARG;
  const MlpLayout L = mlp_layout();
  const size_t nc = c->nc;
  MlpBufs &m = c->mb;
  if (!c->mlp_alloc) {
    int rc = 0;
    const int np = mlp_parts((int)nc);
    const int *d_off = nullptr, *d_idx = nullptr;
    const float *d_geo = nullptr;
    std::vector<unsigned char> mask(L.n, 1);
    for (int k = 5; k < 16; ++k) {
      for (int h = 0; h < 64; ++h) mask[L.w3 + k * 64 + h] = 0;
      mask[L.b3 + k] = 0;
    }
    const unsigned char *d_mask = nullptr;
    std::vector<float> zero64(64, 0.f);
    const float *d_zero = nullptr;
    if ((rc = upload(c, &d_off, c->mlp_nb_off)) || (rc = upload(c, &d_idx, c->mlp_nb_idx)) ||
        (rc = upload(c, &d_geo, c->mlp_geo)) || (rc = upload(c, &d_mask, mask)) || (rc = upload(c, &d_zero, zero64)) ||
        (rc = dalloc(c, &m.raw, 32 * nc)) || (rc = dalloc(c, &m.norm, 32)) || (rc = dalloc(c, &m.h0, 64 * nc)) ||
        (rc = dalloc(c, &m.a1, 64 * nc)) || (rc = dalloc(c, &m.h1, 64 * nc)) || (rc = dalloc(c, &m.a2, 64 * nc)) ||
        (rc = dalloc(c, &m.h2, 64 * nc)) || (rc = dalloc(c, &m.out, 16 * nc)) || (rc = dalloc(c, &m.dout, 32 * nc)) ||
        (rc = dalloc(c, &m.g0, 64 * nc)) || (rc = dalloc(c, &m.g1, 64 * nc)) || (rc = dalloc(c, &m.g2, 64 * nc)) ||
        (rc = dalloc(c, &m.da, 64 * nc)) || (rc = dalloc(c, &m.dcat, 128 * nc)) || (rc = dalloc(c, &m.w, (size_t)L.n)) ||
        (rc = dalloc(c, &m.grad, (size_t)L.n)) || (rc = dalloc(c, &m.am, (size_t)L.n)) ||
        (rc = dalloc(c, &m.av, (size_t)L.n)) || (rc = dalloc(c, &m.wb, (size_t)L.n)) ||
        (rc = dalloc(c, &m.wt1, 64 * 128)) || (rc = dalloc(c, &m.wt2, 64 * 128)) || (rc = dalloc(c, &m.wt3, 64 * 32)) ||
        (rc = dalloc(c, &m.wpart, (size_t)np * 64 * 128)) || (rc = dalloc(c, &m.bpart, (size_t)np * 64)) ||
        (rc = dalloc(c, &m.lpart, (size_t)std::max(mlp_parts((int)nc), std::max(norm_parts(c->g), 1)) * 32 + 32)) ||
        (rc = dalloc(c, &m.loss, 32)) ||
        (rc = dalloc(c, &c->mlp_xcur, 32 * nc)) || (rc = dalloc(c, &c->mlp_xprev, 5 * nc)))
      return rc;
    m.nb_off = d_off;
    m.nb_idx = d_idx;
    m.geo = d_geo;
    m.mask = d_mask;
    m.zero64 = d_zero;
    c->mlp_alloc = true;
  }
  // He-normal weights, zero biases; the output layer starts small so the
  // first guess is close to the Jacobi start
  std::vector<float> w(L.n, 0.f);
  unsigned long long st = seed ? seed : 0x5eedull;
  auto fill = [&](int off, int rows, int cols, double scale) {
    for (int i = 0; i < rows * cols; ++i) w[off + i] = (float)(scale * gauss(st));
  };
  fill(L.w0, 64, 32, std::sqrt(2.0 / 29.0));  // 29 live input features
  fill(L.w1, 64, 128, std::sqrt(2.0 / 128.0));
  fill(L.w2, 64, 128, std::sqrt(2.0 / 128.0));
  fill(L.w3, 5, 64, 0.1 * std::sqrt(1.0 / 64.0));
  HIPCHK(c, hipMemcpyAsync(m.w, w.data(), L.n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(m.am, 0, L.n * sizeof(float), c->stream));
  launch_zero_d(5 * (long)nc, c->mlp_xprev, c->stream);
  HIPCHK(c, hipMemsetAsync(m.av, 0, L.n * sizeof(float), c->stream));
  mlp_refresh_bf16(m, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mlp_norm = false;
  c->mlp_nsamp = 0;
  c->mlp_next = 0;
  c->mlp_step = 0;
  return 0;
}

extern "C" int fpsi_mlp_get_params(fpsi_ctx *c, float *p) {
  if (!c || !p) return FPSI_ERR_ARG;
  if (!c->mlp_alloc) return fail(c, FPSI_ERR_STATE, "fpsi_mlp_init not called");
  HIPCHK(c, hipMemcpy(p, c->mb.w, mlp_layout().n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int fpsi_mlp_set_params(fpsi_ctx *c, const float *p) {
  if (!c || !p) return FPSI_ERR_ARG;
  if (!c->mlp_alloc) return fail(c, FPSI_ERR_STATE, "fpsi_mlp_init not called");
  HIPCHK(c, hipMemcpy(c->mb.w, p, mlp_layout().n * sizeof(float), hipMemcpyHostToDevice));
  mlp_refresh_bf16(c->mb, c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int fpsi_mlp_harvest(fpsi_ctx *c, int target_sweeps, int *slot) {
  if (!c) return FPSI_ERR_ARG;
  if (int rc = need_tol_solver(c)) return rc;
  if (!c->mlp_alloc) return fail(c, FPSI_ERR_STATE, "fpsi_mlp_init not called");
  if (!c->assembled) return fail(c, FPSI_ERR_STATE, "system not assembled");
  const int nc = c->nc;
  MlpBufs m = c->mb;
  const int sl = c->mlp_next;
  if (!c->mlp_xs[sl]) {  // sample slots are allocated as they are first filled
    if (int rc = dalloc(c, &c->mlp_xs[sl], 32 * (size_t)nc)) return rc;
    if (int rc = dalloc(c, &c->mlp_ys[sl], 8 * (size_t)nc)) return rc;
  }
  if (int rc = mlp_raw(c)) return rc;
  if (int rc = mlp_set_norm(c)) return rc;
  mlp_normalise(m, nc, c->mlp_xs[sl], c->stream);
  // the target: the solver's own iterate after target_sweeps sweeps
  if (c->line_order()) {
    LinTol t{0.0, 1.0, target_sweeps, 0, 1.0, target_sweeps};
    if (int rc = enqueue_lsgs_fused(c, c->prm.lsgs_maxiter, &t, false)) return rc;
  } else if (int rc = enqueue_sweeps(c, target_sweeps)) {
    return rc;
  }
  mlp_target(m, nc, c->s.X, c->mlp_ys[sl], c->stream);
  launch_copy_d(5 * (long)nc, c->s.X, c->mlp_xprev, c->stream);  // the next step's previous update
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mlp_next = (sl + 1) % fpsi_ctx::MLP_SLOTS;
  c->mlp_nsamp = std::min(c->mlp_nsamp + 1, (int)fpsi_ctx::MLP_SLOTS);
  c->assembled = false;
  if (slot) *slot = sl;
  return 0;
}

extern "C" int fpsi_mlp_train(fpsi_ctx *c, int epochs, float lr, double *loss) {
  if (!c) return FPSI_ERR_ARG;
  if (!c->mlp_alloc || c->mlp_nsamp == 0) return fail(c, FPSI_ERR_STATE, "no harvested samples");
  const int nc = c->nc;
  double last = 0;
  for (int e = 0; e < epochs; ++e) {
    last = 0;
    for (int s = 0; s < c->mlp_nsamp; ++s) {
      MlpBufs m = c->mb;
      m.x = c->mlp_xs[s];
      m.y = c->mlp_ys[s];
      mlp_forward(m, nc, c->stream);
      mlp_backward(m, nc, c->stream);
      mlp_adam(m, lr, ++c->mlp_step, c->stream);
      if (e == epochs - 1) {
        float l = 0;
        HIPCHK(c, hipMemcpyAsync(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        last += l / (5.0 * nc);
      }
    }
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = last / c->mlp_nsamp;
  return 0;
}

extern "C" int fpsi_mlp_eval(fpsi_ctx *c, int slot, float *out, float *grad, double *loss) {
  if (!c) return FPSI_ERR_ARG;
  if (!c->mlp_alloc || slot < 0 || slot >= c->mlp_nsamp) return fail(c, FPSI_ERR_ARG, "no such sample");
  const int nc = c->nc;
  MlpBufs m = c->mb;
  m.x = c->mlp_xs[slot];
  m.y = c->mlp_ys[slot];
  mlp_forward(m, nc, c->stream);
  mlp_backward(m, nc, c->stream);
  HIPCHK(c, hipGetLastError());
  if (out) {
    std::vector<float> h(16 * (size_t)nc);
    HIPCHK(c, hipMemcpyAsync(h.data(), m.out, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int ni = 0; ni < nc; ++ni) std::memcpy(out + 16 * (size_t)c->cperm[ni], &h[16 * (size_t)ni], 16 * sizeof(float));
  }
  if (grad) HIPCHK(c, hipMemcpy(grad, m.grad, mlp_layout().n * sizeof(float), hipMemcpyDeviceToHost));
  float l = 0;
  HIPCHK(c, hipMemcpy(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost));
  if (loss) *loss = l / (5.0 * nc);
  return 0;
}

extern "C" int fpsi_mlp_get_sample(fpsi_ctx *c, int slot, float *x, float *y, int *nb_off, int *nb_idx) {
  if (!c) return FPSI_ERR_ARG;
  if (!c->mlp_alloc || slot < 0 || slot >= c->mlp_nsamp) return fail(c, FPSI_ERR_ARG, "no such sample");
  const int nc = c->nc;
  if (x) {
    std::vector<float> h(32 * (size_t)nc);
    HIPCHK(c, hipMemcpy(h.data(), c->mlp_xs[slot], h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int ni = 0; ni < nc; ++ni) std::memcpy(x + 32 * (size_t)c->cperm[ni], &h[32 * (size_t)ni], 32 * sizeof(float));
  }
  if (y) {
    std::vector<float> h(8 * (size_t)nc);
    HIPCHK(c, hipMemcpy(h.data(), c->mlp_ys[slot], h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int ni = 0; ni < nc; ++ni) std::memcpy(y + 8 * (size_t)c->cperm[ni], &h[8 * (size_t)ni], 8 * sizeof(float));
  }
  if (nb_off) {
    nb_off[0] = 0;
    for (int i = 0; i < nc; ++i) {
      const int ni = c->icperm[i];
      nb_off[i + 1] = nb_off[i] + (c->mlp_nb_off[ni + 1] - c->mlp_nb_off[ni]);
    }
    if (nb_idx)
      for (int i = 0; i < nc; ++i) {
        const int ni = c->icperm[i];
        int o = nb_off[i];
        for (int j = c->mlp_nb_off[ni]; j < c->mlp_nb_off[ni + 1]; ++j) nb_idx[o++] = c->cperm[c->mlp_nb_idx[j]];
      }
  }
  return 0;
}
