// Tiled block Gauss-Seidel half-sweeps for gfx950.
//
// A level-per-launch sweep pays one kernel launch and two dependent memory
// round trips per dependency level (382 levels per half-sweep on a 128^3
// grid).  Here the cells are cut into tiles of at most TILE_MAX cells whose
// dependency graph is acyclic; one workgroup stages a tile's rows, factors
// and right-hand sides in LDS, waits for the tiles it reads from, and runs
// the tile's own levels out of LDS with workgroup barriers in between, so
// only tile boundaries cross workgroups.  Every cell is updated with the same
// arithmetic in the same order as in the level-per-launch kernel and the
// sequential sweep, and reads final values upstream and old values
// downstream exactly as they do, so the result is bitwise the same.
//
// One cooperative launch per half-sweep; workgroup w takes tiles w, w+W, ...
// in sweep order (uniform trip count), so the workgroup holding the smallest
// unfinished tile has finished all its earlier tiles and every tile it waits
// for is smaller: every wait ends.  Each tile publishes by storing its epoch
// into its flag after the device-coherent X stores of all its waves have
// completed (vmcnt = 0); readers poll the flags and read X with
// device-coherent loads.
#include "tiles.h"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "blocks.h"

namespace fpsi {

// ---------------------------------------------------------------------------
// host: tile construction
// ---------------------------------------------------------------------------
std::string build_tiles(int nc, const std::vector<double> &cx, const std::vector<double> &cy,
                        const std::vector<double> &cz, const std::vector<int> &lvl,
                        const std::vector<int> &m_off, const std::vector<int> &m_nbr,
                        const std::vector<int> &nbr_pad, TileHost &out) {
  out = TileHost{};
  for (int c = 0; c < nc; ++c)
    if (m_off[c + 1] - m_off[c] > MAXR) return "a row is wider than 6 blocks";
  // per axis: rank of each cell's centre coordinate among the distinct values
  const std::vector<double> *xs[3] = {&cx, &cy, &cz};
  std::vector<int> rank[3];
  int nd[3];
  for (int d = 0; d < 3; ++d) {
    std::vector<double> v(xs[d]->begin(), xs[d]->end());
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    nd[d] = (int)v.size();
    rank[d].resize(nc);
    for (int c = 0; c < nc; ++c) rank[d][c] = (int)(std::lower_bound(v.begin(), v.end(), (*xs[d])[c]) - v.begin());
  }
  // box extents: whole planes (e_d planes per box, product <= TILE_MAX) when
  // the centres form a lattice, else about 24 cells per box by quantiles
  int nt[3];
  const bool lattice = (long)nd[0] * nd[1] * nd[2] == (long)nc;
  int e[3] = {1, 1, 1};
  if (lattice) {
    while (e[0] * e[1] * e[2] * 2 <= TILE_MAX) {
      int best = 0;
      for (int d = 1; d < 3; ++d)
        if ((double)nd[d] / e[d] > (double)nd[best] / e[best]) best = d;
      if (nd[best] / e[best] < 2) break;
      e[best] *= 2;
    }
    for (int d = 0; d < 3; ++d) nt[d] = (nd[d] + e[d] - 1) / e[d];
  } else {
    const int k = std::max(1, (int)std::lround(std::cbrt(nc / 24.0)));
    nt[0] = nt[1] = nt[2] = k;
  }
  std::vector<long> box(nc);
  for (int c = 0; c < nc; ++c) {
    long b[3];
    for (int d = 0; d < 3; ++d)
      b[d] = lattice ? rank[d][c] / e[d] : (long)((double)rank[d][c] * nt[d] / nd[d]);
    box[c] = b[0] + (long)nt[0] * (b[1] + (long)nt[1] * b[2]);
  }
  // cells by (box, level, index); boxes cut into runs of <= TILE_MAX cells
  // (a cut along the level order never closes a cycle)
  std::vector<int> ord(nc);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) {
    if (box[a] != box[b]) return box[a] < box[b];
    if (lvl[a] != lvl[b]) return lvl[a] < lvl[b];
    return a < b;
  });
  std::vector<int> tile_of(nc), t_off{0};
  for (int k = 0; k < nc; ++k) {
    const int c = ord[k];
    const bool fresh = k == 0 || box[c] != box[ord[k - 1]] || k - t_off.back() == TILE_MAX;
    if (fresh && k > 0) t_off.push_back(k);
    tile_of[c] = (int)t_off.size() - 1;
  }
  t_off.push_back(nc);
  const int ntl = (int)t_off.size() - 1;
  // tile graph: an edge for every interior neighbour pair in different tiles,
  // from the lower level to the higher
  std::vector<std::pair<int, int>> edges;
  for (int c = 0; c < nc; ++c)
    for (int j = m_off[c]; j < m_off[c + 1]; ++j) {
      const int v = m_nbr[j];
      if (v >= nc || lvl[v] <= lvl[c] || tile_of[v] == tile_of[c]) continue;
      edges.push_back({tile_of[c], tile_of[v]});
    }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  std::vector<int> eo(ntl + 1, 0), indeg(ntl, 0);
  for (auto &p : edges) { eo[p.first + 1]++; indeg[p.second]++; }
  for (int t = 0; t < ntl; ++t) eo[t + 1] += eo[t];
  // longest-path level of each tile (Kahn); a leftover tile means a cycle
  std::vector<int> tl(ntl, 0), queue;
  for (int t = 0; t < ntl; ++t)
    if (indeg[t] == 0) queue.push_back(t);
  for (size_t h = 0; h < queue.size(); ++h) {
    const int t = queue[h];
    for (int k = eo[t]; k < eo[t + 1]; ++k) {
      const int u = edges[k].second;
      tl[u] = std::max(tl[u], tl[t] + 1);
      if (--indeg[u] == 0) queue.push_back(u);
    }
  }
  if ((int)queue.size() != ntl) return "the tile graph has a cycle";
  // sweep order: by tile level, then by tile
  std::vector<int> torder(ntl), tnew(ntl);
  std::iota(torder.begin(), torder.end(), 0);
  std::stable_sort(torder.begin(), torder.end(), [&](int a, int b) { return tl[a] < tl[b]; });
  for (int k = 0; k < ntl; ++k) tnew[torder[k]] = k;
  // device tables in sweep order
  out.ntiles = ntl;
  out.cell_off.assign(ntl + 1, 0);
  out.lev_off.assign(ntl + 1, 0);
  std::vector<int> local(nc);
  for (int k = 0; k < ntl; ++k) {
    const int t = torder[k];
    const int a = t_off[t], b = t_off[t + 1];
    out.cell_off[k + 1] = out.cell_off[k] + (b - a);
    for (int x = a; x < b; ++x) {
      local[ord[x]] = x - a;
      out.cells.push_back(ord[x]);
      out.row.push_back(m_off[ord[x]]);
      out.len.push_back(m_off[ord[x] + 1] - m_off[ord[x]]);
    }
    int li = -1;
    for (int x = a; x < b; ++x) {
      if (x == a || lvl[ord[x]] != lvl[ord[x - 1]]) {
        out.lev.push_back(x - a);
        ++li;
      }
      out.lvi.push_back(li);
    }
    out.lev.push_back(b - a);
    out.lev_off[k + 1] = (int)out.lev.size();
  }
  for (int k = 0; k < ntl; ++k) {
    const int t = torder[k];
    for (int x = t_off[t]; x < t_off[t + 1]; ++x) {
      const int c = ord[x];
      for (int j = 0; j < MAXR; ++j) {
        const int v = nbr_pad[(size_t)c * MAXR + j];
        out.nbr.push_back(v < nc && tile_of[v] == t ? ~local[v] : v);
      }
    }
  }
  std::vector<std::vector<int>> up(ntl), dn(ntl);
  for (auto &p : edges) {
    dn[tnew[p.first]].push_back(tnew[p.second]);
    up[tnew[p.second]].push_back(tnew[p.first]);
  }
  out.up_off.assign(ntl + 1, 0);
  out.dn_off.assign(ntl + 1, 0);
  for (int k = 0; k < ntl; ++k) {
    std::sort(up[k].begin(), up[k].end());
    std::sort(dn[k].begin(), dn[k].end());
    out.up.insert(out.up.end(), up[k].begin(), up[k].end());
    out.dn.insert(out.dn.end(), dn[k].begin(), dn[k].end());
    out.up_off[k + 1] = (int)out.up.size();
    out.dn_off[k + 1] = (int)out.dn.size();
  }
  return "";
}

// ---------------------------------------------------------------------------
// device: one half-sweep
// ---------------------------------------------------------------------------
static constexpr int TWAVES = 3;  // 3 x 12 cell slots >= TILE_MAX
static constexpr int TTHREADS = 64 * TWAVES;
static_assert(TWAVES * CPW >= TILE_MAX, "a tile level must fit one pass");

template <typename F>
__global__ void __launch_bounds__(TTHREADS)
    k_sgs_tiles(DevGrid g, TileGrid tg, int backward, int epoch, const F *__restrict__ off,
                const F *__restrict__ lu, const int *__restrict__ piv, const F *__restrict__ B,
                double *__restrict__ X) {
  // the five lanes of group q of wave w own tile cell k = 12 w + q for the
  // whole tile: its row stays in registers, its factors and the X values of
  // its neighbours in LDS
  __shared__ F slu[TILE_MAX * 25];
  __shared__ double sxl[TILE_MAX * 5];          // the tile's own X
  __shared__ double sxe[TILE_MAX * MAXR * 5];   // neighbours outside the tile
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int q = lane / 5, i = lane - 5 * q;
  const int k = wave * CPW + q;
  const int gbase = 5 * (q < CPW ? q : 0);
  const int zslot = g.nc + g.nper;  // the X entry that stays zero
  const int nt = tg.ntiles;
  for (int t = blockIdx.x; t < nt; t += gridDim.x) {
    const int tile = backward ? nt - 1 - t : t;
    const int c0 = tg.cell_off[tile];
    const int n = tg.cell_off[tile + 1] - c0;
    const int nl = tg.lev_off[tile + 1] - tg.lev_off[tile] - 1;
    const bool own = q < CPW && k < n;
    // the tiles this one waits for: wave 0 fetches their ids first, so the
    // poll does not start behind the staging loads
    const int *dl = backward ? tg.dn : tg.up;
    const int d0 = backward ? tg.dn_off[tile] : tg.up_off[tile];
    const int nd = (backward ? tg.dn_off[tile + 1] : tg.up_off[tile + 1]) - d0;
    const int dep0 = wave == 0 && lane < nd ? dl[d0 + lane] : -1;
    // rows (padded like the level kernel: the last block repeated against the
    // zero X entry), factors, right-hand side, old X, neighbour table and
    // level index; none of it depends on this half-sweep.  Lanes without a
    // cell load the tile's last cell and drop it, so every load is issued
    // unconditionally, back to back.
    const int x = c0 + (own ? k : n - 1);
    const int c = tg.cells[x];
    const int r0 = tg.row[x], len = tg.len[x];
    const int my_lv = tg.lvi[x];
    int nb[MAXR];
    F mrow[MAXR][5];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) {
      nb[j] = tg.nbr[(size_t)x * MAXR + j];
      const int slot = len > 0 ? r0 + (j < len ? j : len - 1) : 0;
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) mrow[j][jj] = off[(size_t)slot * 25 + i + 5 * jj];
    }
    F lrow[5];
#pragma unroll
    for (int jj = 0; jj < 5; ++jj) lrow[jj] = lu[(size_t)c * 25 + i + 5 * jj];
    const double bi = -(double)B[(size_t)c * 5 + i];
    const double xold = X[(size_t)c * 5 + i];
    const int pk = piv[c];
    if (own) {
#pragma unroll
      for (int jj = 0; jj < 5; ++jj) slu[k * 25 + i + 5 * jj] = lrow[jj];
      sxl[k * 5 + i] = xold;
    }
    // wait for every tile this one reads final values from: wave 0 polls
    // (one lane per tile, with a pause between rounds to keep the polling
    // traffic small), the other waves wait at the barrier
    if (wave == 0) {
      for (int base = 0; base < nd; base += 64) {
        const int d = base == 0 ? dep0 : (base + lane < nd ? dl[d0 + base + lane] : -1);
        while (true) {
          const bool ok =
              d < 0 || __hip_atomic_load(tg.flag + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epoch;
          if (__all(ok)) break;
          __builtin_amdgcn_s_sleep(8);
        }
      }
    }
    __syncthreads();
    // neighbours outside the tile (final upstream, old downstream, ghosts):
    // lane i fetches component i of each; in-tile entries read the zero entry
    double xe[MAXR];
#pragma unroll
    for (int j = 0; j < MAXR; ++j) xe[j] = ldx<true>(X + (size_t)(nb[j] >= 0 ? nb[j] : zslot) * 5 + i);
    if (own) {
#pragma unroll
      for (int j = 0; j < MAXR; ++j) sxe[(k * MAXR + j) * 5 + i] = xe[j];
    }
    __syncthreads();
    // the tile's own levels, in sweep order
    for (int li = 0; li < nl; ++li) {
      const int L = backward ? nl - 1 - li : li;
      const bool act = own && my_lv == L;
      double tt = 0.0;
      if (act) {
        tt = bi;
#pragma unroll
        for (int j = 0; j < MAXR; ++j) {
          const double *xn = nb[j] < 0 ? sxl + (~nb[j]) * 5 : sxe + (k * MAXR + j) * 5;
#pragma unroll
          for (int jj = 0; jj < 5; ++jj) tt += mrow[j][jj] * xn[jj];
        }
      }
      double v5[5];
#pragma unroll
      for (int kk = 0; kk < 5; ++kk) v5[kk] = -__shfl(tt, gbase + kk, 64);
      if (act) {
        double xs[5];
        lusolve5<F>(slu + k * 25, pk, v5, xs);
        const double xi = i == 0 ? xs[0] : i == 1 ? xs[1] : i == 2 ? xs[2] : i == 3 ? xs[3] : xs[4];
        sxl[k * 5 + i] = xi;
      }
      // LDS only inside the tile: a global store here would make every
      // barrier wait for its write-through
      __syncthreads();
    }
    if (own) stx<true>(X + (size_t)c * 5 + i, sxl[k * 5 + i]);
    // publish once every wave's X stores have completed
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (wave == 0) __hip_atomic_store(tg.flag + tile, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

int sgs_tiles_groups(int fp32, int ncu) {
  int per_cu = 0;
  hipError_t e = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgs_tiles<float>, TTHREADS, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgs_tiles<double>, TTHREADS, 0);
  if (e != hipSuccess || per_cu <= 0) return 0;
  return per_cu * ncu;
}

hipError_t launch_sgs_tiles(const DevGrid &g, const TileGrid &tg, int fp32, int backward, int epoch, int ngroups,
                            DevState &s, hipStream_t st) {
  DevGrid gg = g;
  TileGrid tt = tg;
  int bw = backward, ep = epoch;
  if (fp32) {
    const float *o = (const float *)s.off, *l = (const float *)s.lu, *b = (const float *)s.B;
    const int *pv = s.piv;
    double *x = s.X;
    void *args[] = {&gg, &tt, &bw, &ep, &o, &l, &pv, &b, &x};
    return hipLaunchCooperativeKernel((const void *)k_sgs_tiles<float>, dim3(ngroups), dim3(TTHREADS), args, 0, st);
  }
  const double *o = (const double *)s.off, *l = (const double *)s.lu, *b = (const double *)s.B;
  const int *pv = s.piv;
  double *x = s.X;
  void *args[] = {&gg, &tt, &bw, &ep, &o, &l, &pv, &b, &x};
  return hipLaunchCooperativeKernel((const void *)k_sgs_tiles<double>, dim3(ngroups), dim3(TTHREADS), args, 0, st);
}

}  // namespace fpsi
