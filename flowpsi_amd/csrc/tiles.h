// Tiled Gauss-Seidel half-sweeps (tiles.hip).
#pragma once
#include <string>
#include <vector>

#include "layout.h"

namespace fpsi {

static constexpr int TILE_MAX = 32;  // cells per tile

// Tiles in forward sweep order (a topological order of the tile graph).
struct TileGrid {
  int ntiles;
  const int *cell_off;  // [ntiles + 1] into cells
  const int *cells;     // cells of each tile, sorted by (level, index)
  const int *lev_off;   // [ntiles + 1] into lev
  const int *lev;       // per tile: local start of each of its levels, then its size
  const int *up_off, *up;  // tiles holding upstream neighbours (forward sweep)
  const int *dn_off, *dn;  // tiles holding downstream neighbours (backward sweep)
  const int *nbr;       // [cells][6]: >= 0 X index outside the tile, < 0 ~(local index)
  const int *row;       // [cells] first block slot of the cell's row
  const int *len;       // [cells] blocks in the row
  const int *lvi;       // [cells] index of the cell's level within its tile
  int *flag;            // [ntiles] last half-sweep (epoch) each tile finished
};

struct TileHost {
  int ntiles = 0;
  std::vector<int> cell_off, cells, lev_off, lev, up_off, up, dn_off, dn, nbr, row, len, lvi;
};

// Cut the cells into tiles of at most TILE_MAX cells: boxes of the cell
// centres (whole planes of a structured grid), each box split along its
// level order when larger, ordered topologically.  Returns "" on success or
// why the grid admits no tiling (rows wider than 6, or a cyclic tile graph).
std::string build_tiles(int nc, const std::vector<double> &cx, const std::vector<double> &cy,
                        const std::vector<double> &cz, const std::vector<int> &lvl,
                        const std::vector<int> &m_off, const std::vector<int> &m_nbr,
                        const std::vector<int> &nbr_pad, TileHost &out);

int sgs_tiles_groups(int fp32, int ncu);
hipError_t launch_sgs_tiles(const DevGrid &g, const TileGrid &tg, int fp32, int backward, int epoch, int ngroups,
                            DevState &s, hipStream_t st);

}  // namespace fpsi
