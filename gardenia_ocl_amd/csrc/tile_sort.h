// tile_sort.h -- plan-time reordering of each tile's edges by source vertex.
//
// The pull kernels gather contrib[col[e]] for the edges of a tile, 64 lanes
// per instruction; every distinct cache line among the 64 addresses is one L2
// request.  Gathering a tile's edges in increasing source order makes lanes
// that read neighbouring sources share lines (the hot, relabelled sources
// especially), so fewer requests per edge.  Each sorted edge keeps its
// position inside the tile (16 bits), and the kernels write the gathered value
// to that position in LDS: every row is still summed in CSR order.
#ifndef GD_TILE_SORT_H_
#define GD_TILE_SORT_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gd_common.h"

namespace gd {

constexpr int kSortMax = 16384;  // longest sorted segment (one workgroup's LDS radix sort)

// Sorts the edges of every work segment of `s` by source: row tiles, and hub
// tiles in chunks of `chunk` edges (whole-row exact hubs) or as they are
// (hub segments).  col_sorted[e] / pos[e] receive, for the i-th edge of a
// segment in source order (ties kept in edge order), its source and its
// offset in the segment.  col and col_sorted may be the same array.
// exact_hubs (a schedule built with exact_len > 0, whose hub tiles are whole
// rows): hub rows stay in edge order (col_sorted = col and pos = offset
// there).
int build_tile_sort(const Schedule& s, int chunk, const int32_t* d_col, int32_t* d_col_sorted,
                    uint16_t* d_pos, hipStream_t stream, bool exact_hubs = false);

}  // namespace gd

#endif  // GD_TILE_SORT_H_
