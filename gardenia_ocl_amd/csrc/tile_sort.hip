// tile_sort.hip -- per-segment sort of edges by source (see tile_sort.h).
#include "tile_sort.h"

#include <algorithm>
#include <vector>

namespace gd {

namespace {

constexpr int kSortBlock = 256;

// One workgroup per segment [seg.x, seg.x + seg.y): keys (source << 14 |
// offset) sorted by an LDS bitonic network, padded to a power of two with
// UINT64_MAX.  Sorting the 64-bit key orders by source, then by offset.
// seg.y < 0: the -seg.y edges stay in edge order.
// MAXN: the longest segment of the launch (LDS of MAXN keys).
template <int MAXN>
__global__ __launch_bounds__(kSortBlock) void segment_sort_kernel(const int2* __restrict__ segs,
                                                                  const int32_t* __restrict__ col,
                                                                  int32_t* __restrict__ col_sorted,
                                                                  uint16_t* __restrict__ pos) {
  __shared__ uint64_t key[MAXN];
  const int2 sg = segs[blockIdx.x];
  const int e0 = sg.x, n = sg.y;
  if (n < 0) {
    for (int i = threadIdx.x; i < -n; i += kSortBlock) {
      col_sorted[e0 + i] = col[e0 + i];
      pos[e0 + i] = (uint16_t)i;
    }
    return;
  }
  int np = 1;
  while (np < n) np <<= 1;
  for (int i = threadIdx.x; i < np; i += kSortBlock)
    key[i] = (i < n) ? (((uint64_t)(uint32_t)col[e0 + i] << 14) | (uint64_t)i) : ~0ull;
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += kSortBlock) {
        const int l = i ^ j;
        if (l > i) {
          const uint64_t a = key[i], b = key[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            key[i] = b;
            key[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // every thread read its keys before the last barrier; col may alias col_sorted
  for (int i = threadIdx.x; i < n; i += kSortBlock) {
    const uint64_t v = key[i];
    col_sorted[e0 + i] = (int32_t)(v >> 14);
    pos[e0 + i] = (uint16_t)(v & 0x3fffu);
  }
}

}  // namespace

int build_tile_sort(const Schedule& s, int chunk, const int32_t* d_col, int32_t* d_col_sorted,
                    uint16_t* d_pos, hipStream_t stream, bool exact_hubs) {
  if (chunk <= 0 || chunk > kSortMax) {
    set_error("build_tile_sort: chunk must be in [1, 16384]");
    return GD_ERR_INVALID;
  }
  std::vector<int2> segs;
  segs.reserve(s.tiles.size() + 16);
  for (const Tile& t : s.tiles) {
    const bool keep = exact_hubs && t.y < 0;  // whole-row exact hubs stay in edge order
    const int seg = chunk;
    for (int32_t e = t.z; e < t.w; e += seg) {
      const int32_t n = std::min(seg, t.w - e);
      if (n > kSortMax) {
        set_error("build_tile_sort: a tile is longer than 16384 edges");
        return GD_ERR_INVALID;
      }
      segs.push_back(make_int2(e, keep ? -n : n));
    }
  }
  if (segs.empty()) return GD_OK;
  DeviceBuffer d_segs;
  int rc = d_segs.alloc(sizeof(int2) * segs.size());
  if (rc) return rc;
  GD_HIP_TRY(hipMemcpyAsync(d_segs.ptr, segs.data(), sizeof(int2) * segs.size(),
                            hipMemcpyHostToDevice, stream));
  if (chunk > 8192)  // 128 KB of keys: one workgroup per CU
    hipLaunchKernelGGL(segment_sort_kernel<kSortMax>, dim3((unsigned)segs.size()),
                       dim3(kSortBlock), 0, stream, d_segs.as<int2>(), d_col, d_col_sorted, d_pos);
  else if (chunk > 4096)  // 64 KB of keys
    hipLaunchKernelGGL(segment_sort_kernel<8192>, dim3((unsigned)segs.size()), dim3(kSortBlock), 0,
                       stream, d_segs.as<int2>(), d_col, d_col_sorted, d_pos);
  else
    hipLaunchKernelGGL(segment_sort_kernel<4096>, dim3((unsigned)segs.size()), dim3(kSortBlock), 0,
                       stream, d_segs.as<int2>(), d_col, d_col_sorted, d_pos);
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipStreamSynchronize(stream));  // d_segs is freed on return
  return GD_OK;
}

}  // namespace gd
