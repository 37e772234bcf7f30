// tile_sort.hip -- per-segment sort of edges by source (see tile_sort.h).
#include "tile_sort.h"

#include <rocprim/block/block_radix_sort.hpp>

#include <algorithm>
#include <vector>

namespace gd {

namespace {

constexpr int kSortIPT = 16;  // keys per thread

// One workgroup of BS threads per segment [seg.x, seg.x + seg.y) of at most
// BS * kSortIPT edges: the sources, with their offsets in the segment as
// values, sorted by a stable LDS radix sort (rocPRIM's block radix sort; ties
// stay in edge order).  Thread t holds the segment's edges [t * kSortIPT,
// + kSortIPT); missing ones are padded with key 0xffffffff, which no source
// exceeds and which, the pads coming last in input order, stays after every
// real key in the stable sort.  seg.y < 0: the -seg.y edges stay in edge
// order.
template <int BS>
__global__ __launch_bounds__(BS) void segment_sort_kernel(const int2* __restrict__ segs,
                                                          const int32_t* __restrict__ col,
                                                          int32_t* __restrict__ col_sorted,
                                                          uint16_t* __restrict__ pos) {
  using Sort = rocprim::block_radix_sort<uint32_t, BS, kSortIPT, uint16_t>;
  __shared__ typename Sort::storage_type storage;
  const int2 sg = segs[blockIdx.x];
  const int e0 = sg.x, n = sg.y;
  if (n < 0) {
    for (int i = threadIdx.x; i < -n; i += BS) {
      col_sorted[e0 + i] = col[e0 + i];
      pos[e0 + i] = (uint16_t)i;
    }
    return;
  }
  uint32_t key[kSortIPT];
  uint16_t off[kSortIPT];
#pragma unroll
  for (int i = 0; i < kSortIPT; i++) {
    const int k = threadIdx.x * kSortIPT + i;
    key[i] = k < n ? (uint32_t)col[e0 + k] : 0xffffffffu;
    off[i] = (uint16_t)k;
  }
  // every key is read before the sort's first barrier; col may alias col_sorted
  Sort().sort(key, off, storage);
#pragma unroll
  for (int i = 0; i < kSortIPT; i++) {
    const int k = threadIdx.x * kSortIPT + i;
    if (k < n) {
      col_sorted[e0 + k] = (int32_t)key[i];
      pos[e0 + k] = off[i];
    }
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
  static_assert(1024 * kSortIPT == kSortMax, "the widest sort covers kSortMax edges");
  const dim3 grid((unsigned)segs.size());
  if (chunk > 8192)
    hipLaunchKernelGGL(segment_sort_kernel<1024>, grid, dim3(1024), 0, stream, d_segs.as<int2>(),
                       d_col, d_col_sorted, d_pos);
  else if (chunk > 4096)
    hipLaunchKernelGGL(segment_sort_kernel<512>, grid, dim3(512), 0, stream, d_segs.as<int2>(),
                       d_col, d_col_sorted, d_pos);
  else
    hipLaunchKernelGGL(segment_sort_kernel<256>, grid, dim3(256), 0, stream, d_segs.as<int2>(),
                       d_col, d_col_sorted, d_pos);
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipStreamSynchronize(stream));  // d_segs is freed on return
  return GD_OK;
}

}  // namespace gd
