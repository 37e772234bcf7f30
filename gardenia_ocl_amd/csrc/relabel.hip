// relabel.hip -- plan-time vertex relabelling by out-degree octave.
//
// Pull PageRank gathers contrib[u] once per out-edge of u.  On an unpermuted
// RMAT graph the most-gathered sources are spread over every cache line of
// contrib (hot ids are those with many zero bits), so the hot set does not fit
// in an XCD's 4 MB L2 and gathers are served from the Infinity Cache.  Giving
// vertices new ids in order of decreasing out-degree packs the hot set into a
// prefix of contrib: at RMAT-24 the first 4 MB (1 M vertices) receive 87% of
// all gathers.  The order is by degree octave (floor(log2(outdeg)),
// decreasing), and by original id inside an octave: degree-based grouping
// (Faldu et al., "A Closer Look at Lightweight Graph Reordering", IISWC
// 2019).  The hot set stays a prefix, and inside an octave the original
// order keeps the locality of the generator's id space (RMAT's recursive
// quadrants correlate a destination's id with its sources' ids), so a sorted
// tile's sources fall on fewer cache lines than under an exact degree
// sort.
//
// The relabelled in-CSR keeps every row's edges in their original order (only
// the ids change), so each row is still summed in the reference's order and
// short rows stay bit-identical.  Scores are permuted into the internal order
// at the start of a run and back at the end.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "gd_common.h"
#include "relabel.h"

namespace gd {

__global__ void iota_kernel(int32_t* __restrict__ a, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = i;
}

// key[v] = the octave of v's out-degree: floor(log2(outdeg)) + 1, 0 for none
__global__ void degree_octave_kernel(int m, const int32_t* __restrict__ outdeg,
                                     int32_t* __restrict__ key) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int d = outdeg[i];
    key[i] = d > 0 ? 32 - __clz(d) : 0;
  }
}

__global__ void relabel_rows_kernel(int m, const int32_t* __restrict__ perm,
                                    const int32_t* __restrict__ off,
                                    const int32_t* __restrict__ outdeg,
                                    int32_t* __restrict__ inv, int32_t* __restrict__ new_len,
                                    int32_t* __restrict__ new_outdeg) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
    const int v = perm[i];
    inv[v] = i;
    new_len[i] = off[v + 1] - off[v];
    new_outdeg[i] = outdeg[v];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) new_len[m] = 0;
}

// One wavefront per new row: copy the old row's edges, mapping ids through inv.
__global__ __launch_bounds__(256) void relabel_cols_kernel(int m, const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ inv,
                                                           const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ col,
                                                           const int32_t* __restrict__ new_off,
                                                           int32_t* __restrict__ new_col) {
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (blockDim.x / 64);
  for (int i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; i < m; i += waves) {
    const int v = perm[i];
    const int b = off[v], n = off[v + 1] - b, nb = new_off[i];
    for (int k = lane; k < n; k += 64) new_col[nb + k] = inv[col[b + k]];
  }
}

__global__ void gather_perm_kernel(int m, const int32_t* __restrict__ perm,
                                   const float* __restrict__ src, float* __restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

__global__ void scatter_perm_kernel(int m, const int32_t* __restrict__ perm,
                                    const float* __restrict__ src, float* __restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
    dst[perm[i]] = src[i];
}

static int grid_for(int n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

int build_relabel(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                  const int32_t* d_outdeg, hipStream_t stream, Relabel* r) {
  if (m <= 0) return GD_OK;
  int rc = r->perm.alloc(4 * (size_t)m);
  if (!rc) rc = r->inv.alloc(4 * (size_t)m);
  if (!rc) rc = r->off.alloc(4 * ((size_t)m + 1));
  if (!rc) rc = r->col.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (!rc) rc = r->outdeg.alloc(4 * (size_t)m);
  if (rc) return rc;
  DeviceBuffer ids, len, tmp, key;
  rc = ids.alloc(4 * (size_t)m);
  if (!rc) rc = key.alloc(4 * (size_t)m);
  if (!rc) rc = len.alloc(4 * ((size_t)m + 1));
  if (rc) return rc;
  // perm = ids sorted by out-degree octave, descending; the radix sort is
  // stable, so one octave keeps increasing ids.
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(m)), dim3(256), 0, stream, ids.as<int32_t>(), m);
  hipLaunchKernelGGL(degree_octave_kernel, dim3(grid_for(m)), dim3(256), 0, stream, m, d_outdeg,
                     key.as<int32_t>());
  size_t sort_bytes = 0, scan_bytes = 0;
  GD_HIP_TRY(rocprim::radix_sort_pairs_desc(nullptr, sort_bytes, key.as<int32_t>(),
                                             r->outdeg.as<int32_t>(), ids.as<int32_t>(),
                                             r->perm.as<int32_t>(), (size_t)m, 0, 6, stream));
  GD_HIP_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, len.as<int32_t>(),
                                     r->off.as<int32_t>(), 0, (size_t)m + 1,
                                     rocprim::plus<int32_t>(), stream));
  rc = tmp.alloc(std::max(sort_bytes, scan_bytes));
  if (rc) return rc;
  // (the sorted keys land in r->outdeg temporarily; overwritten below)
  GD_HIP_TRY(rocprim::radix_sort_pairs_desc(tmp.ptr, sort_bytes, key.as<int32_t>(),
                                             r->outdeg.as<int32_t>(), ids.as<int32_t>(),
                                             r->perm.as<int32_t>(), (size_t)m, 0, 6, stream));
  hipLaunchKernelGGL(relabel_rows_kernel, dim3(grid_for(m)), dim3(256), 0, stream, m,
                     r->perm.as<int32_t>(), d_off, d_outdeg, r->inv.as<int32_t>(),
                     len.as<int32_t>(), r->outdeg.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(rocprim::exclusive_scan(tmp.ptr, scan_bytes, len.as<int32_t>(),
                                     r->off.as<int32_t>(), 0, (size_t)m + 1,
                                     rocprim::plus<int32_t>(), stream));
  hipLaunchKernelGGL(relabel_cols_kernel, dim3(grid_for(m)), dim3(256), 0, stream, m,
                     r->perm.as<int32_t>(), r->inv.as<int32_t>(), d_off, d_col,
                     r->off.as<int32_t>(), r->col.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipStreamSynchronize(stream));
  return GD_OK;
}

int relabel_to_internal(const Relabel& r, int m, const float* user, float* internal,
                        hipStream_t stream) {
  hipLaunchKernelGGL(gather_perm_kernel, dim3(grid_for(m)), dim3(256), 0, stream, m,
                     r.perm.as<int32_t>(), user, internal);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

int relabel_to_user(const Relabel& r, int m, const float* internal, float* user,
                    hipStream_t stream) {
  hipLaunchKernelGGL(scatter_perm_kernel, dim3(grid_for(m)), dim3(256), 0, stream, m,
                     r.perm.as<int32_t>(), internal, user);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

}  // namespace gd
