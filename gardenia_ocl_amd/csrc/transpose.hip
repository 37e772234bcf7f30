// transpose.hip -- plan-time CSR -> CSC on the device, with edge positions.
//
// The adjoints pull over the transposed matrix (dx = A^T g; the PageRank
// adjoint iterates on the reversed graph).  A stable radix sort of the column
// indices, carrying the edge index, gives the CSC with the rows of each column
// in increasing order -- the same order a counting-sort transpose of the
// reference (graph_io.h:165-170 with transpose=true) produces -- and the
// permutation that moves per-edge values between the two orders.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "gd_common.h"
#include "transpose.h"

namespace gd {

__global__ void edge_rows_kernel(int m, const int32_t* __restrict__ off,
                                 int32_t* __restrict__ row_of_edge) {
  const int lane = threadIdx.x & 63;
  for (int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < m;
       r += gridDim.x * (blockDim.x / 64))
    for (int e = off[r] + lane; e < off[r + 1]; e += 64) row_of_edge[e] = r;
}

__global__ void iota64_kernel(int32_t* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (int32_t)i;
}

// col_count[c] = number of edges in column c (from the sorted keys), as offsets.
__global__ void csc_offsets_kernel(int64_t nnz, int ncols, const int32_t* __restrict__ sorted_col,
                                   int32_t* __restrict__ csc_off) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = sorted_col[i];
    const int cp = (i == 0) ? -1 : sorted_col[i - 1];
    for (int q = cp + 1; q <= c; q++) csc_off[q] = (int32_t)i;
    if (i == nnz - 1)
      for (int q = c + 1; q <= ncols; q++) csc_off[q] = (int32_t)nnz;
  }
}

__global__ void gather_rows_kernel(int64_t nnz, const int32_t* __restrict__ perm,
                                   const int32_t* __restrict__ row_of_edge,
                                   int32_t* __restrict__ csc_row) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x)
    csc_row[i] = row_of_edge[perm[i]];
}

__global__ void fill_int_kernel(int32_t* __restrict__ a, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

static int grid64(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

int build_transpose(int rows, int cols, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                    hipStream_t stream, Transpose* t) {
  int rc = t->off.alloc(4 * ((size_t)cols + 1));
  if (!rc) rc = t->row.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (!rc) rc = t->perm.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (rc) return rc;
  if (nnz == 0) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid64(cols + 1)), dim3(256), 0, stream,
                       t->off.as<int32_t>(), cols + 1, 0);
    GD_HIP_TRY(hipGetLastError());
    GD_HIP_TRY(hipStreamSynchronize(stream));
    return GD_OK;
  }
  DeviceBuffer roe, ids, keys, tmp;
  rc = roe.alloc(4 * (size_t)nnz);
  if (!rc) rc = ids.alloc(4 * (size_t)nnz);
  if (!rc) rc = keys.alloc(4 * (size_t)nnz);
  if (rc) return rc;
  hipLaunchKernelGGL(edge_rows_kernel, dim3(grid64(rows / 4 + 1)), dim3(256), 0, stream, rows,
                     d_off, roe.as<int32_t>());
  hipLaunchKernelGGL(iota64_kernel, dim3(grid64(nnz)), dim3(256), 0, stream, ids.as<int32_t>(),
                     nnz);
  GD_HIP_TRY(hipGetLastError());
  int end_bit = 1;
  while ((1ll << end_bit) < (int64_t)cols + 1 && end_bit < 32) end_bit++;
  size_t bytes = 0;
  GD_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, d_col, keys.as<int32_t>(),
                                       ids.as<int32_t>(), t->perm.as<int32_t>(), (size_t)nnz, 0,
                                       end_bit, stream));
  rc = tmp.alloc(bytes);
  if (rc) return rc;
  GD_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, bytes, d_col, keys.as<int32_t>(),
                                       ids.as<int32_t>(), t->perm.as<int32_t>(), (size_t)nnz, 0,
                                       end_bit, stream));
  // columns with no edges before the first key
  hipLaunchKernelGGL(fill_int_kernel, dim3(grid64(cols + 1)), dim3(256), 0, stream,
                     t->off.as<int32_t>(), cols + 1, 0);
  hipLaunchKernelGGL(csc_offsets_kernel, dim3(grid64(nnz)), dim3(256), 0, stream, nnz, cols,
                     keys.as<int32_t>(), t->off.as<int32_t>());
  hipLaunchKernelGGL(gather_rows_kernel, dim3(grid64(nnz)), dim3(256), 0, stream, nnz,
                     t->perm.as<int32_t>(), roe.as<int32_t>(), t->row.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipStreamSynchronize(stream));
  return GD_OK;
}

__global__ void permute_values_kernel(int64_t nnz, const int32_t* __restrict__ perm,
                                      const float* __restrict__ src, float* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

int transpose_values(const Transpose& t, int64_t nnz, const float* csr_vals, float* csc_vals,
                     hipStream_t stream) {
  if (nnz == 0) return GD_OK;
  hipLaunchKernelGGL(permute_values_kernel, dim3(grid64(nnz)), dim3(256), 0, stream, nnz,
                     t.perm.as<int32_t>(), csr_vals, csc_vals);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

}  // namespace gd
