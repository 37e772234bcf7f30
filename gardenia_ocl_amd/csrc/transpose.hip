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

void edge_rows(int m, const int32_t* d_off, int32_t* d_row_of_edge, hipStream_t stream) {
  hipLaunchKernelGGL(edge_rows_kernel, dim3(grid64(m / 4 + 1)), dim3(256), 0, stream, m, d_off,
                     d_row_of_edge);
}

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

// Chunks of kPermChunk positions, kPermU gathers in flight per thread.
// Workgroup b takes chunk xcd_chunk(), so the positions in flight on one XCD
// stay in one narrow window (of the destination for a gather, of the source
// for a scatter).  The grid is a multiple of 8 (G = 8 * ceil(chunks / 8));
// chunks past the end are skipped.
constexpr int kPermU = 8;
constexpr int kPermChunk = 256 * kPermU;

static int xcd_grid(int64_t n) {
  const int64_t chunks = (n + kPermChunk - 1) / kPermChunk;
  return (int)(kXcds * ((chunks + kXcds - 1) / kXcds));
}

// dst[i] = src[perm[i]]
template <class T>
__global__ __launch_bounds__(256) void permute_kernel(int64_t n, const int32_t* __restrict__ perm,
                                                      const T* __restrict__ src,
                                                      T* __restrict__ dst) {
  const int64_t c = xcd_chunk() * kPermChunk;
  if (c >= n) return;
  const int64_t last = n - 1;
  int p[kPermU];
  T v[kPermU];
#pragma unroll
  for (int q = 0; q < kPermU; q++) p[q] = perm[min(c + q * 256 + threadIdx.x, last)];
#pragma unroll
  for (int q = 0; q < kPermU; q++) v[q] = src[p[q]];
#pragma unroll
  for (int q = 0; q < kPermU; q++)
    if (c + q * 256 + threadIdx.x <= last) dst[c + q * 256 + threadIdx.x] = v[q];
}

// Chunk-local scatter: the chunk's source values are staged in LDS, then
// written in the order of their destinations (lidx[t]: the chunk-local source
// index of the chunk's t-th destination in increasing order, pos_sorted[t]:
// that destination), so consecutive lanes store to consecutive addresses
// wherever a chunk sends a run of entries to one bucket.
__global__ __launch_bounds__(256) void scatter_sorted_kernel(int64_t n,
                                                             const uint16_t* __restrict__ lidx,
                                                             const int32_t* __restrict__ pos_sorted,
                                                             const float* __restrict__ src,
                                                             float* __restrict__ dst) {
  __shared__ float s_v[kPermChunk];
  const int64_t c = xcd_chunk() * kPermChunk;
  if (c >= n) return;
#pragma unroll
  for (int q = 0; q < kPermU; q++) {
    const int64_t k = c + q * 256 + threadIdx.x;
    if (k < n) s_v[q * 256 + threadIdx.x] = src[k];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPermU; q++) {
    const int64_t t = c + q * 256 + threadIdx.x;
    if (t < n) dst[pos_sorted[t]] = s_v[lidx[t]];
  }
}

// keys2[k] = (chunk of k) << key_bits | keys[k]
__global__ void chunk_keys_kernel(int64_t n, const int32_t* __restrict__ keys, int key_bits,
                                  int32_t* __restrict__ keys2) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    keys2[k] = (int32_t)(((k / kPermChunk) << key_bits) | keys[k]);
}

// lidx[t] = order[t] - chunk start, pos_sorted[t] = pos[order[t]]
__global__ void chunk_order_kernel(int64_t n, const int32_t* __restrict__ order,
                                   const int32_t* __restrict__ pos, uint16_t* __restrict__ lidx,
                                   int32_t* __restrict__ pos_sorted) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t k = order[t];
    lidx[t] = (uint16_t)(k % kPermChunk);
    pos_sorted[t] = pos[k];
  }
}

// pos[order[t]] = t
__global__ void invert_kernel(int64_t n, const int32_t* __restrict__ order,
                              int32_t* __restrict__ pos) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    pos[order[t]] = (int32_t)t;
}

int transpose_values(const Transpose& t, int64_t nnz, const float* csr_vals, float* csc_vals,
                     hipStream_t stream) {
  if (nnz == 0) return GD_OK;
  return permute_values(nnz, t.perm.as<int32_t>(), csr_vals, csc_vals, stream);
}

int permute_values(int64_t nnz, const int32_t* perm, const float* src, float* dst,
                   hipStream_t stream) {
  if (nnz == 0) return GD_OK;
  hipLaunchKernelGGL(permute_kernel<float>, dim3(xcd_grid(nnz)), dim3(256), 0, stream, nnz, perm,
                     src, dst);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

int build_bucketed_permutation(int64_t n, const int32_t* d_keys, int key_bits,
                               const int32_t* d_perm, hipStream_t stream, BucketedPermutation* b) {
  b->n = n;
  // the per-chunk store order sorts (chunk, bucket) keys of 31 bits: past that
  // (about 2^29.5 entries) the permutation is one direct gather
  int chunk_bits = 1;
  while ((1ll << chunk_bits) < (n + kPermChunk - 1) / kPermChunk) chunk_bits++;
  if (chunk_bits + key_bits > 31) {
    b->direct = true;
    int rc = b->perm.alloc(4 * (size_t)std::max<int64_t>(n, 1));
    if (rc) return rc;
    if (n > 0)
      GD_HIP_TRY(hipMemcpyAsync(b->perm.ptr, d_perm, 4 * (size_t)n, hipMemcpyDeviceToDevice,
                                stream));
    GD_HIP_TRY(hipStreamSynchronize(stream));
    return GD_OK;
  }
  int rc = b->pos.alloc(4 * (size_t)std::max<int64_t>(n, 1));
  if (!rc) rc = b->perm.alloc(4 * (size_t)std::max<int64_t>(n, 1));
  if (!rc) rc = b->buf.alloc(4 * (size_t)std::max<int64_t>(n, 1));
  if (rc || n == 0) return rc;
  DeviceBuffer ids, order, keys_out, tmp;
  rc = ids.alloc(4 * (size_t)n);
  if (!rc) rc = order.alloc(4 * (size_t)n);
  if (!rc) rc = keys_out.alloc(4 * (size_t)n);
  if (rc) return rc;
  hipLaunchKernelGGL(iota64_kernel, dim3(grid64(n)), dim3(256), 0, stream, ids.as<int32_t>(), n);
  GD_HIP_TRY(hipGetLastError());
  // stable: the entries of one bucket keep their source order
  size_t bytes = 0;
  GD_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, d_keys, keys_out.as<int32_t>(),
                                       ids.as<int32_t>(), order.as<int32_t>(), (size_t)n, 0,
                                       key_bits, stream));
  rc = tmp.alloc(bytes);
  if (rc) return rc;
  GD_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, bytes, d_keys, keys_out.as<int32_t>(),
                                       ids.as<int32_t>(), order.as<int32_t>(), (size_t)n, 0,
                                       key_bits, stream));
  hipLaunchKernelGGL(invert_kernel, dim3(grid64(n)), dim3(256), 0, stream, n,
                     order.as<int32_t>(), b->pos.as<int32_t>());
  hipLaunchKernelGGL(permute_kernel<int32_t>, dim3(xcd_grid(n)), dim3(256), 0, stream, n, d_perm,
                     b->pos.as<int32_t>(), b->perm.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  // each chunk's entries in the order of their bucket-buffer positions: a
  // stable sort by (chunk, bucket) (the positions increase with the source
  // index inside a bucket)
  rc = b->lidx.alloc(2 * (size_t)n);
  if (!rc) rc = b->pos_sorted.alloc(4 * (size_t)n);
  if (rc) return rc;
  hipLaunchKernelGGL(chunk_keys_kernel, dim3(grid64(n)), dim3(256), 0, stream, n, d_keys,
                     key_bits, keys_out.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  DeviceBuffer keys_sorted;
  rc = keys_sorted.alloc(4 * (size_t)n);
  if (rc) return rc;
  size_t bytes2 = 0;
  GD_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes2, keys_out.as<int32_t>(),
                                       keys_sorted.as<int32_t>(), ids.as<int32_t>(),
                                       order.as<int32_t>(), (size_t)n, 0,
                                       chunk_bits + key_bits, stream));
  if (bytes2 > bytes) {
    rc = tmp.alloc(bytes2);  // frees the smaller one first
    if (rc) return rc;
  }
  GD_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, bytes2, keys_out.as<int32_t>(),
                                       keys_sorted.as<int32_t>(), ids.as<int32_t>(),
                                       order.as<int32_t>(), (size_t)n, 0,
                                       chunk_bits + key_bits, stream));
  hipLaunchKernelGGL(chunk_order_kernel, dim3(grid64(n)), dim3(256), 0, stream, n,
                     order.as<int32_t>(), b->pos.as<int32_t>(), b->lidx.as<uint16_t>(),
                     b->pos_sorted.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipStreamSynchronize(stream));  // the temporaries are freed on return
  b->pos.release();  // only the chunk-sorted form is used from here on
  return GD_OK;
}

int run_bucketed_permutation(const BucketedPermutation& b, const float* src, float* dst,
                             hipStream_t stream) {
  if (b.n == 0) return GD_OK;
  if (b.direct) return permute_values(b.n, b.perm.as<int32_t>(), src, dst, stream);
  hipLaunchKernelGGL(scatter_sorted_kernel, dim3(xcd_grid(b.n)), dim3(256), 0, stream, b.n,
                     b.lidx.as<uint16_t>(), b.pos_sorted.as<int32_t>(), src, b.buf.as<float>());
  hipLaunchKernelGGL(permute_kernel<float>, dim3(xcd_grid(b.n)), dim3(256), 0, stream, b.n,
                     b.perm.as<int32_t>(), b.buf.as<float>(), dst);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

}  // namespace gd
