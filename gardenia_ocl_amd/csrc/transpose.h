// transpose.h -- plan-time CSR -> CSC (device), used by the adjoints.
#ifndef GD_TRANSPOSE_H_
#define GD_TRANSPOSE_H_

#include <hip/hip_runtime.h>

#include "gd_common.h"

namespace gd {

// Workgroups are dispatched to the 8 XCDs round-robin.  With a grid that is a
// multiple of 8, workgroup b taking chunk xcd_chunk() makes XCD x walk the
// x-th eighth of the chunks in order, so the data in flight on one XCD stays
// in one narrow window.
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_chunk() {
  const int b = blockIdx.x, per_xcd = gridDim.x / kXcds;
  return (int64_t)(b % kXcds) * per_xcd + b / kXcds;
}

// row_of_edge[e] = the row r with off[r] <= e < off[r + 1], a wavefront per
// row (async on stream).
void edge_rows(int m, const int32_t* d_off, int32_t* d_row_of_edge, hipStream_t stream);

struct Transpose {
  DeviceBuffer off;   // [cols+1] CSC column offsets
  DeviceBuffer row;   // [nnz] row index of each CSC entry (increasing within a column)
  DeviceBuffer perm;  // [nnz] CSR edge index of each CSC entry
};

// Synchronises `stream`.  d_col must already be validated (in [0, cols)).
int build_transpose(int rows, int cols, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                    hipStream_t stream, Transpose* t);

// csc_vals[k] = csr_vals[perm[k]]  (async on stream)
int transpose_values(const Transpose& t, int64_t nnz, const float* csr_vals, float* csc_vals,
                     hipStream_t stream);

// dst[i] = src[perm[i]] for i < nnz  (async on stream)
int permute_values(int64_t nnz, const int32_t* perm, const float* src, float* dst,
                   hipStream_t stream);

// dst[i] = src[perm[i]] in two passes through a bucket buffer: pass 1
// scatters src in its own order into buckets (source entry k to bucket
// keys[k], in source order within a bucket), pass 2 gathers dst from the
// bucket buffer.  Where the destination order reads the buckets one after
// another, the gathers of pass 2 stay within about one bucket per XCD (hits
// in its L2), while a direct random gather over a large src pulls a cache
// line from memory per element.
struct BucketedPermutation {
  int64_t n = 0;
  bool direct = false;      // too large for the chunk sort: perm holds src positions, one gather
  DeviceBuffer pos;         // [n] bucket-buffer position of source entry k (build time only)
  DeviceBuffer lidx;        // [n] uint16: per 2048-entry chunk, its entries in position order
  DeviceBuffer pos_sorted;  // [n] the bucket-buffer position of each of them
  DeviceBuffer perm;        // [n] bucket-buffer position read for destination entry i
  DeviceBuffer buf;         // [n] the bucket buffer (values)
};

// keys[k] < 2^key_bits.  Synchronises `stream`.  When (chunks x buckets) do
// not fit a 31-bit sort key (about 2^29.5 entries), the permutation is kept
// as one direct gather instead.
int build_bucketed_permutation(int64_t n, const int32_t* d_keys, int key_bits,
                               const int32_t* d_perm, hipStream_t stream, BucketedPermutation* b);
// dst[i] = src[perm[i]]  (async on stream)
int run_bucketed_permutation(const BucketedPermutation& b, const float* src, float* dst,
                             hipStream_t stream);

}  // namespace gd

#endif  // GD_TRANSPOSE_H_
