// transpose.h -- plan-time CSR -> CSC (device), used by the adjoints.
#ifndef GD_TRANSPOSE_H_
#define GD_TRANSPOSE_H_

#include <hip/hip_runtime.h>

#include "gd_common.h"

namespace gd {

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

}  // namespace gd

#endif  // GD_TRANSPOSE_H_
