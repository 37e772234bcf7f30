// relabel.h -- plan-time vertex relabelling by decreasing out-degree octave.
#ifndef GD_RELABEL_H_
#define GD_RELABEL_H_

#include <hip/hip_runtime.h>

#include "gd_common.h"

namespace gd {

struct Relabel {
  DeviceBuffer perm;    // perm[new] = old
  DeviceBuffer inv;     // inv[old] = new
  DeviceBuffer off;     // relabelled in-CSR row offsets [m+1]
  DeviceBuffer col;     // relabelled in-CSR column indices [nnz] (row order kept)
  DeviceBuffer outdeg;  // out-degree in the new order [m]
};

// Builds the relabelled in-CSR on the device (synchronises `stream`).
int build_relabel(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                  const int32_t* d_outdeg, hipStream_t stream, Relabel* r);

// internal[i] = user[perm[i]]  /  user[perm[i]] = internal[i]  (async on stream)
int relabel_to_internal(const Relabel& r, int m, const float* user, float* internal,
                        hipStream_t stream);
int relabel_to_user(const Relabel& r, int m, const float* internal, float* user,
                    hipStream_t stream);

}  // namespace gd

#endif  // GD_RELABEL_H_
