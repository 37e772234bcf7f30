// spmv_internal.h -- hooks of the SpMV plan for the other plans of this
// library (not part of the C ABI).
#ifndef GD_SPMV_INTERNAL_H_
#define GD_SPMV_INTERNAL_H_

#include <stdint.h>

#include "gardenia_hip.h"

namespace gd {

// The slice-major value copy of a plan in GD_SPMV_SLICED mode, nullptr in
// the tile mode.  A caller may write it directly instead of calling
// gd_spmv_plan_refresh_values: position i holds the value of the CSR entry
// that spmv_slice_order sends to i.
float* spmv_sliced_values(gd_spmv_plan* p);

// d_out[slice-major position of CSR entry k] = d_in[k] for every entry of a
// sliced plan (any 4-byte payload; async on the plan's stream).
int spmv_slice_order(gd_spmv_plan* p, const int32_t* d_in, int32_t* d_out);

// For a sliced plan: d_Xs[i] = d_x[column of i] for every slice-major
// position i, each slice's gathers on one XCD (async on the plan's stream).
int spmv_sliced_gather(gd_spmv_plan* p, const float* d_x, float* d_Xs);

}  // namespace gd

#endif  // GD_SPMV_INTERNAL_H_
