// verify.hip -- device PageRank verifier: the reference PRVerifier's push
// residual (src/pr/verifier.cc:50-68) on the GPU, bit-exact per vertex.
//
// The reference pushes one iteration over the out-CSR: for src = 0..m-1 in
// order, contrib = scores[src] / degree[src] is added (FP32) to sums[dst] for
// each out-edge, then error += |base + kDamp*sums[v] - scores[v]| (FP64).
// Each destination therefore accumulates its sources in increasing source
// order.  Here the out-CSR is transposed with a stable sort (each
// destination's sources in exactly that order) and pulled by the iteration
// engine with every row of more than 4096 sources summed in order with
// sequential FP32 rounding by its own workgroup (exact_hub_kernel), every
// other row by its owner thread in order (kSeqAll): every per-vertex term is
// the reference's bit for bit; only the FP64 sum of the terms runs in a
// different (fixed) order.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "iter_engine.h"
#include "pull_engine.h"
#include "transpose.h"

namespace gd {
namespace {

struct VerifyGather {
  const float* __restrict__ contrib;
  __device__ __forceinline__ float value(int /*e*/, int u) const { return contrib[u]; }
};

struct VerifyEpilogue {
  const float* __restrict__ scores;
  float base;
  float damp;
  struct Pre {
    float old;
  };
  __device__ __forceinline__ Pre pre(int v) const { return Pre{scores[v]}; }
  __device__ __forceinline__ float init(const Pre&) const { return 0.0f; }
  // new_score = base_score + kDamp * sums[v]; error += fabs(new_score - scores[v])
  __device__ __forceinline__ double finish(int /*v*/, float sum, const Pre& p) const {
    const float nv = base + damp * sum;
    return (double)fabsf(nv - p.old);
  }
};

struct VerifyPolicy {
  static constexpr bool kExactHubs = true;
  static constexpr bool kSeqAll = true;  // every row in source order, bit for bit
  struct Args {
    const float* contrib;
    const float* scores;
    float base;
    float damp;
  };
  template <int NT>
  __device__ static VerifyGather gather(const Args& a, int) {
    return VerifyGather{a.contrib};
  }
  template <int NT>
  __device__ static VerifyEpilogue epilogue(const Args& a, int) {
    return VerifyEpilogue{a.scores, a.base, a.damp};
  }
};

// outgoing_contrib = scores_to_test[src] / degree[src]  (verifier.cc:53)
__global__ void verify_contrib_kernel(int m, const float* __restrict__ scores,
                                      const int32_t* __restrict__ degree,
                                      float* __restrict__ contrib) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < m; u += gridDim.x * blockDim.x)
    contrib[u] = scores[u] / (float)degree[u];
}

}  // namespace
}  // namespace gd

using namespace gd;

extern "C" int gd_pr_verify(int m, int64_t nnz, const int32_t* d_out_row_offsets,
                            const int32_t* d_out_column_indices, const int32_t* d_degree,
                            const float* d_scores, float damping, void* stream,
                            double* residual) {
  GD_CHECK_ARG(residual != nullptr, "NULL residual output");
  GD_CHECK_ARG(m >= 0 && nnz >= 0 && nnz < (1ll << 31), "m/nnz out of range");
  GD_CHECK_ARG(m == 0 || (d_out_row_offsets && d_degree && d_scores), "NULL device array");
  GD_CHECK_ARG(nnz == 0 || d_out_column_indices, "NULL column indices");
  *residual = 0.0;
  if (m == 0) return GD_OK;
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = check_columns(d_out_column_indices, nnz, m, s);
  if (!rc) rc = check_offsets(d_out_row_offsets, m, nnz, s);  // before the transpose reads them
  if (rc) return rc;
  Transpose tr;  // destination-major, sources in increasing order
  rc = build_transpose(m, m, nnz, d_out_row_offsets, d_out_column_indices, s, &tr);
  if (rc) return rc;
  DeviceBuffer contrib;
  rc = contrib.alloc(4 * (size_t)m);
  if (rc) return rc;
  hipLaunchKernelGGL(verify_contrib_kernel, dim3(std::min(4096, (m + 255) / 256)), dim3(256), 0,
                     s, m, d_scores, d_degree, contrib.as<float>());
  GD_HIP_TRY(hipGetLastError());
  IterEngine eng;
  eng.ipt = 16;
  eng.nt = 0;
  eng.persistent = false;
  eng.exact_len = kBlock * 16;  // longer rows: in-order exact sums by their own workgroup
  rc = eng.build(m, nnz, tr.off.as<int32_t>(), tr.row.as<int32_t>(), s);
  if (!rc) rc = eng.setup_grid<VerifyPolicy>();
  if (rc) return rc;
  const VerifyPolicy::Args a{contrib.as<float>(), d_scores, (1.0f - damping) / (float)m, damping};
  gd_pr_stats st;
  memset(&st, 0, sizeof(st));
  double err = 0.0;
  rc = eng.run<VerifyPolicy>(a, 1, 0.0, 1, false, &st, &err);
  if (rc) return rc;
  GD_HIP_TRY(hipStreamSynchronize(s));  // buffers are freed on return
  *residual = err;
  return GD_OK;
}
