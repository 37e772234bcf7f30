// push.hip -- push-style PageRank over the out-CSR (src/pr/scatter.cu:11-41,
// host loop :57-78): the reference's alternative to the pull path, kept as a
// cross-check of it.  Per iteration:
//   next[v] = base for every v; every out-edge (u, v) adds
//   kDamp * score[u] / outdeg[u] to next[v] (a hardware FP32 atomic add in
//   L2: the order of the additions is the hardware's, as with the reference's
//   atomicAdd, so the scores match the pull path's to FP32 summation-order
//   noise, not bit for bit); then |next - score| summed, score = next.
// The loop continues while the change is > tol, at least once (the
// reference's do-while).  Unlike the reference, the L1 change is summed in
// FP64 in a fixed order (per-workgroup partials over a fixed grid, added in
// index order), so it does not depend on scheduling.
//
// MI355X layout: the scatter is edge-parallel over a COO copy of the sources
// (built once per call), so a source with 10^5 out-edges is spread over the
// whole grid instead of one thread (the reference's thread per source,
// :16-27); the per-source value is computed once per iteration, not once per
// edge.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"

namespace gd {
namespace {

constexpr int kPushBlock = 256;
constexpr int kPushParts = 1024;  // workgroups of the update pass (fixed: a fixed L1 grouping)

// src[e] = u for every edge e of u's out-row (once per call).
__global__ void push_sources_kernel(int m, const int32_t* __restrict__ off,
                                    int32_t* __restrict__ src) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < m; u += gridDim.x * blockDim.x)
    for (int e = off[u]; e < off[u + 1]; e++) src[e] = u;
}

// value[u] = kDamp * score[u] / outdeg[u] ((kDamp * score) first, as
// scatter.cu:22), next[v] = base (initialize, :11-14).
__global__ void push_prepare_kernel(int m, const int32_t* __restrict__ off,
                                    const float* __restrict__ scores, float damp, float base,
                                    float* __restrict__ value, float* __restrict__ next) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m; v += gridDim.x * blockDim.x) {
    value[v] = damp * scores[v] / (float)(off[v + 1] - off[v]);
    next[v] = base;
  }
}

// scatter (:16-27), edge-parallel: next[col[e]] += value[src[e]], four edges
// per thread in flight.
__global__ __launch_bounds__(kPushBlock) void push_edges_kernel(int64_t nnz,
                                                                const int32_t* __restrict__ src,
                                                                const int32_t* __restrict__ col,
                                                                const float* __restrict__ value,
                                                                float* __restrict__ next) {
  const int64_t stride = (int64_t)gridDim.x * kPushBlock;
  int64_t e = (int64_t)blockIdx.x * kPushBlock + threadIdx.x;
  for (; e + 3 * stride < nnz; e += 4 * stride) {
    int d[4];
    float x[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      d[q] = col[e + q * stride];
      x[q] = value[src[e + q * stride]];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) unsafeAtomicAdd(next + d[q], x[q]);
  }
  for (; e < nnz; e += stride) unsafeAtomicAdd(next + col[e], value[src[e]]);
}

// reduce (:30-41): per-workgroup FP64 sum of |next - score| (fixed grid and
// order), score = next, then the next iteration's value and next = base.
__global__ __launch_bounds__(kPushBlock) void push_update_kernel(
    int m, const int32_t* __restrict__ off, float* __restrict__ scores, float* __restrict__ next,
    float* __restrict__ value, float damp, float base, double* __restrict__ part) {
  __shared__ double s_tmp[kPushBlock / 64];
  double err = 0.0;
  for (int v = blockIdx.x * kPushBlock + threadIdx.x; v < m; v += gridDim.x * kPushBlock) {
    const float nv = next[v];
    err += (double)fabsf(nv - scores[v]);
    scores[v] = nv;
    value[v] = damp * nv / (float)(off[v + 1] - off[v]);
    next[v] = base;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) err += __shfl_xor(err, o, 64);
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x / 64] = err;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kPushBlock / 64; w++) t += s_tmp[w];
    part[blockIdx.x] = t;
  }
}

// L1 of the iteration: the workgroup partials added in index order.
__global__ __launch_bounds__(kPushBlock) void push_total_kernel(int n, const double* __restrict__ part,
                                                                double* __restrict__ out) {
  __shared__ double s_part[kPushParts];
  for (int i = threadIdx.x; i < n; i += kPushBlock) s_part[i] = part[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < n; i++) t += s_part[i];
    *out = t;
  }
}

}  // namespace
}  // namespace gd

using namespace gd;

extern "C" int gd_pr_push_solve(int m, int64_t nnz, const int32_t* d_out_row_offsets,
                                const int32_t* d_out_column_indices, float* d_scores,
                                const gd_pr_options* opts, void* stream, gd_pr_stats* stats,
                                double* errors) {
  GD_CHECK_ARG(m >= 0 && nnz >= 0 && nnz < (1ll << 31), "m/nnz out of range");
  GD_CHECK_ARG(m == 0 || (d_out_row_offsets && d_scores), "NULL device array");
  GD_CHECK_ARG(nnz == 0 || d_out_column_indices, "NULL column indices");
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  GD_CHECK_ARG(o.max_iter >= 0, "max_iter must be >= 0");
  GD_CHECK_ARG(o.seed_vertex == -1, "the push solver has no personalised form (seed_vertex -1)");
  gd_pr_stats st;
  memset(&st, 0, sizeof(st));
  if (m == 0) {
    if (stats) *stats = st;
    return GD_OK;
  }
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = check_offsets(d_out_row_offsets, m, nnz, s);
  if (!rc && nnz > 0) rc = check_columns(d_out_column_indices, nnz, m, s);
  if (rc) return rc;
  DeviceBuffer src, value, next, part, total;
  rc = src.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (!rc) rc = value.alloc(4 * (size_t)m);
  if (!rc) rc = next.alloc(4 * (size_t)m);
  if (!rc) rc = part.alloc(sizeof(double) * kPushParts);
  if (!rc) rc = total.alloc(sizeof(double));
  if (rc) return rc;
  const int vgrid = std::max(1, std::min(16384, (m + kPushBlock - 1) / kPushBlock));
  const int ugrid = std::max(1, std::min(kPushParts, (m + kPushBlock - 1) / kPushBlock));
  const int egrid =
      (int)std::max<int64_t>(1, std::min<int64_t>(32768, (nnz + 4 * kPushBlock - 1) / (4 * kPushBlock)));
  const float damp = o.damping;
  const float base = (1.0f - damp) / (float)m;  // scatter.cu:57
  EventPair ev;
  GD_HIP_TRY(ev.create());
  if (nnz > 0)
    hipLaunchKernelGGL(push_sources_kernel, dim3(vgrid), dim3(kPushBlock), 0, s, m,
                       d_out_row_offsets, src.as<int32_t>());
  GD_HIP_TRY(hipEventRecord(ev.a, s));
  hipLaunchKernelGGL(push_prepare_kernel, dim3(vgrid), dim3(kPushBlock), 0, s, m,
                     d_out_row_offsets, d_scores, damp, base, value.as<float>(), next.as<float>());
  GD_HIP_TRY(hipGetLastError());
  int iter = 0;
  double diff = 0.0;
  std::vector<double> log;
  do {  // :66-78
    ++iter;
    if (nnz > 0)
      hipLaunchKernelGGL(push_edges_kernel, dim3(egrid), dim3(kPushBlock), 0, s, nnz,
                         src.as<int32_t>(), d_out_column_indices, value.as<float>(),
                         next.as<float>());
    hipLaunchKernelGGL(push_update_kernel, dim3(ugrid), dim3(kPushBlock), 0, s, m,
                       d_out_row_offsets, d_scores, next.as<float>(), value.as<float>(), damp, base,
                       part.as<double>());
    hipLaunchKernelGGL(push_total_kernel, dim3(1), dim3(kPushBlock), 0, s, ugrid,
                       part.as<double>(), total.as<double>());
    GD_HIP_TRY(hipGetLastError());
    GD_HIP_TRY(hipMemcpyAsync(&diff, total.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    GD_HIP_TRY(hipStreamSynchronize(s));
    log.push_back(diff);
  } while (diff > o.tol && iter < o.max_iter);
  GD_HIP_TRY(hipEventRecord(ev.b, s));
  GD_HIP_TRY(hipEventSynchronize(ev.b));
  GD_HIP_TRY(hipEventElapsedTime(&st.solve_ms, ev.a, ev.b));
  st.iterations = iter;
  st.final_error = diff;
  st.converged = !(diff > o.tol);
  if (errors) std::copy(log.begin(), log.end(), errors);
  if (stats) *stats = st;
  return GD_OK;
}
