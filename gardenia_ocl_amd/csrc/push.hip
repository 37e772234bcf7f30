// push.hip -- push-style PageRank over the out-CSR (src/pr/scatter.cu:11-41,
// host loop :57-78): the reference's alternative to the pull path, kept as a
// cross-check of it.  Per iteration:
//   next[v] = base for every v; every out-edge (u, v) adds
//   kDamp * score[u] / outdeg[u] to next[v]; then |next - score| summed,
//   score = next.
// The loop continues while the change is > tol, at least once (the
// reference's do-while).  The additions into one vertex happen in an order
// the hardware picks (LDS atomic adds; the reference's atomicAdd likewise),
// so the scores match the pull path's to FP32 summation-order noise, not bit
// for bit.  Unlike the reference, the L1 change is summed in FP64 in a fixed
// order (per-workgroup partials over a fixed grid, added in index order), so
// it does not depend on scheduling.
//
// MI355X layout: propagation blocking (Beamer, Asanovic and Patterson,
// "Reducing PageRank Communication via Propagation Blocking", IPDPS 2017).
// The reference's atomicAdd per edge into next[] (67 MB at RMAT-24) misses
// every cache.  Here the destinations are cut into windows of kWin vertices;
// once per call, a stable radix sort gives every edge a slot in window order
// (out-CSR order inside a window).  Per iteration:
//   1. bin: every edge writes value[u] into its slot.  Workgroups walk the
//      out-CSR in XCD-contiguous chunks, value[u] is read in source order;
//      a chunk's values are staged in LDS and stored in slot order, so the
//      stores into one window are runs of consecutive slots;
//   2. accumulate: a workgroup per piece of a window (at most kPieceEdges
//      slots) adds the piece's values into the window in LDS (ds_add_f32)
//      and stores the window's partial sums;
//   3. update: next[v] = base + the partial sums of v's window, pieces in
//      order; the L1 change, score = next and the next value[] in the same
//      pass.
// No global atomics: the accumulation of a piece is one workgroup's.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "transpose.h"

namespace gd {
namespace {

constexpr int kPushBlock = 256;
constexpr int kPushParts = 4096;  // workgroups of the update pass (fixed: a fixed L1 grouping)
constexpr int kWinBits = 15;
constexpr int kWin = 1 << kWinBits;   // destination window: 32,768 vertices, 128 KB of LDS
constexpr int kPieceEdges = 1 << 18;  // slots per accumulating workgroup at most
constexpr int kAccBlock = 1024;
constexpr int kAccU = 4;              // slots in flight per thread of the accumulation
constexpr int kBinU = 8;              // edges per thread in flight in the binning pass
constexpr int kSortChunk = 4096;      // edges per workgroup of the binning pass

// keys[e] = destination window of edge e, ids[e] = e (once per call).
__global__ void window_keys_kernel(int64_t nnz, const int32_t* __restrict__ col,
                                   int32_t* __restrict__ keys, int32_t* __restrict__ ids) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    keys[e] = col[e] >> kWinBits;
    ids[e] = (int32_t)e;
  }
}

// order[i] = edge in slot i: slot[e] = i, low[i] = its destination inside the
// window (once per call).
__global__ void window_layout_kernel(int64_t nnz, const int32_t* __restrict__ order,
                                     const int32_t* __restrict__ col, int32_t* __restrict__ slot,
                                     uint16_t* __restrict__ low) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = order[i];
    slot[e] = (int32_t)i;
    low[i] = (uint16_t)(col[e] & (kWin - 1));
  }
}

// start[w] = first slot of window w, start[windows] = nnz, from the sorted
// keys (nnz >= 1; each entry written once).
__global__ void window_starts_kernel(int64_t nnz, const int32_t* __restrict__ keys, int windows,
                                     int32_t* __restrict__ start) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int k = keys[i], prev = i ? keys[i - 1] : -1;
    for (int w = prev + 1; w <= k; w++) start[w] = (int32_t)i;
    if (i == nnz - 1)
      for (int w = k + 1; w <= windows; w++) start[w] = (int32_t)nnz;
  }
}

// value[u] = kDamp * score[u] / outdeg[u] ((kDamp * score) first, as
// scatter.cu:22).
__global__ void push_prepare_kernel(int m, const int32_t* __restrict__ off,
                                    const float* __restrict__ scores, float damp,
                                    float* __restrict__ value) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m; v += gridDim.x * blockDim.x)
    value[v] = damp * scores[v] / (float)(off[v + 1] - off[v]);
}

// 1. bin: binval[slot[e]] = value[src[e]] over a chunk of kSortChunk edges.
// The chunk's values are staged in LDS in edge order, then stored in slot
// order (lidx[t]: chunk-local edge of the chunk's t-th slot, sslot[t]: that
// slot), so the stores into one window come out as runs of consecutive
// slots.
__global__ __launch_bounds__(kPushBlock) void push_bin_kernel(
    int64_t nnz, const int32_t* __restrict__ src, const uint16_t* __restrict__ lidx,
    const int32_t* __restrict__ sslot, const float* __restrict__ value,
    float* __restrict__ binval) {
  __shared__ float s_v[kSortChunk];
  const int64_t c = xcd_chunk() * kSortChunk;
  if (c >= nnz) return;
  const int n = (int)min<int64_t>(kSortChunk, nnz - c);
  for (int q0 = 0; q0 < kSortChunk; q0 += kPushBlock * kBinU) {
    int u[kBinU];
#pragma unroll
    for (int t = 0; t < kBinU; t++)
      u[t] = src[c + min(q0 + t * kPushBlock + (int)threadIdx.x, n - 1)];
#pragma unroll
    for (int t = 0; t < kBinU; t++) {
      const int k = q0 + t * kPushBlock + threadIdx.x;
      if (k < n) s_v[k] = value[u[t]];
    }
  }
  __syncthreads();
  for (int q0 = 0; q0 < kSortChunk; q0 += kPushBlock * kBinU) {
#pragma unroll
    for (int t = 0; t < kBinU; t++) {
      const int k = q0 + t * kPushBlock + threadIdx.x;
      if (k < n) binval[sslot[c + k]] = s_v[lidx[c + k]];
    }
  }
}

// keys2[e] = (e's sort chunk) << wbits | e's window (once per call)
__global__ void chunk_window_keys_kernel(int64_t nnz, const int32_t* __restrict__ col, int wbits,
                                         int64_t* __restrict__ keys2) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x)
    keys2[e] = ((e / kSortChunk) << wbits) | (col[e] >> kWinBits);
}

// lidx[t] = order[t] - its chunk's start, sslot[t] = slot[order[t]]
__global__ void chunk_slots_kernel(int64_t nnz, const int32_t* __restrict__ order,
                                   const int32_t* __restrict__ slot, uint16_t* __restrict__ lidx,
                                   int32_t* __restrict__ sslot) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nnz;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = order[t];
    lidx[t] = (uint16_t)(e % kSortChunk);
    sslot[t] = slot[e];
  }
}

// 2. accumulate: piece p = (window, first slot, end slot) -> part[p][0, kWin).
__global__ __launch_bounds__(kAccBlock) void push_window_kernel(const int4* __restrict__ pieces,
                                                                const uint16_t* __restrict__ low,
                                                                const float* __restrict__ binval,
                                                                float* __restrict__ part) {
  __shared__ float acc[kWin];
  for (int j = threadIdx.x; j < kWin; j += kAccBlock) acc[j] = 0.0f;
  __syncthreads();
  const int4 p = pieces[blockIdx.x];
  int i = p.y + threadIdx.x;
  for (; (int64_t)i + (kAccU - 1) * kAccBlock < p.z; i += kAccU * kAccBlock) {
    int d[kAccU];
    float x[kAccU];
#pragma unroll
    for (int t = 0; t < kAccU; t++) {
      d[t] = low[i + t * kAccBlock];
      x[t] = binval[i + t * kAccBlock];
    }
#pragma unroll
    for (int t = 0; t < kAccU; t++) atomicAdd(&acc[d[t]], x[t]);
  }
  for (; i < p.z; i += kAccBlock) atomicAdd(&acc[low[i]], binval[i]);
  __syncthreads();
  float* out = part + (int64_t)blockIdx.x * kWin;
  for (int j = threadIdx.x; j < kWin; j += kAccBlock) out[j] = acc[j];
}

// 3. update (reduce, :30-41): next = base + v's window partials (pieces in
// order), per-workgroup FP64 sum of |next - score| (fixed grid and order),
// score = next, then the next iteration's value.
__global__ __launch_bounds__(kPushBlock) void push_update_kernel(
    int m, const int32_t* __restrict__ off, float* __restrict__ scores, float* __restrict__ value,
    const int32_t* __restrict__ pfirst, const float* __restrict__ part, float damp, float base,
    double* __restrict__ partial) {
  __shared__ double s_tmp[kPushBlock / 64];
  double err = 0.0;
  for (int v = blockIdx.x * kPushBlock + threadIdx.x; v < m; v += gridDim.x * kPushBlock) {
    const int w = v >> kWinBits, j = v & (kWin - 1);
    float sum = 0.0f;
    for (int q = pfirst[w]; q < pfirst[w + 1]; q++) sum += part[(int64_t)q * kWin + j];
    const float nv = base + sum;
    err += (double)fabsf(nv - scores[v]);
    scores[v] = nv;
    value[v] = damp * nv / (float)(off[v + 1] - off[v]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) err += __shfl_xor(err, o, 64);
  if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x / 64] = err;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kPushBlock / 64; w++) t += s_tmp[w];
    partial[blockIdx.x] = t;
  }
}

// L1 of the iteration from the workgroup partials, in a fixed order: thread t
// adds partials [t * per, + per) in index order, thread 0 adds those sums in
// thread order.
__global__ __launch_bounds__(kPushBlock) void push_total_kernel(int n, const double* __restrict__ part,
                                                                double* __restrict__ out) {
  __shared__ double s_sum[kPushBlock];
  const int per = (n + kPushBlock - 1) / kPushBlock;
  double t = 0.0;
  for (int i = threadIdx.x * per; i < min(n, (threadIdx.x + 1) * per); i++) t += part[i];
  s_sum[threadIdx.x] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int i = 0; i < kPushBlock; i++) a += s_sum[i];
    *out = a;
  }
}

// The window layout of an out-CSR (once per call): each chunk's slots in
// order (lidx + sslot), low, the pieces and each window's first piece.
// Synchronises `s`.
struct WindowLayout {
  DeviceBuffer lidx, sslot, low, pieces, pfirst;
  int n_pieces = 0;
};

int build_windows(int m, int64_t nnz, const int32_t* col, hipStream_t s, WindowLayout* L) {
  const int windows = (int)(((int64_t)m + kWin - 1) >> kWinBits);
  std::vector<int32_t> start(windows + 1, 0);
  if (nnz > 0) {
    DeviceBuffer keys, keys_out, ids, order, wstart, slot, tmp;
    int rc = keys.alloc(4 * (size_t)nnz);
    if (!rc) rc = keys_out.alloc(4 * (size_t)nnz);
    if (!rc) rc = ids.alloc(4 * (size_t)nnz);
    if (!rc) rc = order.alloc(4 * (size_t)nnz);
    if (!rc) rc = wstart.alloc(4 * ((size_t)windows + 1));
    if (!rc) rc = slot.alloc(4 * (size_t)nnz);
    if (!rc) rc = L->low.alloc(2 * (size_t)nnz);
    if (rc) return rc;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(65536, (nnz + 255) / 256));
    hipLaunchKernelGGL(window_keys_kernel, dim3(g), dim3(256), 0, s, nnz, col,
                       keys.as<int32_t>(), ids.as<int32_t>());
    GD_HIP_TRY(hipGetLastError());
    int bits = 1;
    while ((1 << bits) < windows) bits++;
    // stable: a window's edges keep their out-CSR order
    size_t bytes = 0;
    GD_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, keys.as<int32_t>(),
                                         keys_out.as<int32_t>(), ids.as<int32_t>(),
                                         order.as<int32_t>(), (size_t)nnz, 0, bits, s));
    rc = tmp.alloc(bytes);
    if (rc) return rc;
    GD_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, bytes, keys.as<int32_t>(),
                                         keys_out.as<int32_t>(), ids.as<int32_t>(),
                                         order.as<int32_t>(), (size_t)nnz, 0, bits, s));
    hipLaunchKernelGGL(window_layout_kernel, dim3(g), dim3(256), 0, s, nnz, order.as<int32_t>(),
                       col, slot.as<int32_t>(), L->low.as<uint16_t>());
    hipLaunchKernelGGL(window_starts_kernel, dim3(g), dim3(256), 0, s, nnz,
                       keys_out.as<int32_t>(), windows, wstart.as<int32_t>());
    GD_HIP_TRY(hipGetLastError());
    GD_HIP_TRY(hipMemcpyAsync(start.data(), wstart.ptr, 4 * ((size_t)windows + 1),
                              hipMemcpyDeviceToHost, s));
    int cbits = 1;
    while ((1ll << cbits) < (nnz + kSortChunk - 1) / kSortChunk) cbits++;
    // each chunk's edges in slot order: a stable sort by (chunk, window)
    // (slots increase with the edge inside a window)
    keys.release();
    keys_out.release();
    DeviceBuffer keys2, keys2_out;
    rc = keys2.alloc(8 * (size_t)nnz);
    if (!rc) rc = keys2_out.alloc(8 * (size_t)nnz);
    if (!rc) rc = L->lidx.alloc(2 * (size_t)nnz);
    if (!rc) rc = L->sslot.alloc(4 * (size_t)nnz);
    if (rc) return rc;
    hipLaunchKernelGGL(chunk_window_keys_kernel, dim3(g), dim3(256), 0, s, nnz, col, bits,
                       keys2.as<int64_t>());
    GD_HIP_TRY(hipGetLastError());
    size_t bytes2 = 0;
    GD_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes2, keys2.as<int64_t>(),
                                         keys2_out.as<int64_t>(), ids.as<int32_t>(),
                                         order.as<int32_t>(), (size_t)nnz, 0, cbits + bits, s));
    if (bytes2 > bytes) {
      rc = tmp.alloc(bytes2);
      if (rc) return rc;
    }
    GD_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, bytes2, keys2.as<int64_t>(),
                                         keys2_out.as<int64_t>(), ids.as<int32_t>(),
                                         order.as<int32_t>(), (size_t)nnz, 0, cbits + bits, s));
    hipLaunchKernelGGL(chunk_slots_kernel, dim3(g), dim3(256), 0, s, nnz, order.as<int32_t>(),
                       slot.as<int32_t>(), L->lidx.as<uint16_t>(), L->sslot.as<int32_t>());
    GD_HIP_TRY(hipGetLastError());
    GD_HIP_TRY(hipStreamSynchronize(s));
  }
  // pieces: each window's slots in ceil(n / kPieceEdges) near-equal parts
  std::vector<int4> pieces;
  std::vector<int32_t> pfirst(windows + 1, 0);
  for (int w = 0; w < windows; w++) {
    pfirst[w] = (int32_t)pieces.size();
    const int64_t a = start[w], n = (int64_t)start[w + 1] - a;
    const int64_t k = (n + kPieceEdges - 1) / kPieceEdges;
    for (int64_t q = 0; q < k; q++)
      pieces.push_back(make_int4(w, (int)(a + n * q / k), (int)(a + n * (q + 1) / k), 0));
  }
  pfirst[windows] = (int32_t)pieces.size();
  L->n_pieces = (int)pieces.size();
  int rc = L->pfirst.alloc(4 * ((size_t)windows + 1));
  if (!rc) rc = L->pieces.alloc(sizeof(int4) * std::max<size_t>(1, pieces.size()));
  if (rc) return rc;
  GD_HIP_TRY(hipMemcpyAsync(L->pfirst.ptr, pfirst.data(), 4 * ((size_t)windows + 1),
                            hipMemcpyHostToDevice, s));
  if (!pieces.empty())
    GD_HIP_TRY(hipMemcpyAsync(L->pieces.ptr, pieces.data(), sizeof(int4) * pieces.size(),
                              hipMemcpyHostToDevice, s));
  GD_HIP_TRY(hipStreamSynchronize(s));
  return GD_OK;
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
  WindowLayout L;
  rc = build_windows(m, nnz, d_out_column_indices, s, &L);
  if (rc) return rc;
  DeviceBuffer src, value, binval, part, partial, total;
  rc = src.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (!rc) rc = binval.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (!rc) rc = value.alloc(4 * (size_t)m);
  if (!rc) rc = part.alloc(4 * (size_t)kWin * std::max(1, L.n_pieces));
  if (!rc) rc = partial.alloc(sizeof(double) * kPushParts);
  if (!rc) rc = total.alloc(sizeof(double));
  if (rc) return rc;
  const int vgrid = std::max(1, std::min(16384, (m + kPushBlock - 1) / kPushBlock));
  const int ugrid = std::max(1, std::min(kPushParts, (m + kPushBlock - 1) / kPushBlock));
  const int64_t schunks = (nnz + kSortChunk - 1) / kSortChunk;
  const int sgrid = (int)(kXcds * ((schunks + kXcds - 1) / kXcds));  // a multiple of 8
  const float damp = o.damping;
  const float base = (1.0f - damp) / (float)m;  // scatter.cu:57
  EventPair ev;
  GD_HIP_TRY(ev.create());
  if (nnz > 0) edge_rows(m, d_out_row_offsets, src.as<int32_t>(), s);  // src[e] = u
  GD_HIP_TRY(hipEventRecord(ev.a, s));
  hipLaunchKernelGGL(push_prepare_kernel, dim3(vgrid), dim3(kPushBlock), 0, s, m,
                     d_out_row_offsets, d_scores, damp, value.as<float>());
  GD_HIP_TRY(hipGetLastError());
  int iter = 0;
  double diff = 0.0;
  std::vector<double> log;
  do {  // :66-78
    ++iter;
    if (nnz > 0) {
      hipLaunchKernelGGL(push_bin_kernel, dim3(sgrid), dim3(kPushBlock), 0, s, nnz,
                         src.as<int32_t>(), L.lidx.as<uint16_t>(), L.sslot.as<int32_t>(),
                         value.as<float>(), binval.as<float>());
      hipLaunchKernelGGL(push_window_kernel, dim3(L.n_pieces), dim3(kAccBlock), 0, s,
                         L.pieces.as<int4>(), L.low.as<uint16_t>(), binval.as<float>(),
                         part.as<float>());
    }
    hipLaunchKernelGGL(push_update_kernel, dim3(ugrid), dim3(kPushBlock), 0, s, m,
                       d_out_row_offsets, d_scores, value.as<float>(), L.pfirst.as<int32_t>(),
                       part.as<float>(), damp, base, partial.as<double>());
    hipLaunchKernelGGL(push_total_kernel, dim3(1), dim3(kPushBlock), 0, s, ugrid,
                       partial.as<double>(), total.as<double>());
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
