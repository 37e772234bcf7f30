// noalloc.hip -- pull PageRank that allocates nothing (gd_pr_run_noalloc).
//
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  The caller
// passes the in-CSR, the out-degree and the scores (all in HBM); everything
// else the iteration needs -- the two contrib vectors, the per-workgroup L1
// partials, the per-iteration L1 log, the iteration state and the validation
// flags -- lives in one file-scope device array, `scratch`, of fixed size,
// one copy per GPU for the life of the process.  Host code never uses the
// array's name as an address: hipGetSymbolAddress gives each device's copy
// once (first call on that device) and every kernel takes that pointer.
//
// No schedule is built (that needs the offsets on the host): workgroup b owns
// rows [256 b, 256 b + 256).  A row of <= 64 in-edges is added by its owner
// thread in CSR order straight from global memory; longer rows are added one
// after the other by the workgroup's hub chain (three waves gather 1,536-edge
// chunks into LDS, one lane adds them in edge order).  Every row is therefore
// summed in the reference's order with its FP32 rounding
// (src/pr/omp_base.cc:29-41), and the epilogue is the plan path's
// (PrEpilogue).  The last workgroup of an iteration adds the L1 partials in a
// fixed order and advances the device iteration state (fused_finish), as the
// small-graph launches of the plan path do.
//
// Calls that overlap on one GPU share the array and race; the array keeps the
// last call's contents, so every call re-initialises what it reads.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "pr_policy.h"
#include "pull_engine.h"
#include "pull_iter.h"

namespace gd {
namespace {

constexpr int kNaMaxV = GD_NOALLOC_MAX_VERTICES;
constexpr int kNaMaxIter = GD_NOALLOC_MAX_ITER;
constexpr int kNaMaxBlocks = kNaMaxV / kBlock;
constexpr int kNaMaxDevices = 64;

// Layout of the scratch array (byte offsets, each 256-byte aligned).
constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t kOffC0 = 0;
constexpr size_t kOffC1 = kOffC0 + align256(sizeof(float) * kNaMaxV);
constexpr size_t kOffPart = kOffC1 + align256(sizeof(float) * kNaMaxV);
constexpr size_t kOffErr = kOffPart + align256(sizeof(double) * kNaMaxBlocks);
constexpr size_t kOffState = kOffErr + align256(sizeof(double) * kNaMaxIter);
constexpr size_t kOffFlags = kOffState + align256(sizeof(IterState));
constexpr size_t kScratchBytes = kOffFlags + 256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

struct NaPtrs {
  float* c0;
  float* c1;
  double* part;
  double* errors;
  IterState* st;
  unsigned* flags;  // [0]: first bad offset row, [1]: first bad column edge
};

NaPtrs carve(unsigned char* base) {
  return NaPtrs{reinterpret_cast<float*>(base + kOffC0), reinterpret_cast<float*>(base + kOffC1),
                reinterpret_cast<double*>(base + kOffPart),
                reinterpret_cast<double*>(base + kOffErr),
                reinterpret_cast<IterState*>(base + kOffState),
                reinterpret_cast<unsigned*>(base + kOffFlags)};
}

// Each device's address of `scratch`, looked up once with that device current.
std::mutex g_mu;
unsigned char* g_base[kNaMaxDevices] = {};

int scratch_base(unsigned char** out) {
  int dev = 0;
  GD_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDevices) {
    set_error("gd_pr_run_noalloc: device ordinal out of range");
    return GD_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_base[dev]) {
    void* p = nullptr;
    GD_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_base[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_base[dev];
  return GD_OK;
}

constexpr unsigned kNone = 0xffffffffu;

// Cleared by a kernel rather than hipMemsetAsync, so a caller may capture the
// whole call into a graph.
__global__ void na_clear_flags_kernel(unsigned* flags) {
  if (threadIdx.x < 2) flags[threadIdx.x] = kNone;
}

// CSR validation: offsets a prefix sum of nnz edges, every column in [0, m).
// Records the first offending row / edge so the host can name it.
__global__ void na_validate_kernel(const int32_t* __restrict__ off,
                                   const int32_t* __restrict__ col, int m, int64_t nnz,
                                   unsigned* flags) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = t0; r <= m; r += stride) {
    const int32_t o = off[r];
    const bool bad = (r == 0 && o != 0) || (r == m && (int64_t)o != nnz) ||
                     (r < m && off[r + 1] < o) || o < 0 || (int64_t)o > nnz;
    if (bad) atomicMin(&flags[0], (unsigned)r);
  }
  for (int64_t e = t0; e < nnz; e += stride) {
    const int32_t u = col[e];
    if (u < 0 || u >= m) atomicMin(&flags[1], (unsigned)e);
  }
}

// contrib of the first iteration (src/pr/omp_base.cc:25-27) and the state.
__global__ __launch_bounds__(kBlock) void na_init_kernel(int m, const float* __restrict__ scores,
                                                         const int32_t* __restrict__ outdeg,
                                                         float* __restrict__ c0, IterState* st,
                                                         int max_iter, double tol) {
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < m; v += gridDim.x * kBlock)
    c0[v] = scores[v] / (float)outdeg[v];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->ticket = 0;
    st->tol = tol;
  }
}

struct NaArgs {
  const int32_t* __restrict__ off;
  const int32_t* __restrict__ col;
  int e_max;
  int m;
  PrPolicy::Args pa;
  IterArgs it;  // err_part, st, fuse_parts, errors are the fields used
};

// One iteration: workgroup b owns rows [kBlock b, kBlock b + kBlock).
__global__ __launch_bounds__(kBlock) void na_pull_kernel(const NaArgs a) {
  if (a.it.st->done) return;
  __shared__ __align__(16) float buf[2 * kHubChunk + 4];
  __shared__ double s_err[kWaves];
  __shared__ uint64_t s_long[kWaves];
  __shared__ int s_last;
  const int odd = a.it.st->iter & 1;
  const auto g = PrPolicy::gather<0>(a.pa, odd);
  const auto e = PrPolicy::epilogue<0>(a.pa, odd);
  const int tid = threadIdx.x;
  const int r = blockIdx.x * kBlock + tid;
  int b = 0, en = 0;
  if (r < a.m) {
    b = a.off[r];
    en = a.off[r + 1];
  }
  const bool lng = en - b > kSeqRow;
  double err = 0.0;
  if (r < a.m && !lng) {
    // owner thread, CSR order, four gathers in flight
    const auto p = e.pre(r);
    float s = e.init(p);
    int k = b;
    for (; k + 4 <= en; k += 4) {
      const int u0 = a.col[k], u1 = a.col[k + 1], u2 = a.col[k + 2], u3 = a.col[k + 3];
      const float v0 = g.value(k, u0), v1 = g.value(k + 1, u1), v2 = g.value(k + 2, u2),
                  v3 = g.value(k + 3, u3);
      s += v0;
      s += v1;
      s += v2;
      s += v3;
    }
    for (; k < en; k++) s += g.value(k, a.col[k]);
    err = e.finish(r, s, p);
  }
  // long rows of this workgroup, in row order, each by the hub chain
  const uint64_t mask = __ballot(lng);
  if ((tid & (kWave - 1)) == 0) s_long[tid / kWave] = mask;
  __syncthreads();
  double err_long = 0.0;
  for (int w = 0; w < kWaves; w++) {
    uint64_t mk = s_long[w];
    while (mk) {
      const int row = blockIdx.x * kBlock + w * kWave + __builtin_ctzll(mk);
      mk &= mk - 1;
      const int rb = a.off[row], re = a.off[row + 1];
      const auto p = e.pre(row);
      const float s = hub_chain_sum<8, false>(a.col, nullptr, a.e_max, rb, re - rb, g, e.init(p),
                                              buf, buf + kHubChunk, buf + 2 * kHubChunk);
      if (tid == 0) err_long += e.finish(row, s, p);
    }
  }
  const double tot = block_sum<kWaves>(err, s_err) + err_long;
  if (tid == 0) a.it.err_part[blockIdx.x] = tot;
  fused_finish(a.it, s_err, &s_last);
}

}  // namespace
}  // namespace gd

using namespace gd;

extern "C" int gd_pr_run_noalloc(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                                 const int32_t* d_outdeg, float* d_scores,
                                 const gd_pr_options* opts, void* stream, gd_pr_stats* stats,
                                 double* errors) {
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  GD_CHECK_ARG(m >= 0 && nnz >= 0 && nnz < (1ll << 31), "m / nnz out of range");
  GD_CHECK_ARG(m <= GD_NOALLOC_MAX_VERTICES,
               "gd_pr_run_noalloc: m exceeds GD_NOALLOC_MAX_VERTICES (2^21), the size of its "
               "static scratch; use a gd_pr_plan");
  GD_CHECK_ARG(o.max_iter >= 0 && o.max_iter <= GD_NOALLOC_MAX_ITER,
               "gd_pr_run_noalloc: max_iter must be in [0, GD_NOALLOC_MAX_ITER]");
  GD_CHECK_ARG(o.seed_vertex >= -1 && o.seed_vertex < std::max(m, 1), "seed_vertex out of range");
  GD_CHECK_ARG(m == 0 || (d_off && d_outdeg && d_scores), "NULL device array");
  GD_CHECK_ARG(nnz == 0 || d_col, "NULL column indices");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  gd_pr_stats s;
  memset(&s, 0, sizeof(s));
  if (m == 0) {  // as gd_pr_plan_run: every iteration of the reference's loop has error 0
    if (o.max_iter > 0) {
      s.converged = 0.0 < o.tol;
      s.iterations = s.converged ? 1 : o.max_iter;
      if (errors)
        for (int i = 0; i < s.iterations; i++) errors[i] = 0.0;
    }
    if (stats) *stats = s;
    return GD_OK;
  }
  unsigned char* base = nullptr;
  int rc = scratch_base(&base);
  if (rc) return rc;
  const NaPtrs sp = carve(base);
  const hipStream_t st = (hipStream_t)stream;
  const auto t0 = std::chrono::steady_clock::now();

  // validate the CSR before any kernel indexes through it
  hipLaunchKernelGGL(na_clear_flags_kernel, dim3(1), dim3(64), 0, st, sp.flags);
  const int64_t work = std::max<int64_t>((int64_t)m + 1, nnz);
  const int vblocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
  hipLaunchKernelGGL(na_validate_kernel, dim3(vblocks), dim3(256), 0, st, d_off,
                     nnz ? d_col : d_off, m, nnz, sp.flags);
  GD_HIP_TRY(hipGetLastError());
  unsigned flags[2] = {kNone, kNone};
  GD_HIP_TRY(hipMemcpyAsync(flags, sp.flags, sizeof(flags), hipMemcpyDeviceToHost, st));
  GD_HIP_TRY(hipStreamSynchronize(st));
  if (flags[0] != kNone) {
    set_error("row offsets are not a prefix sum of nnz edges (first bad entry " +
              std::to_string(flags[0]) + ")");
    return GD_ERR_INVALID;
  }
  if (flags[1] != kNone) {
    set_error("column index out of range at edge " + std::to_string(flags[1]));
    return GD_ERR_INVALID;
  }

  const float damp = o.damping;
  const float bscore = (o.seed_vertex < 0) ? (1.0f - damp) / (float)m : (1.0f - damp);
  const int nblocks = (m + kBlock - 1) / kBlock;
  NaArgs a;
  memset(&a, 0, sizeof(a));
  a.off = d_off;
  a.col = nnz ? d_col : d_off;  // never read past a row's edges; any valid pointer
  a.e_max = (int)std::max<int64_t>(nnz - 1, 0);
  a.m = m;
  a.pa = PrPolicy::Args{d_scores, sp.c0, sp.c1, d_outdeg, bscore, damp, o.seed_vertex};
  a.it.err_part = sp.part;
  a.it.st = sp.st;
  a.it.fuse_parts = nblocks;
  a.it.errors = sp.errors;
  hipLaunchKernelGGL(na_init_kernel, dim3(std::min(nblocks, 4096)), dim3(kBlock), 0, st, m,
                     d_scores, d_outdeg, sp.c0, sp.st, o.max_iter, o.tol);
  GD_HIP_TRY(hipGetLastError());
  const int chunk = o.chunk > 0 ? o.chunk : (o.tol > 0 ? 8 : std::max(1, o.max_iter));
  IterState h;
  memset(&h, 0, sizeof(h));
  int launched = 0;
  while (launched < o.max_iter) {
    const int n = std::min(chunk, o.max_iter - launched);
    for (int k = 0; k < n; k++)
      hipLaunchKernelGGL(na_pull_kernel, dim3(nblocks), dim3(kBlock), 0, st, a);
    GD_HIP_TRY(hipGetLastError());
    launched += n;
    GD_HIP_TRY(hipMemcpyAsync(&h, sp.st, sizeof(IterState), hipMemcpyDeviceToHost, st));
    GD_HIP_TRY(hipStreamSynchronize(st));
    if (h.done) break;
  }
  s.iterations = h.iter;
  if (h.iter > 0) {
    double last = 0.0;
    GD_HIP_TRY(hipMemcpyAsync(&last, sp.errors + (h.iter - 1), sizeof(double),
                              hipMemcpyDeviceToHost, st));
    if (errors)
      GD_HIP_TRY(hipMemcpyAsync(errors, sp.errors, sizeof(double) * h.iter,
                                hipMemcpyDeviceToHost, st));
    GD_HIP_TRY(hipStreamSynchronize(st));
    s.final_error = last;
    s.converged = last < o.tol;
  }
  s.solve_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  if (stats) *stats = s;
  return GD_OK;
}
