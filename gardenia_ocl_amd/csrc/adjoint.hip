// adjoint.hip -- differentiable weighted PageRank and SpMV backward (gfx950).
//
// Weighted PageRank (no counterpart in the reference, whose PageRank is
// unweighted): with raw edge weights a_uv (stored in in-CSR order, row v,
// column u) and out-strengths s_u = sum_v a_uv,
//     x = (1-d)/m + d P x,          P[v,u] = a_uv / s_u,
// which for a == 1 is exactly the reference iteration (s_u = out-degree).
// For a loss L(x) with dL/dx = w, the fixed point is differentiated
// implicitly: the adjoint
//     lambda = w + d P^T lambda
// is iterated by the same pull engine over the transposed graph (a CSC built
// once per plan), and
//     dL/da_uv = d * (x_u / s_u) * (lambda_v - q_u),   q = P^T lambda,
// where the q_u term is the derivative through the row normalisation.  Memory
// stays O(V + E): no forward iterates are stored.
//
// SpMV backward for y = A x:  dx = A^T g  (the SpMV tile engine over the CSC)
// and dA[k] = g[row(k)] * x[col(k)]  (an edge kernel over the CSR).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "iter_engine.h"
#include "pull_engine.h"
#include "spmv_internal.h"
#include "transpose.h"

namespace gd {

// ---------------------------------------------------------------- policies

struct WGather {
  const float* __restrict__ w;  // per-edge weight, in the order of the pulled CSR
  const float* __restrict__ c;  // gathered vector
  __device__ __forceinline__ float value(int e, int u) const { return w[e] * c[u]; }
};

// forward: x_v = base + d * sum_u a_uv c_u,  c_v = x_v / s_v
template <int NT>
struct WprEpilogue {
  float* __restrict__ x;
  float* __restrict__ c_next;
  const float* __restrict__ s;
  float base;
  float damp;
  struct Pre {
    float old;
    float s;
  };
  __device__ __forceinline__ Pre pre(int v) const { return Pre{x[v], s[v]}; }
  __device__ __forceinline__ float init(const Pre&) const { return 0.0f; }
  __device__ __forceinline__ double finish(int v, float sum, const Pre& p) const {
    const float nv = base + damp * sum;
    x[v] = nv;
    c_next[v] = p.s > 0.0f ? nv / p.s : 0.0f;
    return (double)fabsf(nv - p.old);
  }
};

struct WprPolicy {
  // as plain PageRank: with unit weights the forward is its arithmetic exactly
  static constexpr bool kExactHubs = true;
  static constexpr bool kSeqAll = true;
  struct Args {
    float* x;
    float* c0;
    float* c1;
    const float* a;  // in-CSR order
    const float* s;
    float base;
    float damp;
  };
  template <int NT>
  __device__ static WGather gather(const Args& a, int odd) {
    return WGather{a.a, odd ? a.c1 : a.c0};
  }
  template <int NT>
  __device__ static WprEpilogue<NT> epilogue(const Args& a, int odd) {
    return WprEpilogue<NT>{a.x, odd ? a.c0 : a.c1, a.s, a.base, a.damp};
  }
};

// adjoint: lambda_u = w_u + d * (sum_v a_uv lambda_v) / s_u  (pull over the CSC)
template <int NT>
struct AdjEpilogue {
  float* __restrict__ lam;
  float* __restrict__ lam_next;  // gathered copy for the next iteration
  const float* __restrict__ s;
  const float* __restrict__ w;
  float damp;
  struct Pre {
    float old;
    float s;
    float w;
  };
  __device__ __forceinline__ Pre pre(int u) const { return Pre{lam[u], s[u], w[u]}; }
  __device__ __forceinline__ float init(const Pre&) const { return 0.0f; }
  __device__ __forceinline__ double finish(int u, float sum, const Pre& p) const {
    const float q = p.s > 0.0f ? sum / p.s : 0.0f;
    const float nl = p.w + damp * q;
    lam[u] = nl;
    lam_next[u] = nl;
    return (double)fabsf(nl - p.old);
  }
};

struct AdjPolicy {
  static constexpr bool kExactHubs = false;
  static constexpr bool kSeqAll = false;
  struct Args {
    float* lam;
    float* c0;
    float* c1;
    const float* at;  // CSC order
    const float* s;
    const float* w;
    float damp;
  };
  template <int NT>
  __device__ static WGather gather(const Args& a, int odd) {
    return WGather{a.at, odd ? a.c1 : a.c0};
  }
  template <int NT>
  __device__ static AdjEpilogue<NT> epilogue(const Args& a, int odd) {
    return AdjEpilogue<NT>{a.lam, odd ? a.c0 : a.c1, a.s, a.w, a.damp};
  }
};

// ---------------------------------------------------------------- small kernels

// s_u = sum of the CSC row u (one wavefront per row, fixed order).
__global__ __launch_bounds__(256) void row_sums_kernel(int n, const int32_t* __restrict__ off,
                                                       const float* __restrict__ val,
                                                       float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  for (int r = blockIdx.x * 4 + threadIdx.x / 64; r < n; r += gridDim.x * 4) {
    float s = 0.0f;
    for (int k = off[r] + lane; k < off[r + 1]; k += 64) s += val[k];
    s = wave_sum(s);
    if (lane == 0) out[r] = s;
  }
}

__global__ void div_or_zero_kernel(int n, const float* __restrict__ a, const float* __restrict__ b,
                                   float* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = b[i] > 0.0f ? a[i] / b[i] : 0.0f;
}

__global__ void copy_kernel(int n, const float* __restrict__ a, float* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = a[i];
}

__global__ void fill_kernel(int n, float* __restrict__ out, float v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = v;
}

// cq[u] = (x_u / s_u, q_u) with q_u = (sum_v a_uv lambda_v) / s_u.
__global__ void pack_cq_kernel(int n, const float* __restrict__ x, const float* __restrict__ s,
                               const float* __restrict__ atl, float2* __restrict__ cq) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float si = s[i];
    cq[i] = si > 0.0f ? make_float2(x[i] / si, atl[i] / si) : make_float2(0.0f, 0.0f);
  }
}

// Row tiles of at most this many edges label their edges' rows in LDS first.
constexpr int kMapEdges = 4096;

// Per-edge map over a CSR tile schedule: f(e, row, col) for every edge.  A row
// tile's edge belongs to the unique row i with s_off[i] <= e < s_off[i + 1]
// (s_off = the tile's offsets in LDS).  Up to kMapEdges edges, thread t finds
// it for edges [t * per, + per) by one binary search and a forward walk and
// stores it in a row table, so each edge then costs one LDS read; larger
// tiles binary-search every edge.  A launch with gridDim.y > 1 splits each
// one-row (hub) tile's edges over gridDim.y workgroups; row tiles need
// gridDim.y == 1.
template <class F>
__global__ __launch_bounds__(kBlock) void edge_map_kernel(const Tile* __restrict__ tiles,
                                                          const int32_t* __restrict__ off,
                                                          const int32_t* __restrict__ col, F f) {
  __shared__ int s_off[kTileRows + 1];
  __shared__ uint16_t s_row[kMapEdges];
  static_assert(kTileRows <= 65536, "row table holds 16-bit tile rows");
  const Tile t = tiles[blockIdx.x];
  if (t.y >= 0) {
    const int nr = t.y - t.x;
    for (int i = threadIdx.x; i <= nr; i += kBlock) s_off[i] = off[t.x + i];
    __syncthreads();
    const int n = t.w - t.z;
    if (n <= kMapEdges) {
      const int per = (n + kBlock - 1) / kBlock;
      const int e0 = t.z + threadIdx.x * per, e1 = min(e0 + per, t.w);
      if (e0 < e1) {
        int lo = 0, hi = nr;  // s_off[0] = t.z <= e0 < t.w = s_off[nr]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (s_off[mid] <= e0) lo = mid; else hi = mid;
        }
        int next = s_off[lo + 1];
        for (int e = e0; e < e1; e++) {
          while (next <= e) next = s_off[++lo + 1];  // stops below nr: s_off[nr] > e
          s_row[e - t.z] = (uint16_t)lo;
        }
      }
      __syncthreads();
      for (int e = t.z + threadIdx.x; e < t.w; e += kBlock) f(e, t.x + s_row[e - t.z], col[e]);
      return;
    }
    for (int e = t.z + threadIdx.x; e < t.w; e += kBlock) {
      int lo = 0, hi = nr;  // largest i with s_off[i] <= e
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= e) lo = mid; else hi = mid;
      }
      f(e, t.x + lo, col[e]);
    }
  } else {
    for (int e = t.z + blockIdx.y * kBlock + threadIdx.x; e < t.w; e += kBlock * gridDim.y)
      f(e, t.x, col[e]);
  }
}

struct WprGradOp {
  const float2* __restrict__ cq;
  const float* __restrict__ lam;
  float* __restrict__ grad;
  float damp;
  __device__ __forceinline__ void operator()(int e, int v, int u) const {
    const float2 p = cq[u];
    grad[e] = damp * p.x * (lam[v] - p.y);
  }
};

struct SpmvDaOp {
  const float* __restrict__ x;
  const float* __restrict__ g;
  float* __restrict__ dA;
  __device__ __forceinline__ void operator()(int k, int r, int c) const { dA[k] = g[r] * x[c]; }
};

// the same product with x[c] read from its slice-major copy: Xs[inv[k]] = x[c]
struct SpmvDaSlicedOp {
  const float* __restrict__ xs;
  const int32_t* __restrict__ inv;
  const float* __restrict__ g;
  float* __restrict__ dA;
  __device__ __forceinline__ void operator()(int k, int r, int) const { dA[k] = g[r] * xs[inv[k]]; }
};

static int grid_n(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }

}  // namespace gd

using namespace gd;

// ================================================================ weighted PR

struct gd_wpr_plan {
  int m = 0;
  int64_t nnz = 0;
  hipStream_t stream = nullptr;
  const int32_t* off = nullptr;
  const int32_t* col = nullptr;
  const float* a = nullptr;  // caller's weights, in-CSR order
  Transpose tr;              // CSC of the in-CSR = out-CSR of the graph
  DeviceBuffer at, s, c0, c1, atl, cq, lam_tmp;
  IterEngine fwd, adj;
  gd_spmv_plan* qplan = nullptr;  // y = A_t lambda over the CSC
};

namespace {

int wpr_refresh_weights(gd_wpr_plan* p) {
  int rc = transpose_values(p->tr, p->nnz, p->a, p->at.as<float>(), p->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(row_sums_kernel, dim3(grid_n(p->m / 4 + 1)), dim3(256), 0, p->stream, p->m,
                     p->tr.off.as<int32_t>(), p->at.as<float>(), p->s.as<float>());
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

}  // namespace

extern "C" {

int gd_wpr_plan_create(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                       const float* d_weight, void* stream, gd_wpr_plan** out) {
  GD_CHECK_ARG(out != nullptr, "plan output pointer is NULL");
  *out = nullptr;
  GD_CHECK_ARG(m > 0 && nnz >= 0 && nnz < (1ll << 31), "m/nnz out of range");
  GD_CHECK_ARG(d_off && (nnz == 0 || (d_col && d_weight)), "NULL device array");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  std::unique_ptr<gd_wpr_plan> owner(new gd_wpr_plan());
  gd_wpr_plan* p = owner.get();
  p->m = m;
  p->nnz = nnz;
  p->stream = (hipStream_t)stream;
  p->off = d_off;
  p->col = d_col;
  p->a = d_weight;
  auto fail = [](int rc) { return rc; };  // owner frees the plan
  int rc = check_columns(d_col, nnz, m, p->stream);
  if (rc) return fail(rc);
  rc = check_offsets(d_off, m, nnz, p->stream);  // before any device walk
  if (rc) return fail(rc);
  rc = build_transpose(m, m, nnz, d_off, d_col, p->stream, &p->tr);
  if (!rc) rc = p->at.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (!rc) rc = p->s.alloc(4 * (size_t)m);
  if (!rc) rc = p->c0.alloc(4 * (size_t)m);
  if (!rc) rc = p->c1.alloc(4 * (size_t)m);
  if (!rc) rc = p->atl.alloc(4 * (size_t)m);
  if (!rc) rc = p->cq.alloc(8 * (size_t)m);
  if (rc) return fail(rc);
  for (IterEngine* e : {&p->fwd, &p->adj}) {
    e->ipt = 8;
    e->nt = 0;
    e->persistent = false;
  }
  // The forward sums every row in edge order (bit-exact), so it takes the
  // plain PageRank plan's tiles: 4096 edges, gathered in source order, hub
  // rows above 4096 edges.  The adjoint tree-sums its longer rows: sorting or
  // resizing its tiles changes that rounding, and at RMAT-20, tol 1e-7, took
  // 145 iterations instead of 128 at the same time per iteration.
  p->fwd.ipt = 16;
  p->fwd.sort_tiles = true;
  p->fwd.exact_len = 4096;  // hub rows in edge order, as the plain PageRank plan
  rc = p->fwd.build(m, nnz, d_off, d_col, p->stream);
  if (!rc) rc = p->fwd.setup_grid<WprPolicy>();
  if (!rc) rc = p->adj.build(m, nnz, p->tr.off.as<int32_t>(), p->tr.row.as<int32_t>(), p->stream);
  if (!rc) rc = p->adj.setup_grid<AdjPolicy>();
  if (!rc)
    rc = gd_spmv_plan_create(m, m, nnz, p->tr.off.as<int32_t>(), p->tr.row.as<int32_t>(),
                             p->at.as<float>(), stream, &p->qplan);
  if (rc) return fail(rc);
  *out = owner.release();
  return GD_OK;
}

int gd_wpr_plan_destroy(gd_wpr_plan* p) {
  if (!p) return GD_OK;
  gd_spmv_plan_destroy(p->qplan);
  delete p;
  return GD_OK;
}

int gd_wpr_forward(gd_wpr_plan* p, float* d_x, const gd_pr_options* opts, gd_pr_stats* stats,
                   double* errors) {
  GD_CHECK_ARG(p != nullptr && d_x != nullptr, "NULL plan / x");
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  GD_CHECK_ARG(o.max_iter >= 0, "max_iter must be >= 0");
  gd_pr_stats s;
  memset(&s, 0, sizeof(s));
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  GD_HIP_TRY(hipEventRecord(t0, p->stream));
  int rc = wpr_refresh_weights(p);  // weights may have changed since the last call
  if (rc) return rc;
  hipLaunchKernelGGL(div_or_zero_kernel, dim3(grid_n(p->m)), dim3(256), 0, p->stream, p->m, d_x,
                     p->s.as<float>(), p->c0.as<float>());
  const float damp = o.damping;
  const WprPolicy::Args a{d_x, p->c0.as<float>(), p->c1.as<float>(), p->a, p->s.as<float>(),
                          (1.0f - damp) / (float)p->m, damp};
  rc = p->fwd.run<WprPolicy>(a, o.max_iter, o.tol, o.chunk, o.profile != 0, &s, errors);
  if (rc) return rc;
  GD_HIP_TRY(hipEventRecord(t1, p->stream));
  GD_HIP_TRY(hipEventSynchronize(t1));
  GD_HIP_TRY(hipEventElapsedTime(&s.solve_ms, t0, t1));
  if (stats) *stats = s;
  return GD_OK;
}

int gd_wpr_backward(gd_wpr_plan* p, const float* d_x, const float* d_w, float* d_lambda,
                    float* d_grad, const gd_pr_options* opts, gd_pr_stats* stats) {
  GD_CHECK_ARG(p != nullptr && d_x && d_w && d_lambda, "NULL plan / x / w / lambda");
  GD_CHECK_ARG(p->nnz == 0 || d_grad, "NULL grad");
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  GD_CHECK_ARG(o.max_iter >= 0, "max_iter must be >= 0");
  gd_pr_stats s;
  memset(&s, 0, sizeof(s));
  const int m = p->m;
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  GD_HIP_TRY(hipEventRecord(t0, p->stream));
  int rc = wpr_refresh_weights(p);
  if (rc) return rc;
  // lambda_0 = w
  hipLaunchKernelGGL(copy_kernel, dim3(grid_n(m)), dim3(256), 0, p->stream, m, d_w, d_lambda);
  hipLaunchKernelGGL(copy_kernel, dim3(grid_n(m)), dim3(256), 0, p->stream, m, d_w,
                     p->c0.as<float>());
  const AdjPolicy::Args a{d_lambda, p->c0.as<float>(), p->c1.as<float>(), p->at.as<float>(),
                          p->s.as<float>(), d_w, o.damping};
  rc = p->adj.run<AdjPolicy>(a, o.max_iter, o.tol, o.chunk, o.profile != 0, &s, nullptr);
  if (rc) return rc;
  if (p->nnz > 0) {
    // q = (A_t lambda) / s,  cq = (x / s, q)
    hipLaunchKernelGGL(fill_kernel, dim3(grid_n(m)), dim3(256), 0, p->stream, m,
                       p->atl.as<float>(), 0.0f);
    rc = gd_spmv_plan_run(p->qplan, d_lambda, p->atl.as<float>());
    if (rc) return rc;
    hipLaunchKernelGGL(pack_cq_kernel, dim3(grid_n(m)), dim3(256), 0, p->stream, m, d_x,
                       p->s.as<float>(), p->atl.as<float>(), p->cq.as<float2>());
    const WprGradOp op{p->cq.as<float2>(), d_lambda, d_grad, o.damping};
    // every edge once: the forward's row tiles, then its whole-row hub tiles
    if (p->fwd.n_tiles)
      hipLaunchKernelGGL(edge_map_kernel<WprGradOp>, dim3(p->fwd.n_tiles), dim3(kBlock), 0,
                         p->stream, p->fwd.tiles.as<Tile>(), p->off, p->col, op);
    if (p->fwd.n_hub_tiles) {
      // one workgroup per 8 edges per thread of the longest hub
      const int split = std::min<int64_t>(
          64, std::max<int64_t>(1, ((int64_t)p->fwd.sched.max_row_len + kBlock * 8 - 1) /
                                       (kBlock * 8)));
      hipLaunchKernelGGL(edge_map_kernel<WprGradOp>, dim3(p->fwd.n_hub_tiles, split),
                         dim3(kBlock), 0, p->stream, p->fwd.hub_tiles.as<Tile>(), p->off, p->col,
                         op);
    }
    GD_HIP_TRY(hipGetLastError());
  }
  GD_HIP_TRY(hipEventRecord(t1, p->stream));
  GD_HIP_TRY(hipEventSynchronize(t1));
  GD_HIP_TRY(hipEventElapsedTime(&s.solve_ms, t0, t1));
  if (stats) *stats = s;
  return GD_OK;
}

}  // extern "C"

// ================================================================ SpMV backward

struct gd_spmv_bwd_plan {
  int rows = 0, cols = 0;
  int64_t nnz = 0;
  hipStream_t stream = nullptr;
  const int32_t* Ap = nullptr;
  const int32_t* Aj = nullptr;
  Transpose tr;
  DeviceBuffer at;
  Schedule sched;
  DeviceBuffer tiles;
  int n_tiles = 0;
  gd_spmv_plan* tplan = nullptr;  // SpMV over the CSC: dx = A^T g
  // tplan in the column-sliced mode: its slice-major value copy is filled from
  // A's values by vperm (two passes through buckets) instead of `at`
  bool sliced = false;
  BucketedPermutation vperm;  // A's values (CSR order) -> the CSC plan's slice-major copy
  // x larger than one XCD's L2 and A's rows sorted: x gathered in A's
  // column-sliced order (fplan; slice s on XCD s, so the gathers hit L2), then
  // dA over A's tiles reads x[c] from that copy (inv[k] = slice-major position
  // of entry k: eight sequential streams, one per slice).  The copy lives in
  // fplan's slice-major value array, which fplan itself never reads (it is
  // never run as an SpMV).
  gd_spmv_plan* fplan = nullptr;
  DeviceBuffer inv;
  ~gd_spmv_bwd_plan() {
    gd_spmv_plan_destroy(tplan);
    gd_spmv_plan_destroy(fplan);
  }
};

namespace gd {
namespace {

// keys[k] = s * nb + (col[k] >> cshift) for the CSR entry k of row r, where
// s = min(r / width, 7) is the column slice of the CSC plan holding the entry:
// a bucket is one slice's entries in a range of 2^cshift columns of A.
__global__ void bucket_keys_kernel(int rows, const int32_t* __restrict__ off,
                                   const int32_t* __restrict__ col, int width, int cshift, int nb,
                                   int32_t* __restrict__ keys) {
  const int lane = threadIdx.x & 63;
  for (int r = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; r < rows;
       r += gridDim.x * (blockDim.x / 64)) {
    const int base = min(r / width, 7) * nb;
    for (int e = off[r] + lane; e < off[r + 1]; e += 64) keys[e] = base + (col[e] >> cshift);
  }
}

__global__ void iota_kernel(int64_t n, int32_t* __restrict__ a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (int32_t)i;
}

// inv[fwd[i]] = i
__global__ void invert_perm_kernel(int64_t n, const int32_t* __restrict__ fwd,
                                   int32_t* __restrict__ inv) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    inv[fwd[i]] = (int32_t)i;
}

}  // namespace
}  // namespace gd

extern "C" {

int gd_spmv_bwd_plan_create(int rows, int cols, int64_t nnz, const int32_t* d_Ap,
                            const int32_t* d_Aj, void* stream, gd_spmv_bwd_plan** out) {
  GD_CHECK_ARG(out != nullptr, "plan output pointer is NULL");
  *out = nullptr;
  GD_CHECK_ARG(rows >= 0 && cols >= 0 && nnz >= 0 && nnz < (1ll << 31), "size out of range");
  GD_CHECK_ARG((rows == 0 || d_Ap) && (nnz == 0 || d_Aj), "NULL device array");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  std::unique_ptr<gd_spmv_bwd_plan> owner(new gd_spmv_bwd_plan());
  gd_spmv_bwd_plan* p = owner.get();
  p->rows = rows;
  p->cols = cols;
  p->nnz = nnz;
  p->stream = (hipStream_t)stream;
  p->Ap = d_Ap;
  p->Aj = d_Aj;
  auto fail = [](int rc) { return rc; };  // owner frees the plan
  int rc = check_columns(d_Aj, nnz, cols, p->stream);
  if (rc) return fail(rc);
  if (rows > 0) {
    std::vector<int32_t> h_off(rows + 1);
    GD_HIP_TRY(hipMemcpyAsync(h_off.data(), d_Ap, 4 * ((size_t)rows + 1), hipMemcpyDeviceToHost,
                              p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    rc = build_schedule(rows, nnz, h_off.data(), kBlock * 8, kTileRows, &p->sched);
    if (rc) return fail(rc);
    p->n_tiles = (int)p->sched.tiles.size();
    rc = p->tiles.alloc(sizeof(Tile) * p->n_tiles);
    if (rc) return fail(rc);
    GD_HIP_TRY(hipMemcpyAsync(p->tiles.ptr, p->sched.tiles.data(), sizeof(Tile) * p->n_tiles,
                              hipMemcpyHostToDevice, p->stream));
  }
  rc = build_transpose(rows, cols, nnz, d_Ap, d_Aj, p->stream, &p->tr);
  if (!rc) rc = p->at.alloc(4 * (size_t)std::max<int64_t>(nnz, 1));
  if (rc) return fail(rc);
  GD_HIP_TRY(hipMemsetAsync(p->at.ptr, 0, 4 * (size_t)std::max<int64_t>(nnz, 1), p->stream));
  if (cols > 0) {
    // g larger than one XCD's L2: the column-sliced mode (the CSC's rows are
    // sorted; a matrix with repeated entries in a row is not, and keeps the
    // tile mode)
    if (nnz > 0 && (int64_t)rows * 4 > (4 << 20)) {
      rc = gd_spmv_plan_create_ex(cols, rows, nnz, p->tr.off.as<int32_t>(),
                                  p->tr.row.as<int32_t>(), p->at.as<float>(), stream,
                                  GD_SPMV_SLICED, &p->tplan);
      if (rc == GD_ERR_INVALID) rc = GD_OK;  // unsorted: p->tplan stays NULL
      if (rc) return fail(rc);
    }
    if (p->tplan) {
      // The values go into the CSC plan's slice-major copy through buckets of
      // about 2^18 entries (a slice's entries in a range of A's columns): the
      // slice-major order reads the buckets one after another.
      p->sliced = true;
      DeviceBuffer perm_s, keys;
      rc = perm_s.alloc(4 * (size_t)nnz);
      if (!rc) rc = keys.alloc(4 * (size_t)nnz);
      if (!rc) rc = spmv_slice_order(p->tplan, p->tr.perm.as<int32_t>(), perm_s.as<int32_t>());
      const int64_t nb_want = std::max<int64_t>(1, nnz / (8ll << 18));
      int cshift = 0;
      while (((int64_t)cols >> cshift) > nb_want) cshift++;
      const int nb = (int)(((int64_t)cols + (1ll << cshift) - 1) >> cshift);
      int key_bits = 1;
      while ((1ll << key_bits) < 8ll * nb) key_bits++;
      if (!rc) {
        const int width = std::max(1, (rows + 7) / 8);  // the CSC plan's slice width
        hipLaunchKernelGGL(bucket_keys_kernel, dim3(grid_n(rows / 4 + 1)), dim3(256), 0,
                           p->stream, rows, d_Ap, d_Aj, width, cshift, nb, keys.as<int32_t>());
        if (hipGetLastError() != hipSuccess) rc = GD_ERR_DEVICE;
      }
      if (!rc)
        rc = build_bucketed_permutation(nnz, keys.as<int32_t>(), key_bits, perm_s.as<int32_t>(),
                                        p->stream, &p->vperm);
    } else {
      rc = gd_spmv_plan_create(cols, rows, nnz, p->tr.off.as<int32_t>(), p->tr.row.as<int32_t>(),
                               p->at.as<float>(), stream, &p->tplan);
    }
  }
  if (rc) return fail(rc);
  if (rows > 0 && nnz > 0 && (int64_t)cols * 4 > (4 << 20)) {
    // the sliced layout's value copy (of `at`, zeros) becomes the x copy
    rc = gd_spmv_plan_create_ex(rows, cols, nnz, d_Ap, d_Aj, p->at.as<float>(), stream,
                                GD_SPMV_SLICED, &p->fplan);
    if (rc == GD_ERR_INVALID) rc = GD_OK;  // unsorted rows: the edge map below
    if (!rc && p->fplan) {
      DeviceBuffer ids, kmap;
      rc = ids.alloc(4 * (size_t)nnz);
      if (!rc) rc = kmap.alloc(4 * (size_t)nnz);
      if (!rc) rc = p->inv.alloc(4 * (size_t)nnz);
      if (!rc) {
        hipLaunchKernelGGL(iota_kernel, dim3(grid_n(nnz)), dim3(256), 0, p->stream, nnz,
                           ids.as<int32_t>());
        rc = spmv_slice_order(p->fplan, ids.as<int32_t>(), kmap.as<int32_t>());
      }
      if (!rc) {
        hipLaunchKernelGGL(invert_perm_kernel, dim3(grid_n(nnz)), dim3(256), 0, p->stream, nnz,
                           kmap.as<int32_t>(), p->inv.as<int32_t>());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(p->stream) != hipSuccess)
          rc = GD_ERR_DEVICE;
      }
    }
    if (rc) return fail(rc);
  }
  *out = owner.release();
  return GD_OK;
}

int gd_spmv_bwd_run(gd_spmv_bwd_plan* p, const float* d_Ax, const float* d_x, const float* d_g,
                    float* d_dx, float* d_dA) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  GD_CHECK_ARG(p->rows == 0 || d_g, "NULL g");
  if (d_dx && p->cols > 0) {
    GD_CHECK_ARG(p->nnz == 0 || d_Ax, "NULL Ax");
    int rc = p->sliced ? run_bucketed_permutation(p->vperm, d_Ax, spmv_sliced_values(p->tplan),
                                                  p->stream)
                       : transpose_values(p->tr, p->nnz, d_Ax, p->at.as<float>(), p->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_n(p->cols)), dim3(256), 0, p->stream, p->cols,
                       d_dx, 0.0f);
    rc = gd_spmv_plan_run(p->tplan, d_g, d_dx);
    if (rc) return rc;
  }
  if (d_dA && p->nnz > 0) {
    GD_CHECK_ARG(d_x != nullptr, "NULL x");
    if (p->fplan) {
      float* xs = spmv_sliced_values(p->fplan);
      const int rc = spmv_sliced_gather(p->fplan, d_x, xs);
      if (rc) return rc;
      const SpmvDaSlicedOp op{xs, p->inv.as<int32_t>(), d_g, d_dA};
      hipLaunchKernelGGL(edge_map_kernel<SpmvDaSlicedOp>, dim3(p->n_tiles), dim3(kBlock), 0,
                         p->stream, p->tiles.as<Tile>(), p->Ap, p->Aj, op);
    } else {
      const SpmvDaOp op{d_x, d_g, d_dA};
      hipLaunchKernelGGL(edge_map_kernel<SpmvDaOp>, dim3(p->n_tiles), dim3(kBlock), 0, p->stream,
                         p->tiles.as<Tile>(), p->Ap, p->Aj, op);
    }
  }
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

int gd_spmv_bwd_plan_destroy(gd_spmv_bwd_plan* p) {
  delete p;  // the destructor frees the CSC plan
  return GD_OK;
}

}  // extern "C"
