// symgs.hip -- multicolor symmetric Gauss-Seidel sweep (the reference's
// SymGSSolver, src/symgs/symgs.h:28; algorithm src/symgs/omp_base.cc:8-39 and
// src/symgs/verifier.cc:50-77), bit-exact with the sequential sweep.
//
// Semantics.  For each color c = 0..C-1, then C-1..0: every row r listed in
// indices[color_offsets[c] .. color_offsets[c+1]) gets
//     rsum = sum over the row's off-diagonal entries, in CSR order, of
//            x[col] * val (product rounded, then added: sequential FP32);
//     diag = the value of the row's LAST entry with column == r (0 if none);
//     x[r] = (b[r] - rsum) / diag   unless diag == 0.
// Rows of one color never reference each other (a proper coloring, checked
// when the plan is built), so updating them in parallel equals the
// reference's sequential loop over the color, bit for bit.
//
// Layout in HBM (built once per plan).  The rows are gathered in sweep order
// into a private CSR without diagonal entries: row i of the plan is row
// indices[color_offsets[0] + i] of the matrix, so every color is a contiguous
// range of plan rows and of edges, with diag[i] and orig[i] beside it.  Each
// color is cut into the pull engine's LDS row tiles (rows of <= 64 off-diagonal
// entries, summed by their owner thread in CSR order) and whole-row tiles
// (longer rows: a workgroup gathers each 2048-entry chunk into LDS while one
// lane adds the previous chunk in order), with a SymGS gather
// (x[col] * val) and epilogue (the row update).  One color = one launch of
// symgs_color_kernel; the stream orders the colors.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "pull_engine.h"
#include "pull_iter.h"

namespace gd {
namespace {

constexpr int kSgIPT = 8;  // edges per thread per row tile (tile = 2048 edges)

struct SymgsGather {
  const float* __restrict__ x;
  const float* __restrict__ val;
  // x[j] * Ax[jj] (src/symgs/verifier.cc:61): rounded here, added in order later
  __device__ __forceinline__ float value(int e, int u) const { return x[u] * val[e]; }
};

struct SymgsEpilogue {
  float* x;
  const float* __restrict__ b;
  const float* __restrict__ diag;
  const int32_t* __restrict__ orig;
  struct Pre {
    float d;
    float bv;
    int r;
  };
  __device__ __forceinline__ Pre pre(int v) const {
    const int r = orig[v];
    return Pre{diag[v], b[r], r};
  }
  __device__ __forceinline__ float init(const Pre&) const { return 0.0f; }
  // if (diag != 0) x[inew] = (b[inew] - rsum) / diag   (verifier.cc:63)
  __device__ __forceinline__ double finish(int /*v*/, float rsum, const Pre& p) const {
    if (p.d != 0.0f) x[p.r] = (p.bv - rsum) / p.d;
    return 0.0;
  }
};

struct SymgsPolicy {
  struct Args {
    float* x;
    const float* b;
    const float* val;
    const float* diag;
    const int32_t* orig;
  };
  template <int NT>
  __device__ static SymgsGather gather(const Args& a, int) {
    return SymgsGather{a.x, a.val};
  }
  template <int NT>
  __device__ static SymgsEpilogue epilogue(const Args& a, int) {
    return SymgsEpilogue{a.x, a.b, a.diag, a.orig};
  }
};


// One color in one launch: workgroups [0, n_hub_tiles) add one long row each
// (hub_chain_sum), the rest one LDS row tile each (owner thread per row, in
// order).  Long rows first: they take longest.  LONG = false (colors without
// long rows) leaves the long-row code, and its registers, out of the kernel.
template <bool LONG>
__global__ __launch_bounds__(kBlock) void symgs_color_kernel(const SymgsPolicy::Args a,
                                                             const IterArgs it) {
  // a row tile's LDS, or a long row's two chunk buffers + the broadcast sum
  static_assert(sizeof(RowTileLds<kSgIPT>) >= (2 * kHubChunk + 1) * sizeof(float), "LDS union");
  __shared__ __align__(16) unsigned char smem[sizeof(RowTileLds<kSgIPT>)];
  RowTileLds<kSgIPT>& lds = *reinterpret_cast<RowTileLds<kSgIPT>*>(smem);
  float* const buf = reinterpret_cast<float*>(smem);
  const auto g = SymgsPolicy::gather<0>(a, 0);
  const auto e = SymgsPolicy::epilogue<0>(a, 0);
  const int b = blockIdx.x;
  if (LONG && b < it.n_hub_tiles) {
    const Tile t = it.hub_tiles[b];
    const auto pre = e.pre(t.x);
    // 16 float4 LDS reads ahead: the chains bound these sweeps (8 ahead: 9% slower)
    const float rsum = hub_chain_sum<16, false>(it.col, nullptr, it.e_max, t.z, t.w - t.z, g, 0.0f,
                                               buf, buf + kHubChunk, buf + 2 * kHubChunk);
    if (threadIdx.x == 0) (void)e.finish(t.x, rsum, pre);
  } else {
    const Tile t = it.tiles[b - it.n_hub_tiles];
    (void)row_tile<kSgIPT, false, false>(it.off, it.col, nullptr, it.e_max, t.x, t.y, t.z,
                                         t.w - t.z, g, e, lds);
  }
}

#define SG_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

int sg_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// Plan row i = matrix row idx[i]: range check, its color (binary search in
// the color offsets), and each matrix row listed at most once.
__global__ void sg_rows_kernel(int n_p, const int32_t* __restrict__ idx, int m,
                               const int32_t* __restrict__ co, int n_colors,
                               int32_t* __restrict__ color_of, int* __restrict__ bad) {
  SG_LOOP(i, n_p) {
    const int r = idx[i];
    if (r < 0 || r >= m) {
      atomicOr(bad, 1);
      continue;
    }
    const int pos = co[0] + (int)i;
    int lo = 0, hi = n_colors;  // last c with co[c] <= pos
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (co[mid] <= pos)
        lo = mid;
      else
        hi = mid;
    }
    if (atomicCAS(color_of + r, -1, lo) != -1) atomicOr(bad, 2);
  }
}

// Off-diagonal count, last diagonal value, original row.
__global__ void sg_count_kernel(int n_p, const int32_t* __restrict__ idx,
                                const int32_t* __restrict__ Ap, const int32_t* __restrict__ Aj,
                                const float* __restrict__ Ax, int32_t* __restrict__ len,
                                float* __restrict__ diag, int32_t* __restrict__ orig) {
  SG_LOOP(i, n_p) {
    const int r = idx[i];
    int n = 0;
    float d = 0.0f;
    for (int k = Ap[r]; k < Ap[r + 1]; k++) {
      if (Aj[k] == r)
        d = Ax[k];
      else
        n++;
    }
    len[i] = n;
    diag[i] = d;
    orig[i] = r;
  }
}

__global__ void sg_fill_kernel(int n_p, const int32_t* __restrict__ idx,
                               const int32_t* __restrict__ Ap, const int32_t* __restrict__ Aj,
                               const float* __restrict__ Ax, const int32_t* __restrict__ off,
                               int32_t* __restrict__ col, float* __restrict__ val) {
  SG_LOOP(i, n_p) {
    const int r = idx[i];
    int o = off[i];
    for (int k = Ap[r]; k < Ap[r + 1]; k++) {
      const int j = Aj[k];
      if (j != r) {
        col[o] = j;
        val[o] = Ax[k];
        o++;
      }
    }
  }
}

// Proper coloring: no off-diagonal entry joins two rows of one color.
__global__ void sg_proper_kernel(int n_p, const int32_t* __restrict__ off,
                                 const int32_t* __restrict__ col, const int32_t* __restrict__ orig,
                                 const int32_t* __restrict__ color_of, int* __restrict__ bad) {
  SG_LOOP(i, n_p) {
    const int c = color_of[orig[i]];
    int f = 0;
    for (int k = off[i]; k < off[i + 1]; k++) f |= color_of[col[k]] == c;
    if (f) atomicOr(bad, 4);
  }
}

}  // namespace
}  // namespace gd

using namespace gd;

struct gd_symgs_plan {
  int m = 0;
  int n_p = 0;  // rows swept (color_offsets[C] - color_offsets[0])
  int64_t nnz_p = 0;  // off-diagonal entries of those rows
  int n_colors = 0;
  int ipt = kSgIPT;
  double plan_ms = 0.0;
  DeviceBuffer off, col, val, diag, orig, tiles, hubs;
  std::vector<int> tile_begin, hub_begin;  // per color, n_colors + 1 entries
  int max_row_len = 0;
};

namespace {

int read_flag(const DeviceBuffer& b, hipStream_t s, int* h) {
  GD_HIP_TRY(hipMemcpyAsync(h, b.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
  GD_HIP_TRY(hipStreamSynchronize(s));
  return GD_OK;
}

int plan_build(gd_symgs_plan* p, int m, int64_t nnz, const int32_t* d_Ap, const int32_t* d_Aj,
               const float* d_Ax, const int32_t* d_idx, const int32_t* co, int C, hipStream_t s) {
  p->m = m;
  p->n_colors = C;
  const int n_p = co[C] - co[0];
  p->n_p = n_p;
  int rc = check_columns(d_Aj, nnz, m, s);
  if (rc) return rc;
  {  // the matrix offsets must be a CSR prefix sum
    std::vector<int32_t> h(m + 1);
    GD_HIP_TRY(hipMemcpyAsync(h.data(), d_Ap, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost, s));
    GD_HIP_TRY(hipStreamSynchronize(s));
    GD_CHECK_ARG(h[0] == 0 && (int64_t)h[m] == nnz, "Ap[0] must be 0 and Ap[m] must equal nnz");
    for (int i = 0; i < m; i++) GD_CHECK_ARG(h[i + 1] >= h[i], "Ap must be non-decreasing");
  }
  const size_t np1 = (size_t)n_p + 1;
  DeviceBuffer d_co, color_of, bad, len;
  rc = d_co.alloc(4 * ((size_t)C + 1));
  if (!rc) rc = color_of.alloc(4 * (size_t)std::max(m, 1));
  if (!rc) rc = bad.alloc(sizeof(int));
  if (!rc) rc = len.alloc(4 * np1);
  if (!rc) rc = p->off.alloc(4 * np1);
  if (!rc) rc = p->diag.alloc(4 * std::max<size_t>(n_p, 1));
  if (!rc) rc = p->orig.alloc(4 * std::max<size_t>(n_p, 1));
  if (rc) return rc;
  GD_HIP_TRY(hipMemcpyAsync(d_co.ptr, co, 4 * ((size_t)C + 1), hipMemcpyHostToDevice, s));
  GD_HIP_TRY(hipMemsetAsync(color_of.ptr, 0xff, 4 * (size_t)std::max(m, 1), s));  // -1
  GD_HIP_TRY(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
  GD_HIP_TRY(hipMemsetAsync(len.ptr, 0, 4 * np1, s));
  int h_bad = 0;
  if (n_p > 0) {
    hipLaunchKernelGGL(sg_rows_kernel, dim3(sg_grid(n_p)), dim3(256), 0, s, n_p, d_idx + co[0], m,
                       d_co.as<int32_t>(), C, color_of.as<int32_t>(), bad.as<int>());
    GD_HIP_TRY(hipGetLastError());
    rc = read_flag(bad, s, &h_bad);
    if (rc) return rc;
    GD_CHECK_ARG(!(h_bad & 1), "indices: row outside [0, m)");
    GD_CHECK_ARG(!(h_bad & 2), "indices: a row is listed more than once");
    hipLaunchKernelGGL(sg_count_kernel, dim3(sg_grid(n_p)), dim3(256), 0, s, n_p, d_idx + co[0],
                       d_Ap, d_Aj, d_Ax, len.as<int32_t>(), p->diag.as<float>(),
                       p->orig.as<int32_t>());
    GD_HIP_TRY(hipGetLastError());
  }
  {
    size_t tb = 0;
    GD_HIP_TRY(rocprim::exclusive_scan(nullptr, tb, len.as<int32_t>(), p->off.as<int32_t>(), 0,
                                       (size_t)np1, rocprim::plus<int32_t>(), s));
    DeviceBuffer tmp;
    rc = tmp.alloc(std::max<size_t>(tb, 1));
    if (rc) return rc;
    GD_HIP_TRY(rocprim::exclusive_scan(tmp.ptr, tb, len.as<int32_t>(), p->off.as<int32_t>(), 0,
                                       (size_t)np1, rocprim::plus<int32_t>(), s));
    GD_HIP_TRY(hipStreamSynchronize(s));
  }
  std::vector<int32_t> h_off(np1);
  GD_HIP_TRY(hipMemcpyAsync(h_off.data(), p->off.ptr, 4 * np1, hipMemcpyDeviceToHost, s));
  GD_HIP_TRY(hipStreamSynchronize(s));
  p->nnz_p = h_off[n_p];
  const size_t ne = (size_t)std::max<int64_t>(p->nnz_p, 1);
  rc = p->col.alloc(4 * ne);
  if (!rc) rc = p->val.alloc(4 * ne);
  if (rc) return rc;
  GD_HIP_TRY(hipMemsetAsync(p->col.ptr, 0, 4 * ne, s));  // the 1-element dummy when empty
  GD_HIP_TRY(hipMemsetAsync(p->val.ptr, 0, 4 * ne, s));
  if (n_p > 0) {
    hipLaunchKernelGGL(sg_fill_kernel, dim3(sg_grid(n_p)), dim3(256), 0, s, n_p, d_idx + co[0],
                       d_Ap, d_Aj, d_Ax, p->off.as<int32_t>(), p->col.as<int32_t>(),
                       p->val.as<float>());
    hipLaunchKernelGGL(sg_proper_kernel, dim3(sg_grid(n_p)), dim3(256), 0, s, n_p,
                       p->off.as<int32_t>(), p->col.as<int32_t>(), p->orig.as<int32_t>(),
                       color_of.as<int32_t>(), bad.as<int>());
    GD_HIP_TRY(hipGetLastError());
    rc = read_flag(bad, s, &h_bad);
    if (rc) return rc;
    GD_CHECK_ARG(!(h_bad & 4), "coloring is not proper: an off-diagonal entry joins two rows of one color");
  }
  // Per-color schedules: LDS row tiles and whole-row exact tiles (rows of
  // more than kSeqRow off-diagonal entries), in plan-row / edge coordinates.
  std::vector<Tile> all_tiles, all_hubs;
  p->tile_begin.assign(1, 0);
  p->hub_begin.assign(1, 0);
  std::vector<int32_t> reb;
  for (int c = 0; c < C; c++) {
    const int r0 = co[c] - co[0], r1 = co[c + 1] - co[0];
    const int nr = r1 - r0;
    const int32_t e0 = h_off[r0];
    if (nr > 0) {
      reb.resize((size_t)nr + 1);
      for (int i = 0; i <= nr; i++) reb[i] = h_off[r0 + i] - e0;
      Schedule sc;
      rc = build_schedule(nr, reb[nr], reb.data(), kBlock * p->ipt, kTileRows, &sc, kSeqRow);
      if (rc) return rc;
      p->max_row_len = std::max(p->max_row_len, (int)sc.max_row_len);
      const size_t h0 = all_hubs.size();
      for (const Tile& t : sc.tiles) {
        const Tile u{t.x + r0, t.y >= 0 ? t.y + r0 : t.y, t.z + e0, t.w + e0};
        (t.y < 0 ? all_hubs : all_tiles).push_back(u);
      }
      // the color lasts as long as its longest row: start the longest first
      std::stable_sort(all_hubs.begin() + h0, all_hubs.end(),
                       [](const Tile& x, const Tile& y) { return x.w - x.z > y.w - y.z; });
    }
    p->tile_begin.push_back((int)all_tiles.size());
    p->hub_begin.push_back((int)all_hubs.size());
  }
  rc = p->tiles.alloc(sizeof(Tile) * std::max<size_t>(all_tiles.size(), 1));
  if (!rc) rc = p->hubs.alloc(sizeof(Tile) * std::max<size_t>(all_hubs.size(), 1));
  if (rc) return rc;
  if (!all_tiles.empty())
    GD_HIP_TRY(hipMemcpyAsync(p->tiles.ptr, all_tiles.data(), sizeof(Tile) * all_tiles.size(),
                              hipMemcpyHostToDevice, s));
  if (!all_hubs.empty())
    GD_HIP_TRY(hipMemcpyAsync(p->hubs.ptr, all_hubs.data(), sizeof(Tile) * all_hubs.size(),
                              hipMemcpyHostToDevice, s));
  GD_HIP_TRY(hipStreamSynchronize(s));  // host vectors and scratch go out of scope
  return GD_OK;
}

// One color: one launch of symgs_color_kernel on s.
int sweep_color(gd_symgs_plan* p, int c, const SymgsPolicy::Args& a, hipStream_t s) {
  const int tb = p->tile_begin[c], nt = p->tile_begin[c + 1] - tb;
  const int hb = p->hub_begin[c], nh = p->hub_begin[c + 1] - hb;
  const IterArgs it{p->tiles.as<Tile>() + tb,
                    nt,
                    p->hubs.as<Tile>() + hb,
                    nh,
                    nt,
                    p->off.as<int32_t>(),
                    p->col.as<int32_t>(),
                    (int)std::max<int64_t>(p->nnz_p - 1, 0),
                    nullptr,
                    nullptr,
                    nullptr,
                    nullptr};
  if (nh > 0)
    hipLaunchKernelGGL(symgs_color_kernel<true>, dim3(nt + nh), dim3(kBlock), 0, s, a, it);
  else if (nt > 0)
    hipLaunchKernelGGL(symgs_color_kernel<false>, dim3(nt), dim3(kBlock), 0, s, a, it);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

}  // namespace

extern "C" {

int gd_symgs_plan_create(int m, int64_t nnz, const int32_t* d_Ap, const int32_t* d_Aj,
                         const float* d_Ax, const int32_t* d_indices,
                         const int32_t* color_offsets, int n_colors, void* stream,
                         gd_symgs_plan** out) {
  GD_CHECK_ARG(out != nullptr, "NULL plan output");
  *out = nullptr;
  GD_CHECK_ARG(m >= 0 && nnz >= 0 && nnz < (1ll << 31), "m / nnz out of range");
  GD_CHECK_ARG(n_colors >= 0 && color_offsets != nullptr, "NULL color offsets");
  GD_CHECK_ARG(d_Ap != nullptr && (nnz == 0 || (d_Aj && d_Ax)), "NULL matrix array");
  for (int c = 0; c < n_colors; c++)
    GD_CHECK_ARG(color_offsets[c] <= color_offsets[c + 1], "color offsets must be non-decreasing");
  GD_CHECK_ARG(color_offsets[0] >= 0 && color_offsets[n_colors] <= m,
               "color offsets must lie in [0, m]");
  GD_CHECK_ARG(d_indices != nullptr || color_offsets[n_colors] == color_offsets[0],
               "NULL indices");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  const auto t0 = std::chrono::steady_clock::now();
  gd_symgs_plan* p = new (std::nothrow) gd_symgs_plan;
  if (!p) return GD_ERR_ALLOC;
  const int rc = plan_build(p, m, nnz, d_Ap, d_Aj, d_Ax, d_indices, color_offsets, n_colors,
                            (hipStream_t)stream);
  if (rc) {
    delete p;
    return rc;
  }
  p->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  *out = p;
  return GD_OK;
}

int gd_symgs_plan_run(gd_symgs_plan* p, float* d_x, const float* d_b, int sweeps, void* stream) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  GD_CHECK_ARG(sweeps >= 0, "negative sweep count");
  GD_CHECK_ARG(p->n_p == 0 || (d_x && d_b), "NULL x / b");
  hipStream_t s = (hipStream_t)stream;
  const SymgsPolicy::Args a{d_x, d_b, p->val.as<float>(), p->diag.as<float>(),
                            p->orig.as<int32_t>()};
  for (int k = 0; k < sweeps; k++) {
    for (int c = 0; c < p->n_colors; c++) {
      const int rc = sweep_color(p, c, a, s);
      if (rc) return rc;
    }
    for (int c = p->n_colors - 1; c >= 0; c--) {
      const int rc = sweep_color(p, c, a, s);
      if (rc) return rc;
    }
  }
  return GD_OK;
}

int gd_symgs_plan_info(const gd_symgs_plan* p, gd_symgs_plan_info_t* info) {
  GD_CHECK_ARG(p != nullptr && info != nullptr, "NULL argument");
  memset(info, 0, sizeof(*info));
  info->rows = p->n_p;
  info->offdiag_nnz = p->nnz_p;
  info->n_colors = p->n_colors;
  info->n_tiles = p->tile_begin.back();
  info->n_exact_rows = p->hub_begin.back();
  info->max_row_len = p->max_row_len;
  info->plan_ms = p->plan_ms;
  return GD_OK;
}

int gd_symgs_plan_destroy(gd_symgs_plan* p) {
  delete p;
  return GD_OK;
}

int gd_symgs_solve(int m, int nnz, const int32_t* Ap, const int32_t* Aj, const int32_t* indices,
                   const float* Ax, float* x, const float* b, const int32_t* color_offsets,
                   int n_colors, float* kernel_ms) {
  GD_CHECK_ARG(m >= 0 && nnz >= 0, "negative size");
  GD_CHECK_ARG(Ap != nullptr && (m == 0 || (x && b && indices)), "NULL host array");
  GD_CHECK_ARG(nnz == 0 || (Aj && Ax), "NULL column indices / values");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  DeviceBuffer dAp, dAj, dAx, didx, dx, db;
  int rc = dAp.alloc(4 * ((size_t)m + 1));
  if (!rc) rc = dAj.alloc(4 * (size_t)nnz);
  if (!rc) rc = dAx.alloc(4 * (size_t)nnz);
  if (!rc) rc = didx.alloc(4 * (size_t)m);
  if (!rc) rc = dx.alloc(4 * (size_t)m);
  if (!rc) rc = db.alloc(4 * (size_t)m);
  if (rc) return rc;
  GD_HIP_TRY(hipMemcpy(dAp.ptr, Ap, 4 * ((size_t)m + 1), hipMemcpyHostToDevice));
  if (m > 0) {
    GD_HIP_TRY(hipMemcpy(didx.ptr, indices, 4 * (size_t)m, hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(dx.ptr, x, 4 * (size_t)m, hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(db.ptr, b, 4 * (size_t)m, hipMemcpyHostToDevice));
  }
  if (nnz > 0) {
    GD_HIP_TRY(hipMemcpy(dAj.ptr, Aj, 4 * (size_t)nnz, hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(dAx.ptr, Ax, 4 * (size_t)nnz, hipMemcpyHostToDevice));
  }
  gd_symgs_plan* plan = nullptr;
  rc = gd_symgs_plan_create(m, nnz, dAp.as<int32_t>(), dAj.as<int32_t>(), dAx.as<float>(),
                            didx.as<int32_t>(), color_offsets, n_colors, nullptr, &plan);
  if (rc) return rc;
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  GD_HIP_TRY(hipEventRecord(t0, nullptr));
  rc = gd_symgs_plan_run(plan, dx.as<float>(), db.as<float>(), 1, nullptr);
  GD_HIP_TRY(hipEventRecord(t1, nullptr));
  GD_HIP_TRY(hipEventSynchronize(t1));
  float ms = 0.0f;
  GD_HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  if (kernel_ms) *kernel_ms = ms;
  gd_symgs_plan_destroy(plan);
  if (rc) return rc;
  if (m > 0) GD_HIP_TRY(hipMemcpy(x, dx.ptr, 4 * (size_t)m, hipMemcpyDeviceToHost));
  return GD_OK;
}

}  // extern "C"
