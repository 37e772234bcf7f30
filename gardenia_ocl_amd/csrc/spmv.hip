// spmv.hip -- CSR SpMV y <- y + A x for gfx950 (src/spmv/omp_base.cc:17-28).
//
// Same tile engine as PageRank: products x[Aj[k]] * Ax[k] are formed during
// the gather (rounded, this file is built with -ffp-contract=off) and staged
// in LDS; rows are then summed by one thread starting from
// y[row] in CSR order, which is exactly the reference's arithmetic; so are
// the longer rows of a tile (the owner thread's chain just runs longer).
// Rows longer than a tile get a workgroup each that feeds one lane's
// sequential chain in CSR order (hub_chain_sum), so every row is the
// reference's sum bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "pull_engine.h"
#include "spmv_internal.h"
#include <rocprim/device/device_scan.hpp>

namespace gd {

constexpr int kSpIPT = 8;
constexpr int kSpTileNnz = kBlock * kSpIPT;

struct SpGather {
  const float* __restrict__ x;
  const float* __restrict__ Ax;
  __device__ __forceinline__ float value(int e, int j) const { return x[j] * Ax[e]; }
};

struct SpEpilogue {
  float* __restrict__ y;
  using Pre = float;  // y[row] before the update
  __device__ __forceinline__ Pre pre(int r) const { return y[r]; }
  __device__ __forceinline__ float init(const Pre& y0) const { return y0; }
  __device__ __forceinline__ double finish(int r, float s, const Pre&) const {
    y[r] = s;
    return 0.0;
  }
};

// Tile mode, one launch: workgroups [0, n_long) each add one row longer than
// a tile (2048 nonzeros) in CSR order on one lane (hub_chain_sum: products of
// both signs), the rest one LDS row tile each (every row summed by its owner
// thread in CSR order).  Every row is the reference's sum bit for bit.  LONG =
// false (no long rows) leaves the long-row code, and its registers, out.
template <bool LONG>
__global__ __launch_bounds__(kBlock) void spmv_pull_kernel(
    const Tile* __restrict__ tiles, const Tile* __restrict__ long_tiles, int n_long,
    const int32_t* __restrict__ Ap, const int32_t* __restrict__ Aj, int e_max,
    const float* __restrict__ Ax, const float* __restrict__ x, float* __restrict__ y) {
  // a row tile's LDS, or a long row's two chunk buffers + the broadcast sum
  static_assert(sizeof(RowTileLds<kSpIPT>) >= (2 * kHubChunk + 1) * sizeof(float), "LDS union");
  __shared__ __align__(16) unsigned char smem[sizeof(RowTileLds<kSpIPT>)];
  RowTileLds<kSpIPT>& lds = *reinterpret_cast<RowTileLds<kSpIPT>*>(smem);
  const SpGather g{x, Ax};
  const int b = blockIdx.x;
  if (LONG && b < n_long) {
    float* const buf = reinterpret_cast<float*>(smem);
    const Tile t = long_tiles[b];
    const float s = hub_chain_sum<16, false>(Aj, nullptr, e_max, t.z, t.w - t.z, g, y[t.x], buf,
                                            buf + kHubChunk, buf + 2 * kHubChunk);
    if (threadIdx.x == 0) y[t.x] = s;
  } else {
    const Tile t = tiles[b - n_long];
    const SpEpilogue e{y};
    (void)row_tile<kSpIPT, false, false, SpGather, SpEpilogue, true>(
        Ap, Aj, nullptr, e_max, t.x, t.y, t.z, t.w - t.z, g, e, lds);
  }
}

// ---------------------------------------------------------------- column slices
//
// For matrices whose x does not fit in one XCD's L2 (ER 2^22: 16.8 MB of x vs
// 4 MB) every gather of the tile engine misses L2.  The sliced mode splits the
// columns into kSlices ranges of x (2 MB each for ER 2^22) and keeps a
// slice-major copy of the matrix: slice 0's entries of every row (rows in
// order, each row's entries in column order), then slice 1's, ...  One SpMV
// runs the row engine once per slice, slice 0 first: pass q sums each row's
// slice-q products by the row's owner thread onto y[r] (LDS-staged tiles of
// that slice's entries; the x gathers of a pass stay in one 2 MB range, which
// every XCD's L2 holds).  For rows with sorted columns the passes add each
// row's products in column order -- the reference's loop
// (src/spmv/omp_base.cc:21-26) -- so y is its result bit for bit, with y
// itself carrying the running sum between passes.
constexpr int kSlices = 8;

// cnt[s * (rows+1) + r] = entries of row r in column slice s.
__global__ void slice_count_kernel(int rows, const int32_t* __restrict__ Ap,
                                   const int32_t* __restrict__ Aj, int width,
                                   int32_t* __restrict__ cnt, int* __restrict__ unsorted) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    int c[kSlices] = {0, 0, 0, 0, 0, 0, 0, 0};
    int prev = -1, bad = 0;
    for (int k = Ap[r]; k < Ap[r + 1]; k++) {
      const int j = Aj[k];
      bad |= j <= prev;
      prev = j;
      c[min(j / width, kSlices - 1)]++;
    }
#pragma unroll
    for (int q = 0; q < kSlices; q++) cnt[(int64_t)q * (rows + 1) + r] = c[q];
    if (bad) atomicOr(unsorted, 1);
  }
}

// Scatter each row's entries into the slice-major arrays (row order kept):
// the column indices unless values_only, and a 4-byte payload per entry (the
// values, or any other per-entry array).
template <class T>
__global__ void slice_scatter_kernel(int rows, const int32_t* __restrict__ Ap,
                                     const int32_t* __restrict__ Aj, const T* __restrict__ Ax,
                                     int width, const int32_t* __restrict__ soff,
                                     const int64_t* __restrict__ sbase, int32_t* __restrict__ Aj_s,
                                     T* __restrict__ Ax_s, int values_only) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    int64_t pos[kSlices];
#pragma unroll
    for (int q = 0; q < kSlices; q++) pos[q] = sbase[q] + soff[(int64_t)q * (rows + 1) + r];
    for (int k = Ap[r]; k < Ap[r + 1]; k++) {
      const int j = Aj[k];
      const int q = min(j / width, kSlices - 1);
      int64_t at = 0;
#pragma unroll
      for (int t = 0; t < kSlices; t++)
        if (t == q) at = pos[t]++;
      if (!values_only) Aj_s[at] = j;
      Ax_s[at] = Ax[k];
    }
  }
}

// Xs[i] = x[Aj_s[i]] for every slice-major entry i.  Workgroup b takes slice
// b % kSlices and its entries [2048 (b / kSlices), + 2048): workgroups are
// dispatched to the 8 XCDs round-robin, so slice s runs on XCD s, whose L2
// then holds that 1/8 of x; index loads and stores are coalesced, 8 gathers
// in flight per thread.
constexpr int kGatherPT = 8;

__global__ __launch_bounds__(256) void slice_gather_kernel(const int64_t* __restrict__ sbase,
                                                           const int32_t* __restrict__ Aj_s,
                                                           const float* __restrict__ x,
                                                           float* __restrict__ Xs) {
  const int q = blockIdx.x % kSlices;
  const int64_t b = sbase[q], e = sbase[q + 1];
  const int64_t c = b + (int64_t)(blockIdx.x / kSlices) * (256 * kGatherPT);
  if (c >= e) return;
  int j[kGatherPT];
  float v[kGatherPT];
#pragma unroll
  for (int t = 0; t < kGatherPT; t++) j[t] = Aj_s[min(c + t * 256 + threadIdx.x, e - 1)];
#pragma unroll
  for (int t = 0; t < kGatherPT; t++) v[t] = x[j[t]];
#pragma unroll
  for (int t = 0; t < kGatherPT; t++)
    if (c + t * 256 + threadIdx.x < e) Xs[c + t * 256 + threadIdx.x] = v[t];
}

}  // namespace gd

using namespace gd;

struct gd_spmv_plan {
  int rows = 0, cols = 0;
  int64_t nnz = 0;
  const int32_t* Ap = nullptr;
  const int32_t* Aj = nullptr;
  const float* Ax = nullptr;
  hipStream_t stream = nullptr;
  Schedule sched;
  int n_tiles = 0, n_hubs = 0;  // row tiles; rows longer than a tile
  DeviceBuffer tiles, long_tiles;
  DeviceBuffer dummy_col;  // 1 element, used as Aj when nnz == 0
  int e_max = 0;
  // column-sliced mode: per slice, its long rows then its row tiles in
  // slice_tiles[slice_first[q] ...]
  bool sliced = false;
  int width = 0;
  DeviceBuffer soff, sbase, Aj_s, Ax_s, slice_tiles;
  std::vector<int64_t> slice_base;
  std::vector<int> slice_first, slice_long, slice_rows;
};

namespace {

int build_sliced(gd_spmv_plan* p) {
  const int rows = p->rows;
  p->width = std::max(1, (p->cols + kSlices - 1) / kSlices);
  int rc = p->soff.alloc(4 * (size_t)kSlices * (rows + 1));
  if (!rc) rc = p->sbase.alloc(8 * (kSlices + 1));
  if (!rc) rc = p->Aj_s.alloc(4 * (size_t)p->nnz);
  if (!rc) rc = p->Ax_s.alloc(4 * (size_t)p->nnz);
  if (rc) return rc;
  DeviceBuffer cnt, bad, tmp;
  rc = cnt.alloc(4 * (size_t)kSlices * (rows + 1));
  if (!rc) rc = bad.alloc(4);
  if (rc) return rc;
  GD_HIP_TRY(hipMemsetAsync(bad.ptr, 0, 4, p->stream));
  GD_HIP_TRY(hipMemsetAsync(cnt.ptr, 0, 4 * (size_t)kSlices * (rows + 1), p->stream));
  const int g = std::max(1, std::min((rows + 255) / 256, 16384));
  hipLaunchKernelGGL(slice_count_kernel, dim3(g), dim3(256), 0, p->stream, rows, p->Ap, p->Aj,
                     p->width, cnt.as<int32_t>(), bad.as<int>());
  GD_HIP_TRY(hipGetLastError());
  int h_bad = 0;
  GD_HIP_TRY(hipMemcpyAsync(&h_bad, bad.ptr, 4, hipMemcpyDeviceToHost, p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  if (h_bad) {
    set_error("column-sliced SpMV needs sorted columns within each row");
    return GD_ERR_INVALID;
  }
  size_t bytes = 0;
  GD_HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, cnt.as<int32_t>(), p->soff.as<int32_t>(), 0,
                                     (size_t)rows + 1, rocprim::plus<int32_t>(), p->stream));
  rc = tmp.alloc(bytes);
  if (rc) return rc;
  for (int q = 0; q < kSlices; q++)  // each slice's row offsets from 0
    GD_HIP_TRY(rocprim::exclusive_scan(tmp.ptr, bytes, cnt.as<int32_t>() + (size_t)q * (rows + 1),
                                       p->soff.as<int32_t>() + (size_t)q * (rows + 1), 0,
                                       (size_t)rows + 1, rocprim::plus<int32_t>(), p->stream));
  std::vector<int32_t> tot(kSlices);
  for (int q = 0; q < kSlices; q++)
    GD_HIP_TRY(hipMemcpyAsync(&tot[q], p->soff.as<int32_t>() + (size_t)q * (rows + 1) + rows, 4,
                              hipMemcpyDeviceToHost, p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  std::vector<int64_t> base(kSlices + 1, 0);
  for (int q = 0; q < kSlices; q++) base[q + 1] = base[q] + tot[q];
  if (base[kSlices] != p->nnz) {
    set_error("column-sliced SpMV: slice counts do not add up to nnz");
    return GD_ERR_INVALID;
  }
  GD_HIP_TRY(hipMemcpyAsync(p->sbase.ptr, base.data(), 8 * (kSlices + 1), hipMemcpyHostToDevice,
                            p->stream));
  hipLaunchKernelGGL(slice_scatter_kernel<float>, dim3(g), dim3(256), 0, p->stream, rows, p->Ap,
                     p->Aj, p->Ax, p->width, p->soff.as<int32_t>(), p->sbase.as<int64_t>(),
                     p->Aj_s.as<int32_t>(), p->Ax_s.as<float>(), 0);
  GD_HIP_TRY(hipGetLastError());
  // one row-tile schedule per slice, over that slice's row offsets
  std::vector<int32_t> h_soff((size_t)kSlices * (rows + 1));
  GD_HIP_TRY(hipMemcpyAsync(h_soff.data(), p->soff.ptr, 4 * h_soff.size(), hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  std::vector<Tile> all;
  p->slice_base.assign(base.begin(), base.end());
  p->slice_first.assign(kSlices, 0);
  p->slice_long.assign(kSlices, 0);
  p->slice_rows.assign(kSlices, 0);
  for (int q = 0; q < kSlices; q++) {
    p->slice_first[q] = (int)all.size();
    if (tot[q] == 0) continue;  // no entries: the pass would leave y unchanged
    Schedule sc;
    rc = build_schedule(rows, tot[q], h_soff.data() + (size_t)q * (rows + 1), kSpTileNnz,
                        kTileRows, &sc, kSpTileNnz);
    if (rc) return rc;
    std::vector<Tile> row_tiles, longs;
    for (const Tile& t : sc.tiles) (t.y < 0 ? longs : row_tiles).push_back(t);
    std::stable_sort(longs.begin(), longs.end(),  // the longest rows start first
                     [](const Tile& a, const Tile& b) { return a.w - a.z > b.w - b.z; });
    p->slice_long[q] = (int)longs.size();
    p->slice_rows[q] = (int)row_tiles.size();
    all.insert(all.end(), longs.begin(), longs.end());
    all.insert(all.end(), row_tiles.begin(), row_tiles.end());
  }
  rc = p->slice_tiles.alloc(sizeof(Tile) * std::max<size_t>(all.size(), 1));
  if (rc) return rc;
  if (!all.empty())
    GD_HIP_TRY(hipMemcpyAsync(p->slice_tiles.ptr, all.data(), sizeof(Tile) * all.size(),
                              hipMemcpyHostToDevice, p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  return GD_OK;
}

}  // namespace

extern "C" {

int gd_spmv_plan_create(int rows, int cols, int64_t nnz, const int32_t* d_Ap, const int32_t* d_Aj,
                        const float* d_Ax, void* stream, gd_spmv_plan** out) {
  return gd_spmv_plan_create_ex(rows, cols, nnz, d_Ap, d_Aj, d_Ax, stream, GD_SPMV_TILES, out);
}

int gd_spmv_plan_create_ex(int rows, int cols, int64_t nnz, const int32_t* d_Ap,
                           const int32_t* d_Aj, const float* d_Ax, void* stream, int mode,
                           gd_spmv_plan** out) {
  GD_CHECK_ARG(mode == GD_SPMV_AUTO || mode == GD_SPMV_TILES || mode == GD_SPMV_SLICED,
               "unknown SpMV mode");
  GD_CHECK_ARG(out != nullptr, "plan output pointer is NULL");
  *out = nullptr;
  GD_CHECK_ARG(rows >= 0 && cols >= 0 && nnz >= 0 && nnz < (1ll << 31), "size out of range");
  GD_CHECK_ARG(rows == 0 || d_Ap, "NULL row offsets");
  GD_CHECK_ARG(nnz == 0 || (d_Aj && d_Ax), "NULL column indices / values");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  gd_spmv_plan* p = new gd_spmv_plan();
  p->rows = rows;
  p->cols = cols;
  p->nnz = nnz;
  p->Ap = d_Ap;
  p->Aj = d_Aj;
  p->Ax = d_Ax;
  p->stream = (hipStream_t)stream;
  int rc = GD_OK;
  if (rows > 0) {
    std::vector<int32_t> h_off(rows + 1);
    hipError_t e = hipMemcpyAsync(h_off.data(), d_Ap, sizeof(int32_t) * (rows + 1),
                                  hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
      set_error(std::string("copying row offsets: ") + hipGetErrorString(e));
      rc = GD_ERR_DEVICE;
    }
    if (!rc)
      rc = build_schedule(rows, nnz, h_off.data(), kSpTileNnz, kTileRows, &p->sched, kSpTileNnz);
    if (!rc) rc = check_columns(d_Aj, nnz, cols, p->stream);
    if (!rc && nnz > 0 && mode != GD_SPMV_TILES) {
      // auto: slice when x spans more than one XCD's L2 and rows are short
      const bool want = mode == GD_SPMV_SLICED ||
                        ((int64_t)cols * 4 > (4 << 20) && p->sched.max_row_len <= kSeqRow);
      if (want) {
        rc = build_sliced(p);
        if (!rc) p->sliced = true;
        if (rc && mode == GD_SPMV_AUTO) rc = GD_OK;  // unsorted rows: tile engine
      }
    }
    if (!rc && nnz == 0) {
      rc = p->dummy_col.alloc(4);
      if (!rc && hipMemsetAsync(p->dummy_col.ptr, 0, 4, p->stream) != hipSuccess) rc = GD_ERR_DEVICE;
      p->Aj = p->dummy_col.as<int32_t>();
    }
    p->e_max = (int)std::max<int64_t>(nnz - 1, 0);
    if (!rc) {
      std::vector<Tile> row_tiles, longs;
      for (const Tile& t : p->sched.tiles) (t.y < 0 ? longs : row_tiles).push_back(t);
      std::stable_sort(longs.begin(), longs.end(),  // the longest rows start first
                       [](const Tile& a, const Tile& b) { return a.w - a.z > b.w - b.z; });
      p->n_tiles = (int)row_tiles.size();
      p->n_hubs = (int)longs.size();
      rc = p->tiles.alloc(sizeof(Tile) * std::max(1, p->n_tiles));
      if (!rc) rc = p->long_tiles.alloc(sizeof(Tile) * std::max(1, p->n_hubs));
      hipError_t e2 = hipSuccess;
      if (!rc && p->n_tiles)
        e2 = hipMemcpyAsync(p->tiles.ptr, row_tiles.data(), sizeof(Tile) * p->n_tiles,
                            hipMemcpyHostToDevice, p->stream);
      if (!rc && e2 == hipSuccess && p->n_hubs)
        e2 = hipMemcpyAsync(p->long_tiles.ptr, longs.data(), sizeof(Tile) * p->n_hubs,
                            hipMemcpyHostToDevice, p->stream);
      if (!rc && e2 == hipSuccess) e2 = hipStreamSynchronize(p->stream);
      if (!rc && e2 != hipSuccess) {
        set_error(std::string("uploading schedule: ") + hipGetErrorString(e2));
        rc = GD_ERR_DEVICE;
      }
    }
  }
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return GD_OK;
}

int gd_spmv_plan_run(gd_spmv_plan* p, const float* d_x, float* d_y) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  if (p->rows == 0) return GD_OK;
  GD_CHECK_ARG(d_y != nullptr && (p->nnz == 0 || d_x != nullptr), "NULL x / y");
  if (p->sliced) {
    const Tile* tl = p->slice_tiles.as<Tile>();
    for (int q = 0; q < kSlices; q++) {  // slice 0 first: each row's column order
      const int nl = p->slice_long[q], nr = p->slice_rows[q];
      if (nl + nr == 0) continue;
      const Tile* lt = tl + p->slice_first[q];
      const int32_t* off = p->soff.as<int32_t>() + (size_t)q * (p->rows + 1);
      const int32_t* aj = p->Aj_s.as<int32_t>() + p->slice_base[q];
      const float* ax = p->Ax_s.as<float>() + p->slice_base[q];
      const int em = (int)std::max<int64_t>(p->slice_base[q + 1] - p->slice_base[q] - 1, 0);
      if (nl)
        hipLaunchKernelGGL(spmv_pull_kernel<true>, dim3(nl + nr), dim3(kBlock), 0, p->stream,
                           lt + nl, lt, nl, off, aj, em, ax, d_x, d_y);
      else
        hipLaunchKernelGGL(spmv_pull_kernel<false>, dim3(nr), dim3(kBlock), 0, p->stream, lt, lt,
                           0, off, aj, em, ax, d_x, d_y);
    }
    GD_HIP_TRY(hipGetLastError());
    return GD_OK;
  }
  if (p->n_hubs)
    hipLaunchKernelGGL(spmv_pull_kernel<true>, dim3(p->n_hubs + p->n_tiles), dim3(kBlock), 0,
                       p->stream, p->tiles.as<Tile>(), p->long_tiles.as<Tile>(), p->n_hubs, p->Ap,
                       p->Aj, p->e_max, p->Ax, d_x, d_y);
  else if (p->n_tiles)
    hipLaunchKernelGGL(spmv_pull_kernel<false>, dim3(p->n_tiles), dim3(kBlock), 0, p->stream,
                       p->tiles.as<Tile>(), p->long_tiles.as<Tile>(), 0, p->Ap, p->Aj, p->e_max,
                       p->Ax, d_x, d_y);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

int gd_spmv_plan_refresh_values(gd_spmv_plan* p) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  if (!p->sliced) return GD_OK;  // the tile engine reads the caller's values directly
  const int g = std::max(1, std::min((p->rows + 255) / 256, 16384));
  hipLaunchKernelGGL(slice_scatter_kernel<float>, dim3(g), dim3(256), 0, p->stream, p->rows, p->Ap,
                     p->Aj, p->Ax, p->width, p->soff.as<int32_t>(), p->sbase.as<int64_t>(),
                     p->Aj_s.as<int32_t>(), p->Ax_s.as<float>(), 1);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

}  // extern "C"

float* gd::spmv_sliced_values(gd_spmv_plan* p) {
  return (p && p->sliced) ? p->Ax_s.as<float>() : nullptr;
}

int gd::spmv_slice_order(gd_spmv_plan* p, const int32_t* d_in, int32_t* d_out) {
  GD_CHECK_ARG(p != nullptr && p->sliced, "not a sliced SpMV plan");
  if (p->nnz == 0) return GD_OK;
  GD_CHECK_ARG(d_in && d_out, "NULL payload");
  const int g = std::max(1, std::min((p->rows + 255) / 256, 16384));
  hipLaunchKernelGGL(slice_scatter_kernel<int32_t>, dim3(g), dim3(256), 0, p->stream, p->rows,
                     p->Ap, p->Aj, d_in, p->width, p->soff.as<int32_t>(), p->sbase.as<int64_t>(),
                     nullptr, d_out, 1);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

int gd::spmv_sliced_gather(gd_spmv_plan* p, const float* d_x, float* d_Xs) {
  GD_CHECK_ARG(p != nullptr && p->sliced, "not a sliced SpMV plan");
  if (p->nnz == 0) return GD_OK;
  GD_CHECK_ARG(d_x && d_Xs, "NULL array");
  int64_t most = 0;
  for (int q = 0; q < kSlices; q++)
    most = std::max<int64_t>(most, p->slice_base[q + 1] - p->slice_base[q]);
  const int64_t chunks = (most + 256 * kGatherPT - 1) / (256 * kGatherPT);
  hipLaunchKernelGGL(slice_gather_kernel, dim3((unsigned)(kSlices * chunks)), dim3(256), 0,
                     p->stream, p->sbase.as<int64_t>(), p->Aj_s.as<int32_t>(), d_x, d_Xs);
  GD_HIP_TRY(hipGetLastError());
  return GD_OK;
}

extern "C" {

int gd_spmv_plan_mode(const gd_spmv_plan* p) {
  return p == nullptr ? -1 : (p->sliced ? GD_SPMV_SLICED : GD_SPMV_TILES);
}

int gd_spmv_plan_info(const gd_spmv_plan* p, int64_t* n_tiles, int64_t* n_hubs,
                      int32_t* max_row_len) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  if (n_tiles) *n_tiles = p->n_tiles;
  if (n_hubs) *n_hubs = p->n_hubs;
  if (max_row_len) *max_row_len = p->sched.max_row_len;
  return GD_OK;
}

int gd_spmv_plan_destroy(gd_spmv_plan* p) {
  delete p;
  return GD_OK;
}

int gd_spmv_solve(int m, int nnz, const int32_t* Ap, const int32_t* Aj, const float* Ax,
                  const float* x, float* y, float* kernel_ms) {
  GD_CHECK_ARG(m >= 0 && nnz >= 0, "negative size");
  GD_CHECK_ARG(m == 0 || (Ap && x && y), "NULL host array");
  GD_CHECK_ARG(nnz == 0 || (Aj && Ax), "NULL column indices / values");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  DeviceBuffer dAp, dAj, dAx, dx, dy;
  int rc = dAp.alloc(4 * ((size_t)m + 1));
  if (!rc) rc = dAj.alloc(4 * (size_t)nnz);
  if (!rc) rc = dAx.alloc(4 * (size_t)nnz);
  if (!rc) rc = dx.alloc(4 * (size_t)m);
  if (!rc) rc = dy.alloc(4 * (size_t)m);
  if (rc) return rc;
  if (m > 0) {
    GD_HIP_TRY(hipMemcpy(dAp.ptr, Ap, 4 * ((size_t)m + 1), hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(dx.ptr, x, 4 * (size_t)m, hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(dy.ptr, y, 4 * (size_t)m, hipMemcpyHostToDevice));
  }
  if (nnz > 0) {
    GD_HIP_TRY(hipMemcpy(dAj.ptr, Aj, 4 * (size_t)nnz, hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(dAx.ptr, Ax, 4 * (size_t)nnz, hipMemcpyHostToDevice));
  }
  gd_spmv_plan* plan = nullptr;
  rc = gd_spmv_plan_create_ex(m, m, nnz, dAp.as<int32_t>(), dAj.as<int32_t>(), dAx.as<float>(),
                              nullptr, GD_SPMV_AUTO, &plan);
  if (rc) return rc;
  std::unique_ptr<gd_spmv_plan> owner(plan);  // destroyed on every return below
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  GD_HIP_TRY(hipEventRecord(t0, nullptr));
  rc = gd_spmv_plan_run(plan, dx.as<float>(), dy.as<float>());
  GD_HIP_TRY(hipEventRecord(t1, nullptr));
  GD_HIP_TRY(hipEventSynchronize(t1));
  float ms = 0.0f;
  GD_HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  if (kernel_ms) *kernel_ms = ms;
  if (rc) return rc;
  if (m > 0) GD_HIP_TRY(hipMemcpy(y, dy.ptr, 4 * (size_t)m, hipMemcpyDeviceToHost));
  return GD_OK;
}

}  // extern "C"
