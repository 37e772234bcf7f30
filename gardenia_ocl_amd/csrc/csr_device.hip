// csr_device.hip -- CSR construction from an edge list on the device, with the
// semantics of the reference's fill_data() (include/graph_io.h:21-130) and
// the same flags as gd_csr_from_edges (the host builder):
//
//   1. placement (graph_io.h:245-257): every edge pushes one record to the
//      row of its source (its destination with transpose), preceded by a
//      mirrored record when symmetrising; once a mirrored record has been
//      pushed the transpose flag stays cleared for the rest of the list (the
//      reference's loop resets it), so placement is a prefix sum over the
//      per-edge record counts plus the index of the first mirrored edge;
//   2. a stable radix sort of the records (64-bit keys carrying row and
//      column, the weight as the value) by row (by (row, column) when
//      `sorted`), so each row keeps its push order.  std::sort in the
//      reference is not stable: with `sorted` and duplicate edges of different
//      weights, the weight kept by step 4 may differ; the structure, and
//      unweighted graphs, are identical;
//   3. self-loops dropped (graph_io.h:32-45);
//   4. ADJACENT duplicates dropped, comparing with the previous kept record of
//      the row (graph_io.h:48-61);
//   5. row offsets from the row boundaries of the sorted records, degree =
//      row length (graph_io.h:72-78, :289-293).
// Each step is a rocPRIM primitive or a one-pass kernel; the edge list never
// leaves HBM.

#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>

#include "gardenia_hip.h"
#include "gd_common.h"

namespace gd {
namespace {

constexpr int kB = 256;

int grid_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kB - 1) / kB, 65536));
}

#define GRID_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kB)

__global__ void check_edges_kernel(int64_t n, const int32_t* __restrict__ src,
                                   const int32_t* __restrict__ dst, int m, int* __restrict__ bad) {
  int f = 0;
  GRID_LOOP(i, n) f |= (src[i] < 0) | (src[i] >= m) | (dst[i] < 0) | (dst[i] >= m);
  if (__any(f) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

// first index with a mirrored record (symmetrise && (mirror || s != d))
__global__ void first_mirror_kernel(int64_t n, const int32_t* __restrict__ src,
                                    const int32_t* __restrict__ dst, int mirror,
                                    unsigned long long* __restrict__ first) {
  GRID_LOOP(i, n) if (mirror || src[i] != dst[i]) atomicMin(first, (unsigned long long)i);
}

__global__ void count_kernel(int64_t n, const int32_t* __restrict__ src,
                             const int32_t* __restrict__ dst, int sym, int mirror,
                             int64_t* __restrict__ cnt) {
  GRID_LOOP(i, n) cnt[i] = 1 + ((sym && (mirror || src[i] != dst[i])) ? 1 : 0);
}

// Records in push order, as one 64-bit sort key each plus the weight.  The
// key carries the column with the row: (col << 32 | row) sorted on the row
// bits only, or (row << 32 | col) sorted on both when `sorted`; the radix sort
// moves the column along with the row, so no permutation gather is needed.
__device__ __forceinline__ uint64_t make_key(int32_t r, int32_t c, int sorted) {
  return sorted ? (((uint64_t)(uint32_t)r << 32) | (uint32_t)c)
                : (((uint64_t)(uint32_t)c << 32) | (uint32_t)r);
}
__device__ __forceinline__ int32_t key_row(uint64_t k, int sorted) {
  return (int32_t)(sorted ? (k >> 32) : (k & 0xffffffffu));
}
__device__ __forceinline__ int32_t key_col(uint64_t k, int sorted) {
  return (int32_t)(sorted ? (k & 0xffffffffu) : (k >> 32));
}

__global__ void place_kernel(int64_t n, const int32_t* __restrict__ src,
                             const int32_t* __restrict__ dst, const float* __restrict__ wt,
                             const int64_t* __restrict__ at, int sym, int mirror, int transpose,
                             unsigned long long first_mirror, int sorted,
                             uint64_t* __restrict__ key, float* __restrict__ rw) {
  GRID_LOOP(i, n) {
    const int32_t s = src[i], d = dst[i];
    const float w = wt ? wt[i] : 1.0f;
    int64_t p = at[i];
    if (sym && (mirror || s != d)) {  // the mirrored record is pushed first
      key[p] = make_key(d, s, sorted);
      rw[p] = w;
      p++;
    }
    const bool tr = transpose && (unsigned long long)i < first_mirror;
    key[p] = tr ? make_key(d, s, sorted) : make_key(s, d, sorted);
    rw[p] = w;
  }
}

// keep[i]: not a self-loop (remove_selfloops)
__global__ void selfloop_flags_kernel(int64_t n, const uint64_t* __restrict__ key, int sorted,
                                      int32_t* __restrict__ keep) {
  GRID_LOOP(i, n) keep[i] = key_row(key[i], sorted) != key_col(key[i], sorted);
}

// keep[i]: not equal to the previous record of its row (same row and column:
// equal keys)
__global__ void dup_flags_kernel(int64_t n, const uint64_t* __restrict__ key,
                                 int32_t* __restrict__ keep) {
  GRID_LOOP(i, n) keep[i] = !(i > 0 && key[i] == key[i - 1]);
}

__global__ void scatter_kept_kernel(int64_t n, const int32_t* __restrict__ keep,
                                    const int32_t* __restrict__ pos,
                                    const uint64_t* __restrict__ key, const float* __restrict__ w,
                                    uint64_t* __restrict__ key2, float* __restrict__ w2) {
  GRID_LOOP(i, n) {
    if (keep[i]) {
      key2[pos[i]] = key[i];
      w2[pos[i]] = w[i];
    }
  }
}

// Row offsets from records sorted by row: record i sets off[q] = i for every
// row q in (row of record i-1, row of record i]; the last record also closes
// rows after it (off[q] = n up to q = m).  Every entry is written once.
__global__ void offsets_kernel(int64_t n, int m, const uint64_t* __restrict__ key, int sorted,
                               int32_t* __restrict__ off, int32_t* __restrict__ col) {
  GRID_LOOP(i, n) {
    const uint64_t k = key[i];
    const int32_t r = key_row(k, sorted);
    col[i] = key_col(k, sorted);
    const int32_t p = (i == 0) ? -1 : key_row(key[i - 1], sorted);
    for (int32_t q = p + 1; q <= r; q++) off[q] = (int32_t)i;
    if (i == n - 1)
      for (int32_t q = r + 1; q <= m; q++) off[q] = (int32_t)n;
  }
}

__global__ void fill_kernel(int64_t n, int32_t v, int32_t* __restrict__ a) { GRID_LOOP(i, n) a[i] = v; }

__global__ void degree_kernel(int m, const int32_t* __restrict__ off, int32_t* __restrict__ deg) {
  GRID_LOOP(i, m) deg[i] = off[i + 1] - off[i];
}

int sync_read(const void* dptr, void* h, size_t bytes, hipStream_t s) {
  GD_HIP_TRY(hipMemcpyAsync(h, dptr, bytes, hipMemcpyDeviceToHost, s));
  GD_HIP_TRY(hipStreamSynchronize(s));
  return GD_OK;
}

// Stable compaction of (key, w) by 0/1 flags: exclusive scan + scatter.
// `keep` is filled by the caller; key / w are replaced by the kept records.
int compact(int64_t n, DeviceBuffer& key, DeviceBuffer& w, const DeviceBuffer& keep,
            hipStream_t s, int64_t* kept) {
  const size_t cap = (size_t)std::max<int64_t>(n, 1);
  DeviceBuffer pos, key2, w2, tmp;
  int rc = pos.alloc(4 * cap);
  if (!rc) rc = key2.alloc(8 * cap);
  if (!rc) rc = w2.alloc(4 * cap);
  if (rc) return rc;
  size_t tb = 0;
  GD_HIP_TRY(rocprim::exclusive_scan(nullptr, tb, keep.as<int32_t>(), pos.as<int32_t>(), 0,
                                     (size_t)n, rocprim::plus<int32_t>(), s));
  rc = tmp.alloc(std::max<size_t>(tb, 1));
  if (rc) return rc;
  GD_HIP_TRY(rocprim::exclusive_scan(tmp.ptr, tb, keep.as<int32_t>(), pos.as<int32_t>(), 0,
                                     (size_t)n, rocprim::plus<int32_t>(), s));
  hipLaunchKernelGGL(scatter_kept_kernel, dim3(grid_for(n)), dim3(kB), 0, s, n,
                     keep.as<int32_t>(), pos.as<int32_t>(), key.as<uint64_t>(), w.as<float>(),
                     key2.as<uint64_t>(), w2.as<float>());
  GD_HIP_TRY(hipGetLastError());
  int32_t last[2] = {0, 0};
  rc = sync_read(pos.as<int32_t>() + (n - 1), &last[0], 4, s);
  if (!rc) rc = sync_read(keep.as<int32_t>() + (n - 1), &last[1], 4, s);
  if (rc) return rc;
  std::swap(key.ptr, key2.ptr);
  std::swap(key.bytes, key2.bytes);
  std::swap(w.ptr, w2.ptr);
  std::swap(w.bytes, w2.bytes);
  *kept = (int64_t)last[0] + last[1];
  return GD_OK;
}

int build(int m, int64_t n, const int32_t* d_src, const int32_t* d_dst, const float* d_wt,
          int sym, int transpose, int sorted, int rm_self, int rm_dup, int mirror, hipStream_t s,
          gd_device_csr* out) {
  if (n > 0) {
    DeviceBuffer bad;
    int rc = bad.alloc(4);
    if (rc) return rc;
    GD_HIP_TRY(hipMemsetAsync(bad.ptr, 0, 4, s));
    hipLaunchKernelGGL(check_edges_kernel, dim3(grid_for(n)), dim3(kB), 0, s, n, d_src, d_dst, m,
                       bad.as<int>());
    int h = 0;
    rc = sync_read(bad.ptr, &h, 4, s);
    if (rc) return rc;
    GD_CHECK_ARG(h == 0, "edge endpoint out of range [0, m)");
  }
  // 1. placement
  unsigned long long first_mirror = (unsigned long long)n;
  int64_t cap = 0;
  DeviceBuffer at;
  int rc = at.alloc(8 * (size_t)std::max<int64_t>(n, 1));
  if (rc) return rc;
  if (n > 0) {
    DeviceBuffer cnt, first, tmp;
    rc = cnt.alloc(8 * (size_t)n);
    if (!rc) rc = first.alloc(8);
    if (rc) return rc;
    if (sym) {
      GD_HIP_TRY(hipMemcpyAsync(first.ptr, &first_mirror, 8, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(first_mirror_kernel, dim3(grid_for(n)), dim3(kB), 0, s, n, d_src, d_dst,
                         mirror, first.as<unsigned long long>());
      rc = sync_read(first.ptr, &first_mirror, 8, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(count_kernel, dim3(grid_for(n)), dim3(kB), 0, s, n, d_src, d_dst, sym,
                       mirror, cnt.as<int64_t>());
    size_t tb = 0;
    GD_HIP_TRY(rocprim::exclusive_scan(nullptr, tb, cnt.as<int64_t>(), at.as<int64_t>(),
                                       (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
    rc = tmp.alloc(std::max<size_t>(tb, 1));
    if (rc) return rc;
    GD_HIP_TRY(rocprim::exclusive_scan(tmp.ptr, tb, cnt.as<int64_t>(), at.as<int64_t>(),
                                       (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), s));
    int64_t last[2] = {0, 0};
    rc = sync_read(at.as<int64_t>() + (n - 1), &last[0], 8, s);
    if (!rc) rc = sync_read(cnt.as<int64_t>() + (n - 1), &last[1], 8, s);
    if (rc) return rc;
    cap = last[0] + last[1];
  }
  GD_CHECK_ARG(cap < (1ll << 31), "more than 2^31 - 1 edges after symmetrisation");
  const size_t capb = (size_t)std::max<int64_t>(cap, 1);
  DeviceBuffer key, w;  // sorted records
  rc = key.alloc(8 * capb);
  if (!rc) rc = w.alloc(4 * capb);
  if (rc) return rc;
  if (cap > 0) {
    DeviceBuffer key0, w0, tmp;
    rc = key0.alloc(8 * capb);
    if (!rc) rc = w0.alloc(4 * capb);
    if (rc) return rc;
    hipLaunchKernelGGL(place_kernel, dim3(grid_for(n)), dim3(kB), 0, s, n, d_src, d_dst, d_wt,
                       at.as<int64_t>(), sym, mirror, transpose, first_mirror, sorted,
                       key0.as<uint64_t>(), w0.as<float>());
    GD_HIP_TRY(hipGetLastError());
    // 2. stable radix sort by row (by row, then column, when sorted): only the
    // bits in use
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < m) bits++;
    const int end_bit = sorted ? 32 + bits : bits;
    size_t tb = 0;
    GD_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tb, key0.as<uint64_t>(), key.as<uint64_t>(),
                                         w0.as<float>(), w.as<float>(), (size_t)cap, 0, end_bit,
                                         s));
    rc = tmp.alloc(std::max<size_t>(tb, 1));
    if (rc) return rc;
    GD_HIP_TRY(rocprim::radix_sort_pairs(tmp.ptr, tb, key0.as<uint64_t>(), key.as<uint64_t>(),
                                         w0.as<float>(), w.as<float>(), (size_t)cap, 0, end_bit,
                                         s));
    GD_HIP_TRY(hipStreamSynchronize(s));  // scratch is freed on scope exit
  }
  at.release();
  // 3. self-loops, 4. adjacent duplicates (against the previous kept record:
  // equal keys once the self-loops are gone)
  int64_t nk = cap;
  if (nk > 0 && (rm_self || rm_dup)) {
    DeviceBuffer keep;
    rc = keep.alloc(4 * capb);
    if (rc) return rc;
    if (rm_self) {
      hipLaunchKernelGGL(selfloop_flags_kernel, dim3(grid_for(nk)), dim3(kB), 0, s, nk,
                         key.as<uint64_t>(), sorted, keep.as<int32_t>());
      rc = compact(nk, key, w, keep, s, &nk);
      if (rc) return rc;
    }
    if (rm_dup && nk > 0) {
      hipLaunchKernelGGL(dup_flags_kernel, dim3(grid_for(nk)), dim3(kB), 0, s, nk,
                         key.as<uint64_t>(), keep.as<int32_t>());
      rc = compact(nk, key, w, keep, s, &nk);
      if (rc) return rc;
    }
  }
  // 5. column array, offsets and degrees
  DeviceBuffer off, col, deg, wout;
  rc = off.alloc(4 * ((size_t)m + 1));
  if (!rc) rc = col.alloc(4 * (size_t)std::max<int64_t>(nk, 1));
  if (!rc) rc = deg.alloc(4 * (size_t)std::max(m, 1));
  if (rc) return rc;
  if (nk > 0)
    hipLaunchKernelGGL(offsets_kernel, dim3(grid_for(nk)), dim3(kB), 0, s, nk, m,
                       key.as<uint64_t>(), sorted, off.as<int32_t>(), col.as<int32_t>());
  else
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for((int64_t)m + 1)), dim3(kB), 0, s,
                       (int64_t)m + 1, 0, off.as<int32_t>());
  if (m > 0)
    hipLaunchKernelGGL(degree_kernel, dim3(grid_for(m)), dim3(kB), 0, s, m, off.as<int32_t>(),
                       deg.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipStreamSynchronize(s));
  // hand the buffers to the caller (w holds the kept weights, in order)
  out->m = m;
  out->nnz = nk;
  out->row_offsets = off.as<int32_t>();
  out->column_indices = col.as<int32_t>();
  out->weight = w.as<float>();
  out->degree = deg.as<int32_t>();
  off.ptr = col.ptr = w.ptr = deg.ptr = nullptr;  // ownership moves to *out
  return GD_OK;
}

}  // namespace
}  // namespace gd

using namespace gd;

extern "C" {

int gd_csr_from_edges_device(int m, int64_t n, const int32_t* d_src, const int32_t* d_dst,
                             const float* d_wt, int symmetrize, int transpose, int sorted,
                             int remove_selfloops, int remove_redundents, int mirror_selfloops,
                             void* stream, gd_device_csr* out) {
  GD_CHECK_ARG(out != nullptr, "NULL output");
  memset(out, 0, sizeof(*out));
  GD_CHECK_ARG(m >= 0 && n >= 0 && n < (1ll << 31), "m / n out of range");
  GD_CHECK_ARG(n == 0 || (d_src && d_dst), "NULL edge arrays");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  gd_device_csr tmp;
  memset(&tmp, 0, sizeof(tmp));
  const int rc = build(m, n, d_src, d_dst, d_wt, symmetrize != 0, transpose != 0, sorted != 0,
                       remove_selfloops != 0, remove_redundents != 0, mirror_selfloops != 0,
                       (hipStream_t)stream, &tmp);
  if (rc) return rc;
  *out = tmp;
  return GD_OK;
}

int gd_device_csr_copy(const gd_device_csr* c, int32_t* d_off, int32_t* d_col, float* d_w,
                       int32_t* d_deg, void* stream) {
  GD_CHECK_ARG(c != nullptr, "NULL csr");
  hipStream_t s = (hipStream_t)stream;
  const size_t m = (size_t)c->m, nnz = (size_t)c->nnz;
  if (d_off) GD_HIP_TRY(hipMemcpyAsync(d_off, c->row_offsets, 4 * (m + 1), hipMemcpyDeviceToDevice, s));
  if (d_col && nnz)
    GD_HIP_TRY(hipMemcpyAsync(d_col, c->column_indices, 4 * nnz, hipMemcpyDeviceToDevice, s));
  if (d_w && nnz) GD_HIP_TRY(hipMemcpyAsync(d_w, c->weight, 4 * nnz, hipMemcpyDeviceToDevice, s));
  if (d_deg && m) GD_HIP_TRY(hipMemcpyAsync(d_deg, c->degree, 4 * m, hipMemcpyDeviceToDevice, s));
  return GD_OK;
}

int gd_device_csr_free(gd_device_csr* c) {
  if (!c) return GD_OK;
  for (void* p : {(void*)c->row_offsets, (void*)c->column_indices, (void*)c->weight,
                  (void*)c->degree})
    if (p) (void)hipFree(p);
  memset(c, 0, sizeof(*c));
  return GD_OK;
}

}  // extern "C"
