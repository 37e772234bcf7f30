// schedule.hip -- host schedule builder, CSR validation kernel, buffers, errors.
#include <algorithm>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "gd_common.h"

namespace gd {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int DeviceBuffer::alloc(size_t n) {
  release();
  if (n == 0) n = 256;
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error(std::string("hipMalloc(") + std::to_string(n) + "): " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GD_ERR_ALLOC : GD_ERR_DEVICE;
  }
  bytes = n;
  return GD_OK;
}

void DeviceBuffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

namespace {

// The greedy tiling of rows [r0, r1): hub slots (and hub indices) numbered
// from 0 within the range; bad_row = first row with a negative length.
struct SchedulePart {
  std::vector<Tile> tiles;
  std::vector<int32_t> hub_row, hub_nseg;
  int32_t n_slots = 0, max_row_len = 0, bad_row = -1;
};

void build_part(int32_t r0, int32_t r1, const int32_t* off, int tile_nnz, int tile_rows,
                int exact_len, SchedulePart* o) {
  int32_t cur = r0, cur_nnz = 0;
  for (int32_t r = r0; r < r1; r++) {
    const int32_t d = off[r + 1] - off[r];
    if (d < 0) {
      o->bad_row = r;
      return;
    }
    if (d > o->max_row_len) o->max_row_len = d;
    if (exact_len > 0 && d > exact_len) {  // whole-row hub tile
      if (r > cur) o->tiles.push_back(Tile{cur, r, off[cur], off[r]});
      const int32_t h = (int32_t)o->hub_row.size();
      o->hub_row.push_back(r);
      o->hub_nseg.push_back(1);
      o->tiles.push_back(Tile{r, -(h + 1), off[r], off[r + 1]});
      o->n_slots += 1;
      cur = r + 1;
      cur_nnz = 0;
      continue;
    }
    if (d > tile_nnz) {
      if (r > cur) o->tiles.push_back(Tile{cur, r, off[cur], off[r]});
      const int32_t nseg = (d + tile_nnz - 1) / tile_nnz;
      o->hub_row.push_back(r);
      o->hub_nseg.push_back(nseg);
      for (int32_t k = 0; k < nseg; k++) {
        const int32_t e0 = off[r] + k * tile_nnz;
        const int32_t e1 = (k + 1 == nseg) ? off[r + 1] : e0 + tile_nnz;
        o->tiles.push_back(Tile{r, -(o->n_slots + k + 1), e0, e1});
      }
      o->n_slots += nseg;
      cur = r + 1;
      cur_nnz = 0;
      continue;
    }
    if (r > cur && (cur_nnz + d > tile_nnz || r - cur >= tile_rows)) {
      o->tiles.push_back(Tile{cur, r, off[cur], off[r]});
      cur = r;
      cur_nnz = 0;
    }
    cur_nnz += d;
  }
  if (r1 > cur) o->tiles.push_back(Tile{cur, r1, off[cur], off[r1]});
}

}  // namespace

int build_schedule(int m, int64_t nnz, const int32_t* off, int tile_nnz, int tile_rows,
                   Schedule* s, int exact_len) {
  s->tiles.clear();
  s->hub_row.clear();
  s->hub_slot.clear();
  s->hub_nseg.clear();
  s->n_slots = 0;
  s->max_row_len = 0;
  if (m < 0 || nnz < 0 || tile_nnz <= 0 || tile_rows <= 0 || exact_len < 0 ||
      exact_len > tile_nnz) {
    set_error("build_schedule: bad sizes");
    return GD_ERR_INVALID;
  }
  if (off[0] != 0 || (int64_t)off[m] != nnz) {
    set_error("row_offsets[0] must be 0 and row_offsets[m] must equal nnz");
    return GD_ERR_INVALID;
  }
  // Large graphs: kParts fixed row ranges, each tiled greedily on its own (a
  // tile never spans two ranges), built by up to 16 threads and joined in
  // range order.  The ranges depend on m only, so the schedule does not
  // depend on the thread count.
  constexpr int kParts = 64;
  const int parts = m >= (1 << 20) ? kParts : 1;
  std::vector<SchedulePart> part(parts);
  auto range = [&](int p) { return (int32_t)((int64_t)m * p / parts); };
  if (parts == 1) {
    build_part(0, m, off, tile_nnz, tile_rows, exact_len, &part[0]);
  } else {
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        for (int p = t; p < parts; p += nt)
          build_part(range(p), range(p + 1), off, tile_nnz, tile_rows, exact_len, &part[p]);
      });
    for (std::thread& th : pool) th.join();
  }
  size_t n_tiles = 0;
  for (const SchedulePart& pt : part) {
    if (pt.bad_row >= 0) {
      set_error("row_offsets must be non-decreasing (row " + std::to_string(pt.bad_row) + ")");
      return GD_ERR_INVALID;
    }
    n_tiles += pt.tiles.size();
  }
  s->tiles.reserve(n_tiles);
  for (const SchedulePart& pt : part) {
    const int32_t base = s->n_slots;  // slots (and hub indices) of the earlier ranges
    for (Tile t : pt.tiles) {
      if (t.y < 0) t.y -= base;
      s->tiles.push_back(t);
    }
    int32_t slot = base;
    for (size_t h = 0; h < pt.hub_row.size(); h++) {
      s->hub_row.push_back(pt.hub_row[h]);
      s->hub_slot.push_back(slot);
      s->hub_nseg.push_back(pt.hub_nseg[h]);
      slot += pt.hub_nseg[h];
    }
    s->n_slots += pt.n_slots;
    s->max_row_len = std::max(s->max_row_len, pt.max_row_len);
  }
  if (exact_len > 0 && !s->hub_row.empty()) {
    // Whole-row hubs of more than kLongHub edges first, longest first (they run
    // longest); the other hubs spread evenly among the row tiles, so their
    // scan work overlaps other tiles' gathers instead of running back to back.
    constexpr int32_t kLongHub = 65536;
    std::vector<Tile> hubs, rows;
    for (const Tile& t : s->tiles) (t.y < 0 ? hubs : rows).push_back(t);
    std::stable_sort(hubs.begin(), hubs.end(),
                     [](const Tile& a, const Tile& b) { return a.w - a.z > b.w - b.z; });
    size_t nlong = 0;
    while (nlong < hubs.size() && hubs[nlong].w - hubs[nlong].z > kLongHub) nlong++;
    s->tiles.assign(hubs.begin(), hubs.begin() + nlong);
    const size_t H = hubs.size() - nlong, R = rows.size();
    size_t h = 0;
    for (size_t r = 0; r <= R; r++) {
      // hub j goes before row tile floor(j * R / H)
      while (h < H && h * R <= r * H) s->tiles.push_back(hubs[nlong + h++]);
      if (r < R) s->tiles.push_back(rows[r]);
    }
  }
  return GD_OK;
}

__global__ void check_columns_kernel(const int32_t* __restrict__ col, int64_t nnz, int32_t ncols,
                                     int* __restrict__ bad) {
  int found = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = col[i];
    found |= (c < 0) | (c >= ncols);
  }
  if (__any(found) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

int check_columns(const int32_t* d_col, int64_t nnz, int32_t ncols, hipStream_t stream) {
  if (nnz == 0) return GD_OK;
  DeviceBuffer flag;
  int rc = flag.alloc(sizeof(int));
  if (rc) return rc;
  GD_HIP_TRY(hipMemsetAsync(flag.ptr, 0, sizeof(int), stream));
  int blocks = (int)std::min<int64_t>((nnz + 255) / 256, 4096);
  hipLaunchKernelGGL(check_columns_kernel, dim3(blocks), dim3(256), 0, stream, d_col, nnz, ncols,
                     flag.as<int>());
  GD_HIP_TRY(hipGetLastError());
  int h = 0;
  GD_HIP_TRY(hipMemcpyAsync(&h, flag.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
  GD_HIP_TRY(hipStreamSynchronize(stream));
  if (h) {
    set_error("column index out of range [0, ncols)");
    return GD_ERR_INVALID;
  }
  return GD_OK;
}

// flag[0]: off[0] != 0 or off[m] != nnz; flag[1]: the first row r with
// off[r+1] < off[r] (INT_MAX if none).
__global__ void check_offsets_kernel(const int32_t* __restrict__ off, int m, int64_t nnz,
                                     int* flag) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x)
    if (off[r + 1] < off[r]) atomicMin(&flag[1], r);
  if (blockIdx.x == 0 && threadIdx.x == 0 && (off[0] != 0 || (int64_t)off[m] != nnz)) flag[0] = 1;
}

int check_offsets(const int32_t* d_off, int m, int64_t nnz, hipStream_t stream) {
  DeviceBuffer flag;
  int rc = flag.alloc(2 * sizeof(int));
  if (rc) return rc;
  const int init[2] = {0, 0x7fffffff};
  GD_HIP_TRY(hipMemcpyAsync(flag.ptr, init, sizeof(init), hipMemcpyHostToDevice, stream));
  const int blocks = std::max(1, std::min((m + 255) / 256, 4096));
  hipLaunchKernelGGL(check_offsets_kernel, dim3(blocks), dim3(256), 0, stream, d_off, m, nnz,
                     flag.as<int>());
  GD_HIP_TRY(hipGetLastError());
  int h[2] = {0, 0};
  GD_HIP_TRY(hipMemcpyAsync(h, flag.ptr, sizeof(h), hipMemcpyDeviceToHost, stream));
  GD_HIP_TRY(hipStreamSynchronize(stream));
  if (h[0]) {
    set_error("row_offsets[0] must be 0 and row_offsets[m] must equal nnz");
    return GD_ERR_INVALID;
  }
  if (h[1] != 0x7fffffff) {
    set_error("row_offsets must be non-decreasing (row " + std::to_string(h[1]) + ")");
    return GD_ERR_INVALID;
  }
  return GD_OK;
}

}  // namespace gd

extern "C" const char* gd_last_error(void) { return gd::g_last_error.c_str(); }
