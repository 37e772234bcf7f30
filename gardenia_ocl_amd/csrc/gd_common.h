// gd_common.h -- internal helpers shared by the HIP translation units.
#ifndef GD_COMMON_H_
#define GD_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "gardenia_graph.h"

namespace gd {

// Thread-local text of the last failure, returned by gd_last_error().
void set_error(const std::string& msg);

#define GD_HIP_TRY(expr)                                                              \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      ::gd::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));             \
      return GD_ERR_DEVICE;                                                           \
    }                                                                                 \
  } while (0)

#define GD_CHECK_ARG(cond, msg)                \
  do {                                         \
    if (!(cond)) {                             \
      ::gd::set_error(msg);                    \
      return GD_ERR_INVALID;                   \
    }                                          \
  } while (0)

// Work unit of the pull engine (one workgroup each).
//   row tile:    x = first row, y = end row (exclusive) > x, z = first edge, w = end edge
//   hub segment: x = row, y = -(partial slot + 1) < 0,     z = first edge, w = end edge
//   (with exact hubs a hub "segment" is the whole row, summed in edge order)
struct Tile {
  int32_t x, y, z, w;
};

// Row-block schedule over a CSR (CSR-Adaptive style, Greathouse & Daga SC'14):
// consecutive rows are packed into tiles of at most `tile_nnz` edges and
// `tile_rows` rows; a row longer than `tile_nnz` becomes a hub and is split
// into ceil(deg/tile_nnz) segments whose partial sums are combined, in segment
// order, by a finalize pass.  With exact_len > 0 (<= tile_nnz) every row longer
// than exact_len is instead one whole-row hub tile (summed in edge order by one
// workgroup); the longest of them come first, the rest are spread evenly
// among the row tiles.
// Whole-row exact hubs of PageRank-like plans are summed by the workgroup's
// binade scan (exact_row_sum); long rows of the SpMV / SymGS / many-graph /
// no-allocation paths by one lane from LDS chunks of kHubChunk edges
// (hub_chain_sum).
constexpr int kHubChunk = 1536;    // 192 gathering lanes x 8 edges

struct Schedule {
  std::vector<Tile> tiles;
  std::vector<int32_t> hub_row, hub_slot, hub_nseg;
  int32_t n_slots = 0;
  int32_t max_row_len = 0;
};

// Builds the schedule from host row offsets.  Returns GD_ERR_INVALID (and sets
// the error text) if the offsets are not a valid CSR prefix sum of length m+1.
int build_schedule(int m, int64_t nnz, const int32_t* row_offsets, int tile_nnz, int tile_rows,
                   Schedule* out, int exact_len = 0);

// Checks on the device that every column index lies in [0, ncols); returns
// GD_OK, GD_ERR_INVALID or GD_ERR_DEVICE.
int check_columns(const int32_t* d_col, int64_t nnz, int32_t ncols, hipStream_t stream);

// Checks on the device that d_off[0..m] is a CSR prefix sum of nnz edges
// (off[0] == 0, off[m] == nnz, non-decreasing), with build_schedule's error
// texts; returns GD_OK, GD_ERR_INVALID or GD_ERR_DEVICE.
int check_offsets(const int32_t* d_off, int m, int64_t nnz, hipStream_t stream);

// Two timing events, destroyed on every exit path of the function that holds
// them (the GD_HIP_TRY early returns included).
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() = default;
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  hipError_t create() {
    hipError_t e = hipEventCreate(&a);
    return e != hipSuccess ? e : hipEventCreate(&b);
  }
};

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  ~DeviceBuffer() { release(); }
  int alloc(size_t n);
  void release();
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

}  // namespace gd

#endif  // GD_COMMON_H_
