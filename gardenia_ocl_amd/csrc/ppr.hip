// ppr.hip -- batched personalised PageRank (k seeds at once), gfx950.
//
// The same pull iteration as pr.hip, k columns wide: column j teleports its
// (1-d) mass to seed vertex seeds[j] instead of (1-d)/m to every vertex.
// Scores are stored vertex-major with the k seeds of a vertex contiguous
// (X[v*k + j]), so one gathered index u feeds k additions: a wavefront owns a
// chunk of kc = 64*VEC columns, lane l holds columns l*VEC..l*VEC+VEC-1, and
// the gathered row segment C[u][chunk] is one coalesced 256*VEC-byte load.
//
// Each column of every row is summed sequentially in CSR order, the scalar
// personalised loop's order and FP32 rounding, so every column is the
// single-seed run's bit for bit.  Rows longer than kSegEdges (hubs) run
// beside the tiles: one workgroup per (hub row, 16 columns), three waves
// gathering chunks of the row into LDS while the fourth adds them in edge
// order, one column per lane.  The per-iteration update is base + d*sum with
// two roundings (built with -ffp-contract=off).
//
// A persistent grid of workgroups walks the tile list; each workgroup keeps
// per-column L1 partials for the whole sweep, and two small kernels reduce
// them (deterministically) to per-column L1 values, their maximum and the
// convergence flag.
//
// Columns stop one by one: after the first iteration whose L1 change of
// column j is below tol, column j is frozen (its scores are final, exactly as
// a single-seed run would stop).  Frozen columns keep their values; between
// host checks the running columns are moved to the front of every vertex's
// k-vector (an in-place column permutation of X and the gathered contrib), so
// the lanes of frozen columns stop gathering.  The permutation is undone when
// the run ends.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "relabel.h"

namespace gd {

namespace {

constexpr int kPBlock = 256;
constexpr int kPWaves = kPBlock / 64;
constexpr int kSegEdges = 4096;   // rows longer than this are hubs (ppr_hub_chain_kernel)
constexpr int kTileEdges = 1024;  // edges per wavefront tile
constexpr int kTileRowsP = 256;   // rows per wavefront tile
constexpr int kUnroll = 8;        // gathers in flight per lane
constexpr int kSlices = 32;       // first-level L1 reduction slices

struct PprState {
  int iter;
  int done;
  int max_iter;
  int pad;
  double tol;
};

template <int VEC>
struct Vec {
  float v[VEC];
};

template <int VEC>
__device__ __forceinline__ Vec<VEC> load_vec(const float* __restrict__ p) {
  Vec<VEC> r;
  if constexpr (VEC == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
  } else if constexpr (VEC == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    r.v[0] = t.x; r.v[1] = t.y;
  } else {
    r.v[0] = *p;
  }
  return r;
}

template <int VEC>
__device__ __forceinline__ void store_vec(float* __restrict__ p, const float (&v)[VEC]) {
  if constexpr (VEC == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (VEC == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
    *p = v[0];
  }
}

// acc[j] += C[u][col0 + j] over the edges [b, e) in order, kUnroll gathers in
// flight.  b, e, k, col0 are wavefront-uniform.
template <int VEC>
__device__ __forceinline__ void sum_edges(const int32_t* __restrict__ col,
                                          const float* __restrict__ C, int64_t k, int col0,
                                          int b, int e, float (&acc)[VEC]) {
  int i = b;
  for (; i + kUnroll <= e; i += kUnroll) {
    Vec<VEC> g[kUnroll];
#pragma unroll
    for (int q = 0; q < kUnroll; q++) {
      const int u = __builtin_amdgcn_readfirstlane(col[i + q]);
      g[q] = load_vec<VEC>(C + (int64_t)u * k + col0);
    }
#pragma unroll
    for (int q = 0; q < kUnroll; q++)
#pragma unroll
      for (int j = 0; j < VEC; j++) acc[j] += g[q].v[j];
  }
  for (; i < e; i++) {
    const int u = __builtin_amdgcn_readfirstlane(col[i]);
    const Vec<VEC> g = load_vec<VEC>(C + (int64_t)u * k + col0);
#pragma unroll
    for (int j = 0; j < VEC; j++) acc[j] += g.v[j];
  }
}

// Frozen columns (fz) keep their score and add nothing to the L1 change.
template <int VEC>
__device__ __forceinline__ void row_update(int v, const float (&acc)[VEC], const int (&seed)[VEC],
                                           const int (&fz)[VEC], float* __restrict__ X,
                                           float* __restrict__ Cn,
                                           const int32_t* __restrict__ outdeg, int64_t k,
                                           int col0, float teleport, float damp,
                                           double (&err)[VEC]) {
  const int64_t o = (int64_t)v * k + col0;
  const Vec<VEC> old = load_vec<VEC>(X + o);
  const float deg = (float)outdeg[v];
  float nx[VEC], nc[VEC];
#pragma unroll
  for (int j = 0; j < VEC; j++) {
    const float b = (seed[j] == v) ? teleport : 0.0f;
    const float upd = b + damp * acc[j];
    nx[j] = fz[j] ? old.v[j] : upd;
    nc[j] = nx[j] / deg;
    err[j] += fz[j] ? 0.0 : (double)fabsf(nx[j] - old.v[j]);
  }
  store_vec<VEC>(X + o, nx);
  store_vec<VEC>(Cn + o, nc);
}

template <int VEC>
__global__ __launch_bounds__(kPBlock) void ppr_pull_kernel(
    const Tile* __restrict__ tiles, int n_tiles, const int32_t* __restrict__ off,
    const int32_t* __restrict__ col, const int32_t* __restrict__ outdeg,
    const int32_t* __restrict__ seeds, const int32_t* __restrict__ frozen, int ka, int64_t k,
    float* __restrict__ X, float* __restrict__ C0, float* __restrict__ C1,
    double* __restrict__ err_part, const PprState* __restrict__ st, float teleport, float damp) {
  if (st->done) return;
  const int odd = st->iter & 1;
  const float* __restrict__ C = odd ? C1 : C0;
  float* __restrict__ Cn = odd ? C0 : C1;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col0 = blockIdx.y * 64 * VEC + lane * VEC;
  int seed[VEC], fz[VEC];
#pragma unroll
  for (int j = 0; j < VEC; j++) {
    seed[j] = seeds[col0 + j];
    fz[j] = frozen[col0 + j];
  }
  double err[VEC];
#pragma unroll
  for (int j = 0; j < VEC; j++) err[j] = 0.0;
  const int nw = gridDim.x * kPWaves;
  if (col0 < ka) {  // lanes past the running columns gather nothing
    for (int t = blockIdx.x * kPWaves + wid; t < n_tiles; t += nw) {
      const Tile tl = tiles[t];
      for (int r = tl.x; r < tl.y; r++) {
        float acc[VEC];
#pragma unroll
        for (int j = 0; j < VEC; j++) acc[j] = 0.0f;
        sum_edges<VEC>(col, C, k, col0, off[r], off[r + 1], acc);
        row_update<VEC>(r, acc, seed, fz, X, Cn, outdeg, k, col0, teleport, damp, err);
      }
    }
  }
  // workgroup partial per column: waves in order, through LDS
  __shared__ double s_err[kPWaves][64 * VEC];
#pragma unroll
  for (int j = 0; j < VEC; j++) s_err[wid][lane * VEC + j] = err[j];
  __syncthreads();
  for (int c = threadIdx.x; c < 64 * VEC; c += kPBlock) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kPWaves; w++) s += s_err[w][c];
    err_part[(int64_t)blockIdx.x * k + blockIdx.y * 64 * VEC + c] = s;
  }
}

// Hub rows: workgroup b adds hub row hub_row[b / G] for the kHubCols columns
// (16, 32 or 64: gd_ppr_plan_config.hub_cols) of group g = b % G (all groups
// of the longest row are dispatched first) in
// edge order, then runs their row updates.  Lane c < kHubCols of wave 0 adds
// column kHubCols*g + c chunk by chunk from LDS (its own dependent chain: the
// scalar loop's order and rounding) while waves 1..3 gather the next chunk.
// A gathering instruction loads 64/kHubCols edges' column groups at once
// (lane l: edge l / kHubCols, column l % kHubCols), so each gathering wave
// keeps kHubLoads * 64/kHubCols edges in flight: the chain of a 10^5-edge row
// is bound by load latency, and narrower column groups put more of its edges
// in flight (and spread the row over more workgroups).  hub_err row b / G
// receives the row's per-column |change|.
constexpr int kHubLoads = 32;  // gathering instructions per wave per chunk

template <int kHubCols>  // columns per workgroup: 16, 32 or 64
__global__ __launch_bounds__(kPBlock) void ppr_hub_chain_kernel(
    const int32_t* __restrict__ hub_row, int groups, const int32_t* __restrict__ off,
    const int32_t* __restrict__ col, const int32_t* __restrict__ outdeg,
    const int32_t* __restrict__ seeds, const int32_t* __restrict__ frozen, int ka, int64_t k,
    float* __restrict__ X, float* __restrict__ C0, float* __restrict__ C1,
    double* __restrict__ hub_err, const PprState* __restrict__ st, float teleport, float damp) {
  constexpr int kHubEpi = 64 / kHubCols;           // edges per gathering instruction
  constexpr int kHubEdgesW = kHubLoads * kHubEpi;  // edges per gathering wave per chunk
  constexpr int kHubChunkE = 3 * kHubEdgesW;       // edges per chunk
  constexpr int kHubIdRegs = (kHubEdgesW + 63) / 64;
  if (st->done) return;
  __shared__ float buf[2][kHubChunkE * kHubCols];
  const int odd = st->iter & 1;
  const float* __restrict__ C = odd ? C1 : C0;
  float* __restrict__ Cn = odd ? C0 : C1;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int h = blockIdx.x / groups;
  const int g0 = (blockIdx.x % groups) * kHubCols;  // first column of the group (< k)
  const int row = hub_row[h];
  const int e0 = off[row];
  const int ne = off[row + 1] - e0;
  const int nchunks = (ne + kHubChunkE - 1) / kHubChunkE;
  const bool running = g0 < ka;  // uniform: a group past the running columns idles
  const int gw = wid - 1;  // gathering wave 0..2
  const int my_col = g0 + (lane % kHubCols);
  // source ids of this wave's edges of chunk c: lane l holds edge l + 64 q
  auto load_ids = [&](int c, int (&ids)[kHubIdRegs]) {
#pragma unroll
    for (int q = 0; q < kHubIdRegs; q++) {
      const int e = c * kHubChunkE + gw * kHubEdgesW + q * 64 + lane;
      ids[q] = (q * 64 + lane < kHubEdgesW && e < ne) ? col[e0 + e] : 0;
    }
  };
  auto produce = [&](int c, const int (&ids)[kHubIdRegs], float* dst) {
    const int base = c * kHubChunkE + gw * kHubEdgesW;  // first edge of this wave's part
    const int cnt = min(kHubEdgesW, ne - base);
    float v[kHubLoads];
#pragma unroll
    for (int i = 0; i < kHubLoads; i++) {
      const int e = i * kHubEpi + lane / kHubCols;  // this lane's edge in the wave's part
      const int u = __shfl(ids[(i * kHubEpi) / 64], e % 64, 64);  // 0 past the row: valid
      v[i] = C[(int64_t)u * k + my_col];
    }
#pragma unroll
    for (int i = 0; i < kHubLoads; i++)
      if (i * kHubEpi + lane / kHubCols < cnt)
        dst[(gw * kHubEdgesW) * kHubCols + i * 64 + lane] = v[i];
  };
  float s = 0.0f;
  if (running) {
    int ids[kHubIdRegs], nxt[kHubIdRegs];
    if (wid != 0 && nchunks > 0) {
      load_ids(0, ids);
      if (nchunks > 1) load_ids(1, nxt);
      produce(0, ids, buf[0]);
#pragma unroll
      for (int q = 0; q < kHubIdRegs; q++) ids[q] = nxt[q];
    }
    __syncthreads();
    for (int c = 0; c < nchunks; c++) {
      if (wid == 0) {
        if (lane < kHubCols) {
          const float* __restrict__ b = buf[c & 1] + lane;
          const int nc = min(kHubChunkE, ne - c * kHubChunkE);
          int e = 0;
          for (; e + 8 <= nc; e += 8) {
            float a[8];
#pragma unroll
            for (int q = 0; q < 8; q++) a[q] = b[(e + q) * kHubCols];
#pragma unroll
            for (int q = 0; q < 8; q++) s += a[q];
          }
          for (; e < nc; e++) s += b[e * kHubCols];
        }
      } else if (c + 1 < nchunks) {
        if (c + 2 < nchunks) load_ids(c + 2, nxt);  // in flight across the barrier
        produce(c + 1, ids, buf[(c + 1) & 1]);
#pragma unroll
        for (int q = 0; q < kHubIdRegs; q++) ids[q] = nxt[q];
      }
      __syncthreads();
    }
  }
  if (wid == 0 && lane < kHubCols) {
    double err[1] = {0.0};
    if (my_col < ka) {
      const float acc[1] = {s};
      const int sd[1] = {seeds[my_col]}, fz[1] = {frozen[my_col]};
      row_update<1>(row, acc, sd, fz, X, Cn, outdeg, k, my_col, teleport, damp, err);
    }
    hub_err[(int64_t)h * k + my_col] = err[0];
  }
}

// First level: fin[s][c] = sum over the rows of err_part in slice s (fixed order).
__global__ __launch_bounds__(kPBlock) void ppr_err_slices_kernel(const double* __restrict__ part,
                                                                 int rows, int64_t k,
                                                                 int64_t kact,
                                                                 double* __restrict__ fin,
                                                                 const PprState* __restrict__ st) {
  if (st->done) return;
  const int64_t c = (int64_t)blockIdx.x * kPBlock + threadIdx.x;
  if (c >= kact) return;
  const int per = (rows + kSlices - 1) / kSlices;
  const int lo = blockIdx.y * per, hi = min(lo + per, rows);
  double s = 0.0;
  for (int r = lo; r < hi; r++) s += part[(int64_t)r * k + c];
  fin[(int64_t)blockIdx.y * k + c] = s;
}

// Second level: per-column L1 (slices in order) of the columns still running
// (c < kact, not frozen), their maximum, the per-column stop and the state
// update.  With column_stop a column whose L1 change is below tol is frozen.
__global__ __launch_bounds__(kPBlock) void ppr_finalize_kernel(
    const double* __restrict__ fin, int64_t k, int64_t kact, double* __restrict__ col_err,
    int32_t* __restrict__ frozen, int32_t* __restrict__ col_iters, int column_stop,
    PprState* __restrict__ st, double* __restrict__ errors) {
  if (st->done) return;
  __shared__ double s_max[kPBlock];
  const int it = st->iter;
  const double tol = st->tol;
  double mx = 0.0;
  for (int64_t c = threadIdx.x; c < kact; c += kPBlock) {
    if (frozen[c]) continue;
    double s = 0.0;
    for (int q = 0; q < 2 * kSlices; q++) s += fin[(int64_t)q * k + c];
    col_err[c] = s;
    col_iters[c] = it + 1;
    mx = fmax(mx, s);
    if (column_stop && s < tol) frozen[c] = 1;
  }
  s_max[threadIdx.x] = mx;
  __syncthreads();
  for (int w = kPBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    errors[it] = s_max[0];
    st->iter = it + 1;
    // every running column below tol <=> the max over them is
    if (s_max[0] < tol || it + 1 >= st->max_iter) st->done = 1;
  }
}

__global__ void ppr_reset_cols_kernel(int k, int32_t* __restrict__ frozen,
                                      int32_t* __restrict__ col_iters) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) {
    frozen[j] = 0;
    col_iters[j] = 0;
  }
}

// In-place column permutation of the rows of A (and of B when not NULL):
// row[c] <- row[newcol[c]] for c < kact, newcol a permutation of [0, kact).
// One wavefront per row through kact floats of LDS; every workgroup runs the
// same number of loop trips, so the barriers are uniform.
__global__ __launch_bounds__(kPBlock) void ppr_permute_cols_kernel(int m, int64_t k, int kact,
                                                                   const int32_t* __restrict__ newcol,
                                                                   float* __restrict__ A,
                                                                   float* __restrict__ B) {
  extern __shared__ float s_rows[];  // kPWaves * kact
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  float* my = s_rows + (size_t)wid * kact;
  const int passes = B ? 2 : 1;
  for (int base = blockIdx.x * kPWaves; base < m; base += gridDim.x * kPWaves) {
    const int v = base + wid;
    for (int pass = 0; pass < passes; pass++) {
      float* row = (pass ? B : A) + (int64_t)min(v, m - 1) * k;
      if (v < m)
        for (int c = lane; c < kact; c += 64) my[c] = row[c];
      __syncthreads();
      if (v < m)
        for (int c = lane; c < kact; c += 64) row[c] = my[newcol[c]];
      __syncthreads();
    }
  }
}

__global__ void ppr_init_kernel(int64_t n, int64_t k, const float* __restrict__ X,
                                const int32_t* __restrict__ outdeg, float* __restrict__ C0,
                                PprState* __restrict__ st, int max_iter, double tol) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    C0[i] = X[i] / (float)outdeg[i / k];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->tol = tol;
  }
}

// rows of X between the user's vertex order and the plan's; on the way back
// (cperm not NULL) internal column j goes to user column cperm[j]
__global__ void ppr_permute_rows_kernel(int m, int64_t k, const int32_t* __restrict__ perm,
                                        const int32_t* __restrict__ cperm,
                                        const float* __restrict__ src, float* __restrict__ dst,
                                        int to_internal) {
  const int lane = threadIdx.x & 63;
  for (int i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; i < m;
       i += gridDim.x * (blockDim.x / 64)) {
    const int v = perm[i];
    const float* s = src + (int64_t)(to_internal ? v : i) * k;
    float* d = dst + (int64_t)(to_internal ? i : v) * k;
    for (int64_t j = lane; j < k; j += 64) d[cperm ? cperm[j] : j] = s[j];
  }
}

__global__ void ppr_map_seeds_kernel(int k, const int32_t* __restrict__ seeds,
                                     const int32_t* __restrict__ inv, int32_t* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) out[j] = inv ? inv[seeds[j]] : seeds[j];
}

__global__ void check_seeds_kernel(int k, const int32_t* __restrict__ seeds, int m,
                                   int* __restrict__ bad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k && (seeds[j] < 0 || seeds[j] >= m)) atomicOr(bad, 1);
}

}  // namespace

}  // namespace gd

using namespace gd;

struct gd_ppr_plan {
  int m = 0, k = 0, kc = 64;
  int hub_cols = 64;  // columns per hub-row workgroup
  int64_t nnz = 0;
  hipStream_t stream = nullptr;
  bool relabel = false;
  const int32_t* off = nullptr;
  const int32_t* col = nullptr;
  const int32_t* outdeg = nullptr;
  Relabel rl;
  Schedule sched;
  int n_tiles = 0, n_hubs = 0, grid = 0;
  // hub rows run on the plan's stream beside the tiles on `side` (fork / join)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  ~gd_ppr_plan() {
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side) (void)hipStreamDestroy(side);
  }
  DeviceBuffer tiles, hub_row, err_part, hub_err, fin, col_err,
      C0, C1, Xint, seeds, state, errors, frozen, col_iters, newcol, cperm;
  int errors_cap = 0;
  bool column_stop = true;
  bool compact = true;
  std::vector<int32_t> last_col_iters;  // user column order
  float plan_ms = 0.0f;
};

namespace {

int ppr_build(gd_ppr_plan* p, const int32_t* d_off, const int32_t* d_col,
              const int32_t* d_outdeg) {
  const int m = p->m;
  int rc = check_columns(d_col, p->nnz, m, p->stream);
  if (rc) return rc;
  rc = check_offsets(d_off, m, p->nnz, p->stream);  // before any device walk
  if (rc) return rc;
  std::vector<int32_t> h_off(m + 1);
  p->off = d_off;
  p->col = d_col;
  p->outdeg = d_outdeg;
  if (p->relabel) {
    rc = build_relabel(m, p->nnz, d_off, d_col, d_outdeg, p->stream, &p->rl);
    if (rc) return rc;
    p->off = p->rl.off.as<int32_t>();
    p->col = p->rl.col.as<int32_t>();
    p->outdeg = p->rl.outdeg.as<int32_t>();
  }
  GD_HIP_TRY(hipMemcpyAsync(h_off.data(), p->off, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  // wavefront tiles: <= kTileEdges edges / kTileRowsP rows; rows > kSegEdges are
  // hubs (ppr_hub_chain_kernel), longest first
  Schedule& s = p->sched;
  {
    Schedule hubs;
    rc = build_schedule(m, p->nnz, h_off.data(), kSegEdges, m, &hubs);
    if (rc) return rc;
    s.hub_row = hubs.hub_row;
    s.max_row_len = hubs.max_row_len;
    std::stable_sort(s.hub_row.begin(), s.hub_row.end(), [&](int32_t a, int32_t b) {
      return h_off[a + 1] - h_off[a] > h_off[b + 1] - h_off[b];
    });
    // re-tile the non-hub rows with the smaller wavefront tiles
    Schedule small;
    rc = build_schedule(m, p->nnz, h_off.data(), kTileEdges, kTileRowsP, &small);
    if (rc) return rc;
    // heaviest work first: single rows of (kTileEdges, kSegEdges] edges, then
    // the packed row tiles
    for (size_t h = 0; h < small.hub_row.size(); h++) {
      const int r = small.hub_row[h];
      const int d = h_off[r + 1] - h_off[r];
      if (d <= kSegEdges) s.tiles.push_back(Tile{r, r + 1, h_off[r], h_off[r + 1]});
    }
    for (const Tile& t : small.tiles)
      if (t.y >= 0) s.tiles.push_back(t);
  }
  p->n_tiles = (int)s.tiles.size();
  p->n_hubs = (int)s.hub_row.size();
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  p->grid = std::max(1, std::min(cus * 8, (p->n_tiles + kPWaves - 1) / kPWaves));
  const size_t mk = (size_t)m * p->k;
  rc = p->tiles.alloc(sizeof(Tile) * p->n_tiles);
  if (!rc) rc = p->hub_row.alloc(4 * (size_t)std::max(1, p->n_hubs));
  if (!rc) rc = p->err_part.alloc(8 * (size_t)p->grid * p->k);
  if (!rc) rc = p->hub_err.alloc(8 * (size_t)std::max(1, p->n_hubs) * p->k);
  if (!rc) rc = p->fin.alloc(8 * (size_t)2 * kSlices * p->k);
  if (!rc) rc = p->col_err.alloc(8 * (size_t)p->k);
  if (!rc) rc = p->C0.alloc(4 * mk);
  if (!rc) rc = p->C1.alloc(4 * mk);
  if (!rc && p->relabel) rc = p->Xint.alloc(4 * mk);
  if (!rc) rc = p->seeds.alloc(4 * (size_t)p->k);
  if (!rc) rc = p->frozen.alloc(4 * (size_t)p->k);
  if (!rc) rc = p->col_iters.alloc(4 * (size_t)p->k);
  if (!rc) rc = p->newcol.alloc(4 * (size_t)p->k);
  if (!rc) rc = p->cperm.alloc(4 * (size_t)p->k);
  if (!rc) rc = p->state.alloc(sizeof(PprState));
  if (rc) return rc;
  GD_HIP_TRY(hipMemcpyAsync(p->tiles.ptr, s.tiles.data(), sizeof(Tile) * p->n_tiles,
                            hipMemcpyHostToDevice, p->stream));
  if (p->n_hubs) {
    GD_HIP_TRY(hipMemcpyAsync(p->hub_row.ptr, s.hub_row.data(), 4 * p->n_hubs,
                              hipMemcpyHostToDevice, p->stream));
    GD_HIP_TRY(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    GD_HIP_TRY(hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming));
    GD_HIP_TRY(hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming));
  }
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  return GD_OK;
}

// One iteration over the running columns [0, ka) (whole chunks of 64*VEC).
template <int VEC>
void ppr_launch_iteration(gd_ppr_plan* p, float* X, PprState* st, float teleport, float damp,
                          int ka) {
  const int kc = 64 * VEC;
  const int chunks = (ka + kc - 1) / kc;
  const int64_t kact = (int64_t)chunks * kc;
  float* C0 = p->C0.as<float>();
  float* C1 = p->C1.as<float>();
  const int32_t* frozen = p->frozen.as<int32_t>();
  // hub rows first, on the plan's stream, so their workgroups are placed before
  // the tiles fill the CUs; the tiles fork onto the side stream and join back
  hipStream_t ts = p->stream;
  if (p->n_hubs) {
    (void)hipEventRecord(p->fork_ev, p->stream);
    (void)hipStreamWaitEvent(p->side, p->fork_ev, 0);
    ts = p->side;
    const int groups = (int)(kact / p->hub_cols);
    const dim3 hg((unsigned)(p->n_hubs * groups));
    if (p->hub_cols == 16)
      hipLaunchKernelGGL(ppr_hub_chain_kernel<16>, hg, dim3(kPBlock), 0, p->stream,
                         p->hub_row.as<int32_t>(), groups, p->off, p->col, p->outdeg,
                         p->seeds.as<int32_t>(), frozen, ka, (int64_t)p->k, X, C0, C1,
                         p->hub_err.as<double>(), st, teleport, damp);
    else if (p->hub_cols == 32)
      hipLaunchKernelGGL(ppr_hub_chain_kernel<32>, hg, dim3(kPBlock), 0, p->stream,
                         p->hub_row.as<int32_t>(), groups, p->off, p->col, p->outdeg,
                         p->seeds.as<int32_t>(), frozen, ka, (int64_t)p->k, X, C0, C1,
                         p->hub_err.as<double>(), st, teleport, damp);
    else
      hipLaunchKernelGGL(ppr_hub_chain_kernel<64>, hg, dim3(kPBlock), 0, p->stream,
                         p->hub_row.as<int32_t>(), groups, p->off, p->col, p->outdeg,
                         p->seeds.as<int32_t>(), frozen, ka, (int64_t)p->k, X, C0, C1,
                         p->hub_err.as<double>(), st, teleport, damp);
  }
  hipLaunchKernelGGL(ppr_pull_kernel<VEC>, dim3(p->grid, chunks), dim3(kPBlock), 0, ts,
                     p->tiles.as<Tile>(), p->n_tiles, p->off, p->col, p->outdeg,
                     p->seeds.as<int32_t>(), frozen, ka, (int64_t)p->k, X, C0, C1,
                     p->err_part.as<double>(), st, teleport, damp);
  if (p->n_hubs) {
    (void)hipEventRecord(p->join_ev, p->side);
    (void)hipStreamWaitEvent(p->stream, p->join_ev, 0);
  }
  const dim3 sg((unsigned)((kact + kPBlock - 1) / kPBlock), kSlices);
  double* fin = p->fin.as<double>();
  hipLaunchKernelGGL(ppr_err_slices_kernel, sg, dim3(kPBlock), 0, p->stream,
                     p->err_part.as<double>(), p->grid, (int64_t)p->k, kact, fin, st);
  hipLaunchKernelGGL(ppr_err_slices_kernel, sg, dim3(kPBlock), 0, p->stream,
                     p->hub_err.as<double>(), p->n_hubs, (int64_t)p->k, kact,
                     fin + (size_t)kSlices * p->k, st);
  hipLaunchKernelGGL(ppr_finalize_kernel, dim3(1), dim3(kPBlock), 0, p->stream, fin,
                     (int64_t)p->k, kact, p->col_err.as<double>(), p->frozen.as<int32_t>(),
                     p->col_iters.as<int32_t>(), p->column_stop ? 1 : 0, st,
                     p->errors.as<double>());
}

void ppr_launch(gd_ppr_plan* p, float* X, PprState* st, float teleport, float damp, int ka) {
  if (p->kc == 256)
    ppr_launch_iteration<4>(p, X, st, teleport, damp, ka);
  else if (p->kc == 128)
    ppr_launch_iteration<2>(p, X, st, teleport, damp, ka);
  else
    ppr_launch_iteration<1>(p, X, st, teleport, damp, ka);
}

// Applies `order` (new column c <- old column order[c], c < ka) to the rows of
// X and of the contrib buffer the next iteration gathers, and to the k-length
// per-column arrays; cperm (internal -> user column) follows.
int ppr_compact(gd_ppr_plan* p, float* X, float* Cnext, int ka, const std::vector<int32_t>& order,
                std::vector<int32_t>& cperm) {
  GD_HIP_TRY(hipMemcpyAsync(p->newcol.ptr, order.data(), 4 * (size_t)ka, hipMemcpyHostToDevice,
                            p->stream));
  const int m = p->m;
  const int grid = std::max(1, std::min(16384, (m + kPWaves - 1) / kPWaves));
  hipLaunchKernelGGL(ppr_permute_cols_kernel, dim3(grid), dim3(kPBlock),
                     (size_t)kPWaves * ka * sizeof(float), p->stream, m, (int64_t)p->k, ka,
                     p->newcol.as<int32_t>(), X, Cnext);
  GD_HIP_TRY(hipGetLastError());
  std::vector<int32_t> seeds(ka), fz(ka), its(ka);
  std::vector<double> ce(ka);
  GD_HIP_TRY(hipMemcpyAsync(seeds.data(), p->seeds.ptr, 4 * (size_t)ka, hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipMemcpyAsync(fz.data(), p->frozen.ptr, 4 * (size_t)ka, hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipMemcpyAsync(its.data(), p->col_iters.ptr, 4 * (size_t)ka, hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipMemcpyAsync(ce.data(), p->col_err.ptr, 8 * (size_t)ka, hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  std::vector<int32_t> seeds2(ka), fz2(ka), its2(ka), cp2(ka);
  std::vector<double> ce2(ka);
  for (int c = 0; c < ka; c++) {
    const int o = order[c];
    seeds2[c] = seeds[o];
    fz2[c] = fz[o];
    its2[c] = its[o];
    ce2[c] = ce[o];
    cp2[c] = cperm[o];
  }
  for (int c = 0; c < ka; c++) cperm[c] = cp2[c];
  GD_HIP_TRY(hipMemcpyAsync(p->seeds.ptr, seeds2.data(), 4 * (size_t)ka, hipMemcpyHostToDevice,
                            p->stream));
  GD_HIP_TRY(hipMemcpyAsync(p->frozen.ptr, fz2.data(), 4 * (size_t)ka, hipMemcpyHostToDevice,
                            p->stream));
  GD_HIP_TRY(hipMemcpyAsync(p->col_iters.ptr, its2.data(), 4 * (size_t)ka, hipMemcpyHostToDevice,
                            p->stream));
  GD_HIP_TRY(hipMemcpyAsync(p->col_err.ptr, ce2.data(), 8 * (size_t)ka, hipMemcpyHostToDevice,
                            p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));  // host vectors go out of scope
  return GD_OK;
}

}  // namespace

extern "C" {

void gd_ppr_default_plan_config(gd_ppr_plan_config* c) {
  c->relabel = -1;
  c->k_chunk = 0;
  c->column_stop = -1;
  c->compact = -1;
  c->hub_cols = 0;
}

int gd_ppr_plan_create(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                       const int32_t* d_outdeg, int k, void* stream, const gd_ppr_plan_config* cfg,
                       gd_ppr_plan** out) {
  GD_CHECK_ARG(out != nullptr, "plan output pointer is NULL");
  *out = nullptr;
  GD_CHECK_ARG(m >= 0 && nnz >= 0 && nnz < (1ll << 31), "m/nnz out of range");
  GD_CHECK_ARG(m == 0 || (d_off && d_outdeg), "NULL device array");
  GD_CHECK_ARG(nnz == 0 || d_col, "NULL column indices");
  GD_CHECK_ARG(k > 0 && k % 64 == 0, "k must be a positive multiple of 64");
  gd_ppr_plan_config c;
  if (cfg)
    c = *cfg;
  else
    gd_ppr_default_plan_config(&c);
  GD_CHECK_ARG(c.k_chunk == 0 || c.k_chunk == 64 || c.k_chunk == 128 || c.k_chunk == 256,
               "k_chunk must be 0, 64, 128 or 256");
  GD_CHECK_ARG(c.k_chunk == 0 || k % c.k_chunk == 0, "k must be a multiple of k_chunk");
  GD_CHECK_ARG(c.hub_cols == 0 || c.hub_cols == 16 || c.hub_cols == 32 || c.hub_cols == 64,
               "hub_cols must be 0, 16, 32 or 64");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  gd_ppr_plan* p = new gd_ppr_plan();
  p->m = m;
  p->nnz = nnz;
  p->k = k;
  p->kc = c.k_chunk ? c.k_chunk : (k % 256 == 0 ? 256 : (k % 128 == 0 ? 128 : 64));
  p->stream = (hipStream_t)stream;
  p->relabel = (c.relabel == 1) || (c.relabel == -1 && nnz >= (1 << 20));
  p->column_stop = c.column_stop != 0;
  // LDS of the in-place column permutation: kPWaves rows of k floats <= 64 KB
  p->compact = p->column_stop && c.compact != 0 && k <= 4096;
  p->hub_cols = c.hub_cols ? c.hub_cols : 64;
  int rc = GD_OK;
  if (m > 0) {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess)
      (void)hipEventRecord(t0, p->stream);
    rc = ppr_build(p, d_off, d_col, d_outdeg);
    if (t0 && t1) {
      (void)hipEventRecord(t1, p->stream);
      (void)hipEventSynchronize(t1);
      (void)hipEventElapsedTime(&p->plan_ms, t0, t1);
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
  }
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return GD_OK;
}

int gd_ppr_plan_info(const gd_ppr_plan* p, gd_ppr_plan_info_t* info) {
  GD_CHECK_ARG(p != nullptr && info != nullptr, "NULL plan / info");
  info->n_tiles = p->n_tiles;
  info->n_hubs = p->n_hubs;
  info->k = p->k;
  info->k_chunk = p->kc;
  info->relabel = p->relabel ? 1 : 0;
  info->grid = p->grid;
  info->plan_ms = p->plan_ms;
  info->hub_cols = p->hub_cols;
  return GD_OK;
}

int gd_ppr_plan_column_iterations(const gd_ppr_plan* p, int32_t* iterations) {
  GD_CHECK_ARG(p != nullptr && iterations != nullptr, "NULL plan / output");
  for (int j = 0; j < p->k; j++)
    iterations[j] = j < (int)p->last_col_iters.size() ? p->last_col_iters[j] : 0;
  return GD_OK;
}

int gd_ppr_plan_destroy(gd_ppr_plan* p) {
  delete p;
  return GD_OK;
}

int gd_ppr_plan_run(gd_ppr_plan* p, const int32_t* d_seeds, float* d_X,
                    const gd_pr_options* opts, gd_pr_stats* stats, double* errors,
                    double* col_errors) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  GD_CHECK_ARG(o.max_iter >= 0, "max_iter must be >= 0");
  gd_pr_stats s;
  memset(&s, 0, sizeof(s));
  const int m = p->m;
  if (m == 0) {
    if (stats) *stats = s;
    return GD_OK;
  }
  GD_CHECK_ARG(d_seeds != nullptr && d_X != nullptr, "NULL seeds / X");
  {
    DeviceBuffer bad;
    int rc = bad.alloc(4);
    if (rc) return rc;
    GD_HIP_TRY(hipMemsetAsync(bad.ptr, 0, 4, p->stream));
    hipLaunchKernelGGL(check_seeds_kernel, dim3((p->k + 255) / 256), dim3(256), 0, p->stream,
                       p->k, d_seeds, m, bad.as<int>());
    int h = 0;
    GD_HIP_TRY(hipMemcpyAsync(&h, bad.ptr, 4, hipMemcpyDeviceToHost, p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    GD_CHECK_ARG(h == 0, "seed vertex out of range");
  }
  if (p->errors_cap < std::max(1, o.max_iter)) {
    int rc = p->errors.alloc(sizeof(double) * std::max(1, o.max_iter));
    if (rc) return rc;
    p->errors_cap = std::max(1, o.max_iter);
  }
  const float damp = o.damping, teleport = 1.0f - damp;
  const int chunk = o.chunk > 0 ? o.chunk : (o.tol > 0 ? 4 : std::max(1, o.max_iter));
  const int k = p->k;
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  PprState* st = p->state.as<PprState>();
  float* X = p->relabel ? p->Xint.as<float>() : d_X;
  const int64_t n = (int64_t)m * k;
  GD_HIP_TRY(hipEventRecord(t0, p->stream));
  hipLaunchKernelGGL(ppr_map_seeds_kernel, dim3((k + 255) / 256), dim3(256), 0, p->stream, k,
                     d_seeds, p->relabel ? p->rl.inv.as<int32_t>() : nullptr,
                     p->seeds.as<int32_t>());
  if (p->relabel)
    hipLaunchKernelGGL(ppr_permute_rows_kernel, dim3(std::min(16384, (m + 3) / 4)), dim3(256), 0,
                       p->stream, m, (int64_t)k, p->rl.perm.as<int32_t>(), nullptr, d_X, X, 1);
  hipLaunchKernelGGL(ppr_init_kernel, dim3((int)std::min<int64_t>((n + 255) / 256, 65536)),
                     dim3(256), 0, p->stream, n, (int64_t)k, X, p->outdeg, p->C0.as<float>(),
                     st, o.max_iter, o.tol);
  hipLaunchKernelGGL(ppr_reset_cols_kernel, dim3((k + 255) / 256), dim3(256), 0, p->stream, k,
                     p->frozen.as<int32_t>(), p->col_iters.as<int32_t>());
  GD_HIP_TRY(hipGetLastError());
  std::vector<int32_t> cperm(k);  // internal column -> user column
  for (int j = 0; j < k; j++) cperm[j] = j;
  int ka = k;  // columns [0, ka) may still be running; the rest are frozen
  int launched = 0;
  PprState h{};
  std::vector<int32_t> fz;
  while (launched < o.max_iter) {
    const int nl = std::min(chunk, o.max_iter - launched);
    for (int q = 0; q < nl; q++) ppr_launch(p, X, st, teleport, damp, ka);
    GD_HIP_TRY(hipGetLastError());
    launched += nl;
    GD_HIP_TRY(hipMemcpyAsync(&h, st, sizeof(PprState), hipMemcpyDeviceToHost, p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    if (h.done) break;
    if (!p->compact) continue;
    fz.resize(ka);
    GD_HIP_TRY(hipMemcpy(fz.data(), p->frozen.ptr, 4 * (size_t)ka, hipMemcpyDeviceToHost));
    int nf = 0;
    for (int c = 0; c < ka; c++) nf += fz[c] != 0;
    // a permutation costs about one pass over X and the contrib; it pays once a
    // few iterations skip the frozen columns' gathers
    if (nf < std::max(32, ka / 8) || nf >= ka) continue;
    std::vector<int32_t> order;
    order.reserve(ka);
    for (int c = 0; c < ka; c++)
      if (!fz[c]) order.push_back(c);
    for (int c = 0; c < ka; c++)
      if (fz[c]) order.push_back(c);
    float* Cnext = (h.iter & 1) ? p->C1.as<float>() : p->C0.as<float>();
    const int rc = ppr_compact(p, X, Cnext, ka, order, cperm);
    if (rc) return rc;
    ka -= nf;
  }
  // back to the user's vertex and column order
  const bool permuted = ka < k;
  if (permuted)
    GD_HIP_TRY(hipMemcpyAsync(p->cperm.ptr, cperm.data(), 4 * (size_t)k, hipMemcpyHostToDevice,
                              p->stream));
  if (p->relabel) {
    hipLaunchKernelGGL(ppr_permute_rows_kernel, dim3(std::min(16384, (m + 3) / 4)), dim3(256), 0,
                       p->stream, m, (int64_t)k, p->rl.perm.as<int32_t>(),
                       permuted ? p->cperm.as<int32_t>() : nullptr, X, d_X, 0);
  } else if (permuted) {
    std::vector<int32_t> inv(k);  // user column u <- internal column inv[u]
    for (int j = 0; j < k; j++) inv[cperm[j]] = j;
    GD_HIP_TRY(hipMemcpyAsync(p->newcol.ptr, inv.data(), 4 * (size_t)k, hipMemcpyHostToDevice,
                              p->stream));
    const int grid = std::max(1, std::min(16384, (m + kPWaves - 1) / kPWaves));
    hipLaunchKernelGGL(ppr_permute_cols_kernel, dim3(grid), dim3(kPBlock),
                       (size_t)kPWaves * k * sizeof(float), p->stream, m, (int64_t)k, k,
                       p->newcol.as<int32_t>(), X, (float*)nullptr);
    GD_HIP_TRY(hipStreamSynchronize(p->stream));  // inv goes out of scope
  }
  GD_HIP_TRY(hipGetLastError());
  GD_HIP_TRY(hipEventRecord(t1, p->stream));
  GD_HIP_TRY(hipEventSynchronize(t1));
  GD_HIP_TRY(hipEventElapsedTime(&s.solve_ms, t0, t1));
  s.iterations = h.iter;
  std::vector<int32_t> its(k);
  std::vector<double> ce(k);
  GD_HIP_TRY(hipMemcpy(its.data(), p->col_iters.ptr, 4 * (size_t)k, hipMemcpyDeviceToHost));
  GD_HIP_TRY(hipMemcpy(ce.data(), p->col_err.ptr, 8 * (size_t)k, hipMemcpyDeviceToHost));
  p->last_col_iters.assign(k, 0);
  for (int j = 0; j < k; j++) p->last_col_iters[cperm[j]] = its[j];
  if (h.iter > 0) {
    std::vector<double> errs(h.iter);
    GD_HIP_TRY(hipMemcpy(errs.data(), p->errors.ptr, 8 * (size_t)h.iter, hipMemcpyDeviceToHost));
    s.final_error = errs[h.iter - 1];
    s.converged = s.final_error < o.tol;
    if (errors)
      for (int i = 0; i < h.iter; i++) errors[i] = errs[i];
    if (col_errors)
      for (int j = 0; j < k; j++) col_errors[cperm[j]] = ce[j];
  }
  if (stats) *stats = s;
  return GD_OK;
}

}  // extern "C"
