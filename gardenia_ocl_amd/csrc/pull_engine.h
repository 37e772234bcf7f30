// pull_engine.h -- device building blocks of the CSR pull engine (wave64).
//
// A workgroup of kBlock threads owns one Tile.  Latency is paid once per tile:
// every load the tile needs (edge indices, row offsets, the epilogue's
// per-row operands) is issued unconditionally up front -- addresses of idle
// lanes are clamped into the tile instead of branching -- and only then does
// the tile wait, gather (IPT independent gathers in flight per thread), stage
// the gathered values in LDS and reduce rows:
//   * rows of <= kSeqRow edges: the owning thread, sequentially in CSR order
//     (the reference's summation order, so the result is bit-identical);
//   * longer rows: one wavefront, lane-strided partials + xor butterfly; the
//     sum goes back through LDS to the owning thread for the epilogue.
// Hub segments gather kBlock*IPT edges of one long row into a partial sum.
//
// Row r of the tile is owned by thread (r - r0) % kBlock, so per-row reads and
// writes are coalesced.
#ifndef GD_PULL_ENGINE_H_
#define GD_PULL_ENGINE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gd_common.h"

namespace gd {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;
constexpr int kTileRows = 1024;
constexpr int kRowsPerThread = kTileRows / kBlock;
constexpr int kOffLoads = (kTileRows + 1 + kBlock - 1) / kBlock;
constexpr int kSeqRow = 64;

// Loads / stores of arrays streamed once per launch.  With NT the accesses
// carry the non-temporal hint, so the 1 GB-per-iteration index stream does not
// evict the cache-resident part of the gathered vector.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Sum of one value per thread over a workgroup of NWAVES waves, fixed order;
// result valid in thread 0.  `s_tmp` holds NWAVES values.
template <int NWAVES, typename T>
__device__ __forceinline__ T block_sum(T v, T* s_tmp) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) s_tmp[wid] = v;
  __syncthreads();
  T t = T(0);
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < NWAVES; w++) t += s_tmp[w];
  return t;
}

// s + v[0] + ... + v[n-1] in order with FP32 rounding, by the calling lane,
// from 16-byte aligned LDS: 4*kQ values per step, the next step's LDS reads in
// flight while the current step is added (the adds are one dependent chain).
template <int kQ>  // float4 reads per step
__device__ __forceinline__ float lane_seq_sum(float s, const float* __restrict__ v, int n) {
  const float4* __restrict__ v4 = reinterpret_cast<const float4*>(v);
  const int nb = n / (4 * kQ);
  float4 A[kQ], B[kQ];
  auto load = [&](float4(&R)[kQ], int blk) {
#pragma unroll
    for (int k = 0; k < kQ; k++) R[k] = v4[blk * kQ + k];
  };
  auto add = [&](const float4(&R)[kQ]) {
#pragma unroll
    for (int k = 0; k < kQ; k++) {
      s += R[k].x;
      s += R[k].y;
      s += R[k].z;
      s += R[k].w;
    }
  };
  if (nb > 0) {
    load(A, 0);
    int b = 0;
    // scheduling barriers keep each block's reads issued before the previous
    // block's adds, so the adds wait only for reads issued one block earlier
    for (; b + 2 <= nb; b += 2) {
      load(B, b + 1);
      __builtin_amdgcn_sched_barrier(0);
      add(A);
      __builtin_amdgcn_sched_barrier(0);
      load(A, min(b + 2, nb - 1));  // unconditional (clamped): A and B keep their registers
      __builtin_amdgcn_sched_barrier(0);
      add(B);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (b < nb) add(A);
  }
  for (int i = nb * 4 * kQ; i < n; i++) s += v[i];
  return s;
}

// Sequential sum s + v[b] + ... + v[e-1] from LDS (`v` 16-byte aligned) by
// the calling lane, in order with FP32 rounding.  More than kChainRow terms:
// scalar adds up to a 16-byte boundary, then float4 reads 16 terms ahead of
// the dependent adds (lane_seq_sum); fewer: 4 reads in flight.
constexpr int kChainRow = 32;

__device__ __forceinline__ float seq_sum(float s, const float* __restrict__ v, int b, int e) {
  int k = b;
  if (e - b > kChainRow) {
    for (; k & 3; k++) s += v[k];
    return lane_seq_sum<4>(s, v + k, e - k);
  }
  for (; k + 4 <= e; k += 4) {
    const float a0 = v[k], a1 = v[k + 1], a2 = v[k + 2], a3 = v[k + 3];
    s += a0;
    s += a1;
    s += a2;
    s += a3;
  }
  for (; k < e; k++) s += v[k];
  return s;
}

// Loads the tile's edge indices col[e0 + tid + j*kBlock] (clamped into the
// tile so all IPT loads issue back to back).  Requires ne >= 1.
template <int IPT, bool NT>
__device__ __forceinline__ void load_indices(const int32_t* __restrict__ col, int e0, int ne,
                                             int (&idx)[IPT]) {
  const int last = ne - 1;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    idx[j] = ld_stream<NT>(col + e0 + min((int)threadIdx.x + j * kBlock, last));
}

// Tiles whose edges are stored sorted by source (SORTED) carry, per edge, its
// offset in the tile: pos[] (16 bits).  The gather functor always sees the
// edge's CSR index.
template <int IPT, bool NT>
__device__ __forceinline__ void load_positions(const uint16_t* __restrict__ pos, int e0, int ne,
                                               int (&p)[IPT]) {
  const int last = ne - 1;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    p[j] = (int)ld_stream<NT>(pos + e0 + min((int)threadIdx.x + j * kBlock, last));
}

template <int IPT, bool SORTED, class G>
__device__ __forceinline__ void gather_values(const G& g, int e0, int ne, const int (&idx)[IPT],
                                              const int (&p)[IPT], float (&v)[IPT]) {
  const int last = ne - 1;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    v[j] = g.value(e0 + (SORTED ? p[j] : min((int)threadIdx.x + j * kBlock, last)), idx[j]);
}

// Sum of the gathered values of a hub segment [e0, e0+ne), ne >= 1: per-thread
// partial in j order, wave butterfly, then waves in order.  Valid in thread 0.
template <int IPT, bool NT, bool SORTED, class G>
__device__ __forceinline__ float segment_sum(const int32_t* __restrict__ col,
                                             const uint16_t* __restrict__ pos, int e0, int ne,
                                             const G& g, float* s_tmp) {
  int idx[IPT], p[IPT];
  float v[IPT];
  load_indices<IPT, NT>(col, e0, ne, idx);
  if constexpr (SORTED) load_positions<IPT, NT>(pos, e0, ne, p);
  gather_values<IPT, SORTED>(g, e0, ne, idx, p, v);
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < IPT; j++) s += ((int)threadIdx.x + j * kBlock < ne) ? v[j] : 0.0f;
  return block_sum<kWaves>(s, s_tmp);
}

// LDS of one row tile.
template <int IPT>
struct RowTileLds {
  alignas(16) float val[kBlock * IPT];  // gathered per-edge values
  int off[kTileRows + 1];     // row offsets, rebased to the tile on use
  float longsum[kTileRows];   // wavefront sums of rows longer than kSeqRow
};

// Registers holding the loads of one tile that do not depend on a gather:
// its edge indices, and for a row tile its row offsets and each owned row's
// epilogue operands.  Issued as early as possible (see pr_pull_persistent).
template <int IPT, class Pre>
struct TileRegs {
  int idx[IPT];
  int pos[IPT];  // sorted tiles only
  int offv[kOffLoads];
  Pre pre[kRowsPerThread];
};

// Phase A: issue every gather-independent load of tile t, without branches
// (a branch here would make the compiler's wait counting conservative and
// serialize the prefetch with the gathers).  Addresses of idle lanes are
// clamped: an empty tile re-reads its first edge, a hub segment reads the
// offsets and operands of its own row.
// `e_max` is the last valid index of `col` (the arrays hold >= 1 element).
template <int IPT, bool NT, bool SORTED, class E>
__device__ __forceinline__ void issue_tile_loads(const Tile& t, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ col,
                                                 const uint16_t* __restrict__ pos, int e_max,
                                                 const E& e, TileRegs<IPT, typename E::Pre>& R) {
  const int tid = threadIdx.x;
  const int last = max(t.w - t.z - 1, 0);
#pragma unroll
  for (int j = 0; j < IPT; j++)
    R.idx[j] = ld_stream<NT>(col + min(t.z + min(tid + j * kBlock, last), e_max));
  if constexpr (SORTED) {
#pragma unroll
    for (int j = 0; j < IPT; j++)
      R.pos[j] = (int)ld_stream<NT>(pos + min(t.z + min(tid + j * kBlock, last), e_max));
  }
  const int nr = (t.y >= 0) ? t.y - t.x : 1;
#pragma unroll
  for (int q = 0; q < kOffLoads; q++)
    R.offv[q] = ld_stream<NT>(off + t.x + min(tid + q * kBlock, nr));
#pragma unroll
  for (int q = 0; q < kRowsPerThread; q++) R.pre[q] = e.pre(t.x + min(tid + q * kBlock, nr - 1));
}

// Phases C-E of a row tile whose gathered values are already in lds.val and
// rebased offsets in lds.off (after a barrier): reduce rows, epilogue.
// SEQ_ALL: every row, whatever its length, is summed by its owner thread in
// CSR order (no wavefront tree): bit-identical sums for any row the tile holds.
template <int IPT, class E, bool SEQ_ALL = false>
__device__ __forceinline__ double reduce_row_tile(int r0, int nr, const E& e,
                                                  const typename E::Pre (&pre)[kRowsPerThread],
                                                  RowTileLds<IPT>& lds) {
  const int tid = threadIdx.x;
  float sum[kRowsPerThread];
  bool have_long = false;
#pragma unroll
  for (int q = 0; q < kRowsPerThread; q++) {
    const int i = tid + q * kBlock;
    sum[q] = 0.0f;
    if (i < nr) {
      const int b = lds.off[i], en = lds.off[i + 1];
      if (SEQ_ALL || en - b <= kSeqRow)
        sum[q] = seq_sum(e.init(pre[q]), lds.val, b, en);
      else
        have_long = true;
    }
  }
  if (!SEQ_ALL && __syncthreads_or(have_long)) {
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    for (int c = wid * kWave; c < nr; c += kBlock) {
      const int i = c + lane;
      const bool lr = (i < nr) && (lds.off[i + 1] - lds.off[i] > kSeqRow);
      uint64_t mask = __ballot(lr);
      while (mask) {
        const int row = c + __builtin_ctzll(mask);
        mask &= mask - 1;
        const int b = lds.off[row], en = lds.off[row + 1];
        float s = 0.0f;
        int k = b + lane;
        for (; k + 3 * kWave < en; k += 4 * kWave) {
          const float a0 = lds.val[k], a1 = lds.val[k + kWave], a2 = lds.val[k + 2 * kWave],
                      a3 = lds.val[k + 3 * kWave];
          s += a0;
          s += a1;
          s += a2;
          s += a3;
        }
        for (; k < en; k += kWave) s += lds.val[k];
        s = wave_sum(s);
        if (lane == 0) lds.longsum[row] = s;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRowsPerThread; q++) {
      const int i = tid + q * kBlock;
      if (i < nr && lds.off[i + 1] - lds.off[i] > kSeqRow) sum[q] = e.init(pre[q]) + lds.longsum[i];
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kRowsPerThread; q++) {
    const int i = tid + q * kBlock;
    if (i < nr) acc += e.finish(r0 + i, sum[q], pre[q]);
  }
  return acc;
}

// Stage a row tile's gathered values and rebased offsets into LDS.
template <int IPT, bool SORTED, class Pre>
__device__ __forceinline__ void stage_row_tile(const Tile& t, const TileRegs<IPT, Pre>& R,
                                               const float (&v)[IPT], RowTileLds<IPT>& lds) {
  const int tid = threadIdx.x;
  const int nr = t.y - t.x, ne = t.w - t.z;
#pragma unroll
  for (int q = 0; q < kOffLoads; q++)
    if (tid + q * kBlock <= nr) lds.off[tid + q * kBlock] = R.offv[q] - t.z;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    if (tid + j * kBlock < ne) lds.val[SORTED ? R.pos[j] : tid + j * kBlock] = v[j];
}

// One row tile: rows [r0, r1) whose edges are [e0, e0+ne).  The epilogue E
// provides
//   typename E::Pre          per-row operands, loaded before the gather
//   Pre pre(int row)         (clamped row: always a valid index)
//   float init(const Pre&)   value the row sum starts from
//   double finish(int row, float sum, const Pre&)   writes the row, returns
//                            its contribution to the tile's double accumulator
template <int IPT, bool NT, bool SORTED, class G, class E, bool SEQ_ALL = false>
__device__ __forceinline__ double row_tile(const int32_t* __restrict__ off,
                                           const int32_t* __restrict__ col,
                                           const uint16_t* __restrict__ pos, int e_max, int r0,
                                           int r1, int e0, int ne, const G& g, const E& e,
                                           RowTileLds<IPT>& lds) {
  const Tile t{r0, r1, e0, e0 + ne};
  TileRegs<IPT, typename E::Pre> R;
  issue_tile_loads<IPT, NT, SORTED>(t, off, col, pos, e_max, e, R);
  float v[IPT];
  if (ne > 0) gather_values<IPT, SORTED>(g, e0, ne, R.idx, R.pos, v);
  stage_row_tile<IPT, SORTED>(t, R, v, lds);
  __syncthreads();
  return reduce_row_tile<IPT, E, SEQ_ALL>(r0, r1 - r0, e, R.pre, lds);
}

// ---------------------------------------------------------------------------
// Sequential-order sums of long rows.
//
// exact_seq_sum returns s0 + v[0] + v[1] + ... + v[n-1] rounded exactly as the
// sequential FP32 loop `s = s0; for (i) s += v[i];` rounds it (the reference's
// per-row loop, src/pr/omp_base.cc:34-37), computed by the whole workgroup:
//   While the running sum s lies in one binade, 2^k <= |s| < 2^(k+1) with
//   ulp u = 2^(k-23), and the exact s + v[i] stays strictly inside it, the
//   rounded step is s + q_i*u with q_i = v[i]/u rounded to the nearest
//   integer: s advances by an integer number of ulps.  A positive term whose
//   fraction is exactly 1/2 (a tie) rounds s to even: that step is the map
//   x -> (x + floor(q)) rounded up to even, and the workgroup scans the steps
//   as maps (TieMap: additions and round-to-even steps compose into maps of
//   the same form), so runs of equal terms that tie at every step -- many
//   sources with the same score and degree -- do not stop the scan.  s
//   advances over every step up to the first one that leaves the binade, is a
//   negative tie or inf / NaN, or starts from a subnormal / non-finite s;
//   that step is one ordinary FP32 addition, and the scan restarts after it.
//   From s == +0 the sum jumps to the first nonzero term.  For sums of
//   positive terms (PageRank) the restarts are the ~log2(n) binade crossings.
// ---------------------------------------------------------------------------

constexpr int kExactHead = 256;  // terms of a row added serially before the scan

struct ExactLds {
  int32_t wsum[kWaves];
  int32_t wtie[kWaves];
  int32_t wb[kWaves];
  int wmin[kWaves];
  float s;
  int tie;  // the step that stopped the scan was a positive tie
};

// Exclusive prefix over the workgroup (lanes, then waves) and the total, in
// wrapping 32-bit arithmetic.  Only prefixes of steps that all stay inside the
// binade are used, and those are exact (each below 2^25 in magnitude).
__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t& total, ExactLds& x) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  uint32_t inc = (uint32_t)v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)inc, o, kWave);
    if (lane >= o) inc += t;
  }
  if (lane == kWave - 1) x.wsum[wid] = (int32_t)inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    const uint32_t ws = (uint32_t)x.wsum[w];
    before += (w < wid) ? ws : 0u;
    tot += ws;
  }
  __syncthreads();
  total = (int32_t)tot;
  return (int32_t)(before + (inc - (uint32_t)v));
}

// A run of steps in integer ulps as one map of the running value x:
//   t == 0:  x -> x + a
//   t == 1:  x -> y + (y & 1) + b,  y = x + a
// A regular step q is (0, q, 0).  A positive term that is exactly half an ulp
// above an integer f rounds to even: x -> (x + f) rounded up to even, the map
// (1, f, 0).  After such a step the value is even, so what follows it only
// depends on constants: maps compose into maps of the same form, and a scan
// over them gives every prefix exactly (wrapping 32-bit arithmetic keeps the
// parity).  Negative ties stay irregular.
struct TieMap {
  uint32_t t, a, b;
};

__device__ __forceinline__ TieMap tie_compose(const TieMap& p, const TieMap& q) {  // p, then q
  if (!p.t) return TieMap{q.t, p.a + q.a, q.b};
  if (!q.t) return TieMap{1u, p.a, p.b + q.a};
  const uint32_t c = p.b + q.a;
  return TieMap{1u, p.a, c + (c & 1u) + q.b};
}

__device__ __forceinline__ uint32_t tie_apply(const TieMap& m, uint32_t x) {
  const uint32_t y = x + m.a;
  return m.t ? y + (y & 1u) + m.b : y;
}

// Exclusive prefix (lanes, then waves, in order) of the maps and their total.
__device__ __forceinline__ TieMap block_excl_scan_map(TieMap v, TieMap& total, ExactLds& x) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  TieMap inc = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    TieMap left;
    left.t = (uint32_t)__shfl_up((int)inc.t, o, kWave);
    left.a = (uint32_t)__shfl_up((int)inc.a, o, kWave);
    left.b = (uint32_t)__shfl_up((int)inc.b, o, kWave);
    if (lane >= o) inc = tie_compose(left, inc);
  }
  // exclusive within the wave: the inclusive value of the previous lane
  TieMap ex;
  ex.t = (uint32_t)__shfl_up((int)inc.t, 1, kWave);
  ex.a = (uint32_t)__shfl_up((int)inc.a, 1, kWave);
  ex.b = (uint32_t)__shfl_up((int)inc.b, 1, kWave);
  if (lane == 0) ex = TieMap{0u, 0u, 0u};
  if (lane == kWave - 1) {
    x.wtie[wid] = (int32_t)inc.t;
    x.wsum[wid] = (int32_t)inc.a;
    x.wb[wid] = (int32_t)inc.b;
  }
  __syncthreads();
  TieMap before{0u, 0u, 0u}, tot{0u, 0u, 0u};
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    const TieMap wm{(uint32_t)x.wtie[w], (uint32_t)x.wsum[w], (uint32_t)x.wb[w]};
    if (w < wid) before = tie_compose(before, wm);
    tot = tie_compose(tot, wm);
  }
  __syncthreads();
  total = tot;
  return tie_compose(before, ex);
}

__device__ __forceinline__ int block_min(int v, ExactLds& x) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
  if (lane == 0) x.wmin[wid] = v;
  __syncthreads();
  int r = x.wmin[0];
#pragma unroll
  for (int w = 1; w < kWaves; w++) r = min(r, x.wmin[w]);
  __syncthreads();
  return r;
}

// v in units of u = 2^(ex-23), rounded to the nearest integer, from the bits of
// v (|q| clamped to 2^26: such a step leaves the binade anyway).  tie: v is
// positive and its fraction is exactly 1/2 (the result holds the floor; the
// step rounds to even, map (1, q, 0)).  irregular: a negative tie, or v is inf
// or NaN.
__device__ __forceinline__ int32_t ulps_of_tie(float v, int ex, bool& tie, bool& irregular) {
  const uint32_t b = __float_as_uint(v);
  const int eb = (int)((b >> 23) & 0xffu);
  const uint32_t mv = eb ? ((b & 0x7fffffu) | 0x800000u) : (b & 0x7fffffu);  // v = mv*2^(ev-23)
  const int d = ex - (eb ? eb - 127 : -126);                                  // q = mv / 2^d
  int32_t q;
  bool half = false;
  if (d <= 0) {
    q = (d <= -3) ? (1 << 26) : (int32_t)min((uint32_t)1 << 26, mv << (-d));
  } else if (d >= 26) {
    q = 0;
  } else {
    const uint32_t hf = 1u << (d - 1), rem = mv & ((1u << d) - 1u);
    half = (rem == hf);
    q = (int32_t)((mv >> d) + (rem > hf ? 1u : 0u));
  }
  const bool neg = (b >> 31) != 0;
  tie = half && !neg;
  irregular = (half && neg) || eb == 0xff;
  return neg ? -q : q;
}

// Position of chunk element i in the padded LDS buffer: thread t reads its IPT
// consecutive elements at stride IPT+1 words (no bank conflicts).
template <int IPT>
__device__ __forceinline__ int exact_pos(int i) {
  return i + i / IPT;
}

// One chunk of a sequential-order sum: nc values already in LDS (val, padded
// layout exact_pos), added to s in index order with sequential FP32 rounding.
// `head`: add the first kExactHead values serially first (the start of a row,
// where s crosses a binade at almost every doubling).  All threads return the
// same s.  Ends with the workgroup synchronised (val may be rewritten).
// tie_mode (uniform, kept across the chunks of a row): once a positive tie has
// stopped the scan, the row is scanned as maps (TieMap), which go through
// ties; before that, the plain integer scan (cheaper) treats a tie as a stop.
template <int IPT>
__device__ float exact_scan_chunk(const float* __restrict__ val, int nc, bool head, float s,
                                  ExactLds& x, bool& tie_mode) {
  const int tid = threadIdx.x;
  const int lo = tid * IPT, hi = min(lo + IPT, nc);  // this thread's elements
  int start = 0;
  if (head) {
    const int ns = min(nc, kExactHead);
    if (tid == 0) {
      // software-pipelined: the next batch's LDS loads are in flight while the
      // current batch is added (the adds are one dependent chain)
      constexpr int B = 16;
      float a = s;
      int i = 0;
      if (ns >= B) {
        float t[B];
#pragma unroll
        for (int k = 0; k < B; k++) t[k] = val[exact_pos<IPT>(k)];
        for (i = B; i + B <= ns; i += B) {
          float u[B];
#pragma unroll
          for (int k = 0; k < B; k++) u[k] = val[exact_pos<IPT>(i + k)];
#pragma unroll
          for (int k = 0; k < B; k++) a += t[k];
#pragma unroll
          for (int k = 0; k < B; k++) t[k] = u[k];
        }
#pragma unroll
        for (int k = 0; k < B; k++) a += t[k];
      }
      for (; i < ns; i++) a += val[exact_pos<IPT>(i)];
      x.s = a;
    }
    __syncthreads();
    s = x.s;
    __syncthreads();
    start = ns;
  }
  while (start < nc) {  // s and start are uniform across the workgroup
    if (!(fabsf(s) >= 0x1p-125f) || !isfinite(s)) {  // zero, subnormal, inf/NaN
      if (s == 0.0f && !signbit(s)) {
        // +0 + (+-0) == +0: jump to the first nonzero term (x != 0: +0 + x == x)
        int nz = 0x7fffffff;
        for (int i = max(lo, start); i < hi; i++)
          if (val[exact_pos<IPT>(i)] != 0.0f) {
            nz = i;
            break;
          }
        const int jz = block_min(nz, x);
        if (jz == 0x7fffffff) {
          start = nc;
        } else {
          s = val[exact_pos<IPT>(jz)];
          start = jz + 1;
        }
        continue;
      }
      s = s + val[exact_pos<IPT>(start)];
      start++;
      continue;
    }
    const int ex = ilogbf(s);
    // s = P * 2^(ex-23) with 2^23 <= |P| < 2^24
    const int32_t P = (int32_t)(__float_as_uint(s) & 0x7fffffu) | 0x800000;
    const int32_t Ps = signbit(s) ? -P : P;
    int32_t run, end;  // the value before this thread's first step; after the last step
    if (tie_mode) {
      TieMap L{0u, 0u, 0u};
      for (int i = max(lo, start); i < hi; i++) {
        bool tie, irr;
        const uint32_t q = (uint32_t)ulps_of_tie(val[exact_pos<IPT>(i)], ex, tie, irr);
        L = tie_compose(L, TieMap{tie ? 1u : 0u, q, 0u});
      }
      TieMap total;
      run = (int32_t)tie_apply(block_excl_scan_map(L, total, x), (uint32_t)Ps);
      end = (int32_t)tie_apply(total, (uint32_t)Ps);
    } else {
      uint32_t L = 0;
      for (int i = max(lo, start); i < hi; i++) {
        bool tie, irr;
        L += (uint32_t)ulps_of_tie(val[exact_pos<IPT>(i)], ex, tie, irr);
      }
      int32_t total;
      run = (int32_t)((uint32_t)Ps + (uint32_t)block_excl_scan((int32_t)L, total, x));
      end = (int32_t)((uint32_t)Ps + (uint32_t)total);
    }
    int bad = 0x7fffffff;
    bool bad_tie = false;
    for (int i = max(lo, start); i < hi; i++) {
      bool tie, irr;
      const uint32_t q = (uint32_t)ulps_of_tie(val[exact_pos<IPT>(i)], ex, tie, irr);
      const uint32_t y = (uint32_t)run + q;
      const int32_t r = (int32_t)(tie ? y + (y & 1u) : y);
      if (irr || (tie && !tie_mode) || !(abs(r) > (1 << 23) && abs(r) < (1 << 24))) {
        bad = i;
        bad_tie = tie && !tie_mode;
        break;
      }
      run = r;
    }
    const int jb = block_min(bad, x);
    if (jb == 0x7fffffff) {
      // exact: an integer below 2^24 in magnitude times a power of two
      s = ldexpf((float)end, ex - 23);
      start = nc;
    } else {
      if (bad == jb) {
        x.s = ldexpf((float)run, ex - 23) + val[exact_pos<IPT>(jb)];  // FP32 step
        x.tie = bad_tie ? 1 : 0;
      }
      __syncthreads();
      s = x.s;
      tie_mode = tie_mode || x.tie != 0;
      __syncthreads();
      start = jb + 1;
    }
  }
  __syncthreads();
  return s;
}

// Sum of the gathered values of edges [e0, e0+ne) in edge order, rounded as the
// sequential FP32 loop starting from s.  All threads return the same value.
// `val` holds kBlock*IPT + kBlock floats (the padded layout exact_pos).
template <int IPT, bool NT, bool SORTED, class G>
__device__ __forceinline__ float exact_seq_sum(const int32_t* __restrict__ col, const uint16_t* __restrict__ pos,
                               int e_max, int e0, int ne, const G& g, float s,
                               float* __restrict__ val, ExactLds& x) {
  constexpr int C = kBlock * IPT;
  const int tid = threadIdx.x;
  bool tie_mode = false;
  for (int c0 = 0; c0 < ne; c0 += C) {
    const int nc = min(C, ne - c0);
    int idx[IPT], p[IPT];
    float v[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++)
      idx[j] = ld_stream<NT>(col + min(e0 + c0 + min(tid + j * kBlock, nc - 1), e_max));
    if constexpr (SORTED) {  // chunks are sorted separately: offsets are chunk-relative
#pragma unroll
      for (int j = 0; j < IPT; j++)
        p[j] = (int)ld_stream<NT>(pos + min(e0 + c0 + min(tid + j * kBlock, nc - 1), e_max));
    }
#pragma unroll
    for (int j = 0; j < IPT; j++)
      v[j] = g.value(e0 + c0 + (SORTED ? p[j] : min(tid + j * kBlock, nc - 1)), idx[j]);
#pragma unroll
    for (int j = 0; j < IPT; j++)
      if (tid + j * kBlock < nc) val[exact_pos<IPT>(SORTED ? p[j] : tid + j * kBlock)] = v[j];
    __syncthreads();
    s = exact_scan_chunk<IPT>(val, nc, c0 == 0, s, x, tie_mode);
  }
  return s;
}

// ---------------------------------------------------------------------------
// Long rows summed in edge order by one lane.
//
// The reference adds a row's terms one after the other in FP32
// (src/pr/omp_base.cc:34-37, src/spmv/omp_base.cc:21-26); for a row longer
// than a tile the fastest way to the same bits is that dependent chain
// itself, on one lane, with the rest of the workgroup keeping it fed
// (hub_chain_sum).  Measured on the MI355X the chain costs 2.8-3.4 ns per
// term (a dependent v_add_f32 every ~4.8 cycles when the terms sit in
// registers; the LDS reads feeding it take the rest), independent of the
// terms' signs, ties or magnitudes.

// s0 + the gathered values of edges [e0, e0+ne) in edge order, rounded as the
// sequential FP32 loop, with the workgroup split by role: lane 0 of wave 0
// adds chunk c (kHubChunk entries, LDS) while waves 1..3 gather chunk c+1
// into the other buffer (their column loads of chunk c+2 already in flight),
// so no gather waits on the chain and the chain wave issues nothing else.
// SORTED: each chunk is stored sorted by source and pos[e] is the edge's
// offset within its chunk; values go back to edge order in LDS.  Buffers:
// two 16-byte aligned arrays of kHubChunk floats, one float for the sum.
// All threads return the sum.
template <int kQ, bool SORTED, class G>
__device__ float hub_chain_sum(const int32_t* __restrict__ col, const uint16_t* __restrict__ pos,
                               int e_max, int e0, int ne, const G& g, float s0,
                               float* __restrict__ buf0, float* __restrict__ buf1,
                               float* __restrict__ s_out) {
  constexpr int kProd = kBlock - kWave;  // gathering lanes
  constexpr int PPT = kHubChunk / kProd;
  static_assert(PPT * kProd == kHubChunk && kHubChunk % (4 * kQ) == 0, "hub chunk shape");
  const int tid = threadIdx.x;
  const bool chain_wave = tid < kWave;
  const int pt = tid - kWave;
  const int nchunks = (ne + kHubChunk - 1) / kHubChunk;
  int idx[PPT], p[PPT];
  auto slot = [&](int c, int j) {  // clamped into the chunk
    return min(pt + j * kProd, min(kHubChunk, ne - c * kHubChunk) - 1);
  };
  auto load_idx = [&](int c) {
#pragma unroll
    for (int j = 0; j < PPT; j++) idx[j] = col[min(e0 + c * kHubChunk + slot(c, j), e_max)];
    if constexpr (SORTED) {
#pragma unroll
      for (int j = 0; j < PPT; j++) p[j] = (int)pos[min(e0 + c * kHubChunk + slot(c, j), e_max)];
    }
  };
  auto produce = [&](int c, float* buf) {  // gathers chunk c (indices loaded), stores it
    const int nc = min(kHubChunk, ne - c * kHubChunk);
    float v[PPT];
    int o[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
      o[j] = SORTED ? p[j] : slot(c, j);
      v[j] = g.value(e0 + c * kHubChunk + o[j], idx[j]);
    }
#pragma unroll
    for (int j = 0; j < PPT; j++)
      if (pt + j * kProd < nc) buf[o[j]] = v[j];
  };
  if (chain_wave) __builtin_amdgcn_s_setprio(3);
  if (!chain_wave && nchunks > 0) {
    load_idx(0);
    produce(0, buf0);
    if (nchunks > 1) load_idx(1);
  }
  __syncthreads();
  float s = s0;
  for (int c = 0; c < nchunks; c++) {
    float* cur = (c & 1) ? buf1 : buf0;
    float* nxt = (c & 1) ? buf0 : buf1;
    if (chain_wave) {
      if (tid == 0) s = lane_seq_sum<kQ>(s, cur, min(kHubChunk, ne - c * kHubChunk));
    } else if (c + 1 < nchunks) {
      produce(c + 1, nxt);
      if (c + 2 < nchunks) load_idx(c + 2);  // in flight across the barrier
    }
    __syncthreads();
  }
  if (chain_wave) __builtin_amdgcn_s_setprio(0);
  if (tid == 0) *s_out = s;
  __syncthreads();
  return *s_out;
}

}  // namespace gd

#endif  // GD_PULL_ENGINE_H_
