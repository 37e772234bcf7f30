// pull_engine.h -- device building blocks of the CSR pull engine (wave64).
//
// A workgroup of NB threads (kBlock = 256 by default) owns one Tile of up to
// NB*IPT edges and 4*NB rows.  Latency is paid once per tile:
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
// Row r of the tile is owned by thread (r - r0) % NB, so per-row reads and
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
constexpr int kRowsPerThread = 4;  // rows per thread of a row tile (any block size)
constexpr int kTileRows = kRowsPerThread * kBlock;
constexpr int kOffLoads = kRowsPerThread + 1;  // ceil((rows + 1) / threads)
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
// kQ = 8 (reads 32 terms ahead) hides the LDS latency behind the chain: at
// RMAT-24 the pull phase took 1.339 ms against 1.367 ms with kQ = 4, at
// RMAT-16 32.8 against 35.5 µs per iteration; kQ = 12 was no faster and costs
// the row tiles 24 more VGPRs.
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
// scalar adds up to a 16-byte boundary, then float4 reads 32 terms ahead of
// the dependent adds (lane_seq_sum); fewer: 4 reads in flight.
constexpr int kChainRow = 32;

__device__ __forceinline__ float seq_sum(float s, const float* __restrict__ v, int b, int e) {
  int k = b;
  if (e - b > kChainRow) {
    for (; k & 3; k++) s += v[k];
    return lane_seq_sum<8>(s, v + k, e - k);
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
template <int IPT, bool NT, int NB = kBlock>
__device__ __forceinline__ void load_indices(const int32_t* __restrict__ col, int e0, int ne,
                                             int (&idx)[IPT]) {
  const int last = ne - 1;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    idx[j] = ld_stream<NT>(col + e0 + min((int)threadIdx.x + j * NB, last));
}

// Tiles whose edges are stored sorted by source (SORTED) carry, per edge, its
// offset in the tile: pos[] (16 bits).  The gather functor always sees the
// edge's CSR index.
template <int IPT, bool NT, int NB = kBlock>
__device__ __forceinline__ void load_positions(const uint16_t* __restrict__ pos, int e0, int ne,
                                               int (&p)[IPT]) {
  const int last = ne - 1;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    p[j] = (int)ld_stream<NT>(pos + e0 + min((int)threadIdx.x + j * NB, last));
}

template <int IPT, bool SORTED, class G, int NB = kBlock>
__device__ __forceinline__ void gather_values(const G& g, int e0, int ne, const int (&idx)[IPT],
                                              const int (&p)[IPT], float (&v)[IPT]) {
  const int last = ne - 1;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    v[j] = g.value(e0 + (SORTED ? p[j] : min((int)threadIdx.x + j * NB, last)), idx[j]);
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

// LDS of one row tile of a workgroup of NB threads.
template <int IPT, int NB = kBlock>
struct RowTileLds {
  alignas(16) float val[NB * IPT];           // gathered per-edge values
  int off[kRowsPerThread * NB + 1];          // row offsets, rebased to the tile on use
  float longsum[kRowsPerThread * NB];        // wavefront sums of rows longer than kSeqRow
};

// Registers holding the loads of one tile that do not depend on a gather:
// its edge indices, and for a row tile its row offsets and each owned row's
// epilogue operands.  Issued as early as possible (see pr_pull_persistent).
template <int IPT, class Pre, int NB = kBlock>
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
template <int IPT, bool NT, bool SORTED, class E, int NB = kBlock>
__device__ __forceinline__ void issue_tile_loads(const Tile& t, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ col,
                                                 const uint16_t* __restrict__ pos, int e_max,
                                                 const E& e, TileRegs<IPT, typename E::Pre, NB>& R) {
  const int tid = threadIdx.x;
  const int last = max(t.w - t.z - 1, 0);
#pragma unroll
  for (int j = 0; j < IPT; j++)
    R.idx[j] = ld_stream<NT>(col + min(t.z + min(tid + j * NB, last), e_max));
  if constexpr (SORTED) {
#pragma unroll
    for (int j = 0; j < IPT; j++)
      R.pos[j] = (int)ld_stream<NT>(pos + min(t.z + min(tid + j * NB, last), e_max));
  }
  const int nr = (t.y >= 0) ? t.y - t.x : 1;
#pragma unroll
  for (int q = 0; q < kOffLoads; q++)
    R.offv[q] = ld_stream<NT>(off + t.x + min(tid + q * NB, nr));
#pragma unroll
  for (int q = 0; q < kRowsPerThread; q++) R.pre[q] = e.pre(t.x + min(tid + q * NB, nr - 1));
}

// Phases C-E of a row tile whose gathered values are already in lds.val and
// rebased offsets in lds.off (after a barrier): reduce rows, epilogue.
// SEQ_ALL: every row, whatever its length, is summed by its owner thread in
// CSR order (no wavefront tree): bit-identical sums for any row the tile holds.
template <int IPT, class E, bool SEQ_ALL = false, int NB = kBlock>
__device__ __forceinline__ double reduce_row_tile(int r0, int nr, const E& e,
                                                  const typename E::Pre (&pre)[kRowsPerThread],
                                                  RowTileLds<IPT, NB>& lds) {
  const int tid = threadIdx.x;
  float sum[kRowsPerThread];
  bool have_long = false;
#pragma unroll
  for (int q = 0; q < kRowsPerThread; q++) {
    const int i = tid + q * NB;
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
    for (int c = wid * kWave; c < nr; c += NB) {
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
      const int i = tid + q * NB;
      if (i < nr && lds.off[i + 1] - lds.off[i] > kSeqRow) sum[q] = e.init(pre[q]) + lds.longsum[i];
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < kRowsPerThread; q++) {
    const int i = tid + q * NB;
    if (i < nr) acc += e.finish(r0 + i, sum[q], pre[q]);
  }
  return acc;
}

// Stage a row tile's gathered values and rebased offsets into LDS.
template <int IPT, bool SORTED, class Pre, int NB = kBlock>
__device__ __forceinline__ void stage_row_tile(const Tile& t, const TileRegs<IPT, Pre, NB>& R,
                                               const float (&v)[IPT], RowTileLds<IPT, NB>& lds) {
  const int tid = threadIdx.x;
  const int nr = t.y - t.x, ne = t.w - t.z;
#pragma unroll
  for (int q = 0; q < kOffLoads; q++)
    if (tid + q * NB <= nr) lds.off[tid + q * NB] = R.offv[q] - t.z;
#pragma unroll
  for (int j = 0; j < IPT; j++)
    if (tid + j * NB < ne) lds.val[SORTED ? R.pos[j] : tid + j * NB] = v[j];
}

// One row tile: rows [r0, r1) whose edges are [e0, e0+ne).  The epilogue E
// provides
//   typename E::Pre          per-row operands, loaded before the gather
//   Pre pre(int row)         (clamped row: always a valid index)
//   float init(const Pre&)   value the row sum starts from
//   double finish(int row, float sum, const Pre&)   writes the row, returns
//                            its contribution to the tile's double accumulator
template <int IPT, bool NT, bool SORTED, class G, class E, bool SEQ_ALL = false,
          int NB = kBlock>
__device__ __forceinline__ double row_tile(const int32_t* __restrict__ off,
                                           const int32_t* __restrict__ col,
                                           const uint16_t* __restrict__ pos, int e_max, int r0,
                                           int r1, int e0, int ne, const G& g, const E& e,
                                           RowTileLds<IPT, NB>& lds) {
  const Tile t{r0, r1, e0, e0 + ne};
  TileRegs<IPT, typename E::Pre, NB> R;
  issue_tile_loads<IPT, NT, SORTED, E, NB>(t, off, col, pos, e_max, e, R);
  float v[IPT];
  if (ne > 0) gather_values<IPT, SORTED, G, NB>(g, e0, ne, R.idx, R.pos, v);
  stage_row_tile<IPT, SORTED, typename E::Pre, NB>(t, R, v, lds);
  __syncthreads();
  return reduce_row_tile<IPT, E, SEQ_ALL, NB>(r0, r1 - r0, e, R.pre, lds);
}

// ---------------------------------------------------------------------------
// Sequential-order sums of long rows by a workgroup.
//
// exact_row_sum returns s0 + v[0] + v[1] + ... + v[n-1] rounded exactly as
// the sequential FP32 loop `s = s0; for (i) s += v[i];` rounds it (the
// reference's per-row loop, src/pr/omp_base.cc:34-37), computed by the whole
// workgroup, kScanChunk terms at a time (thread t owns terms [16t, 16t+16)):
//   While the running sum s lies in one binade, 2^k <= |s| < 2^(k+1) with
//   ulp u = 2^(k-23), and the exact s + v[i] stays strictly inside it, the
//   rounded step is s + q_i*u with q_i = v[i]/u rounded to the nearest
//   integer: s advances by an integer number of ulps.  A positive term whose
//   fraction is exactly 1/2 (a tie) rounds s to even: that step is the map
//   x -> (x + floor(q)) rounded up to even.  Steps are maps of the running
//   value that compose into maps of the same form (TieMap), so one workgroup
//   prefix over the maps gives every running value of the chunk, ties
//   included.  s advances over every step up to the first one that leaves
//   the binade, is a negative tie or inf / NaN; that step is one ordinary
//   FP32 addition and the scan restarts after it in the new binade.  From
//   s == +0 the sum jumps to the first nonzero term; from a subnormal, -0 or
//   non-finite s it takes one ordinary addition.  For sums of positive terms
//   (PageRank) the restarts are the ~log2(n) binade crossings, most of them
//   in the first terms of a row, which one lane adds serially (kExactHead).
//   The next chunk's terms are gathered into registers while a chunk is
//   scanned, so the gathers' latency overlaps the scan.
// A host emulation of these steps (integer and float forms of q) matched the
// sequential loop bit for bit on rows built to hit ties, binade crossings,
// zeros, signs, subnormals and wide ranges.
// ---------------------------------------------------------------------------

constexpr int kScanPT = 16;                   // terms per thread per chunk
constexpr int kScanChunk = kBlock * kScanPT;  // 4096
constexpr int kExactHead = kScanChunk;       // terms of a row added by one lane first
constexpr int kNoBad = 0x7fffffff;

// A run of steps in integer ulps as one map of the running value x:
//   t == 0:  x -> x + a
//   t == 1:  x -> y + (y & 1) + b,  y = x + a
// A regular step q is (0, q, 0); a positive tie above floor f is (1, f, 0).
// After a round-to-even step the value is even, so what follows it only
// depends on constants: maps compose into maps of the same form (wrapping
// 32-bit arithmetic keeps the parity).
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

// The step of term v (sign bit `neg`) in ulps of the binade with exponent ex:
// T = v / u exactly (a power-of-two scaling; inf when v is far above the
// binade), rounded to the nearest integer, |q| clamped to 2^26 (such a step
// leaves the binade anyway).  irregular: a negative tie, or T not finite.
// The sign comes in as an integer so that no per-term comparison of v is
// loop-invariant (hoisted lane masks would crowd the scalar registers).
__device__ __forceinline__ TieMap step_map(float v, uint32_t neg, int ex, uint32_t& irregular) {
  const float T = ldexpf(v, 23 - ex);
  const float fl = floorf(T);
  const float fr = T - fl;  // exact
  const float qf = fminf(fmaxf(fr > 0.5f ? fl + 1.0f : fl, -0x1p26f), 0x1p26f);
  const uint32_t half = fr == 0.5f;  // T == +-0 gives fr == 0: no tie
  irregular = (half & neg) | (uint32_t)!isfinite(T);
  return TieMap{half & (neg ^ 1u), (uint32_t)(int32_t)qf, 0u};
}

// Inclusive prefix of the maps over the 64 lanes of a wavefront: DPP row
// shifts inside each row of 16 lanes, then row broadcasts across rows (lanes
// whose source is outside the row read the identity map (0, 0, 0)).
#define GD_MAP_SCAN_STEP(CTRL, ROWS)                                                          \
  {                                                                                           \
    TieMap l;                                                                                 \
    l.t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)inc.t, CTRL, ROWS, 0xf, false);       \
    l.a = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)inc.a, CTRL, ROWS, 0xf, false);       \
    l.b = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)inc.b, CTRL, ROWS, 0xf, false);       \
    inc = tie_compose(l, inc);                                                                \
  }

__device__ __forceinline__ TieMap wave_incl_scan_map(TieMap inc) {
  GD_MAP_SCAN_STEP(0x111, 0xf)  // row_shr:1
  GD_MAP_SCAN_STEP(0x112, 0xf)  // row_shr:2
  GD_MAP_SCAN_STEP(0x114, 0xf)  // row_shr:4
  GD_MAP_SCAN_STEP(0x118, 0xf)  // row_shr:8
  GD_MAP_SCAN_STEP(0x142, 0xa)  // row_bcast:15 -> rows 1, 3
  GD_MAP_SCAN_STEP(0x143, 0xc)  // row_bcast:31 -> rows 2, 3
  return inc;
}
#undef GD_MAP_SCAN_STEP

struct ExactLds {
  uint32_t wt[kWaves], wa[kWaves], wb[kWaves];  // wave totals of the maps
  int wbad[kWaves];                             // wave's first irregular term
  float s;
};

// Workgroup-uniform values kept in scalar registers: branches on them are
// scalar branches, not exec-mask updates.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// First value over the workgroup of a per-thread int that increases with the
// thread index where present (kNoBad elsewhere): ballot per wave, then waves
// in order.  Uniform result.  The cross-lane read runs with the whole wave
// active: inside `if (lane == 0)` the compiler may compute v in lane 0 only
// (it is used nowhere else), and the read would return another lane's stale
// register.
__device__ __forceinline__ int block_first(int v, ExactLds& x) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const uint64_t m = __ballot(v != kNoBad);
  const int first = m ? __builtin_amdgcn_readlane(v, __builtin_ctzll(m)) : kNoBad;
  if (lane == 0) x.wbad[wid] = first;
  __syncthreads();
  int r = kNoBad;
#pragma unroll
  for (int w = kWaves - 1; w >= 0; w--) r = (x.wbad[w] != kNoBad) ? x.wbad[w] : r;
  return uniform(r);
}

// The scan of one chunk: nc terms in LDS (`val`, thread t's terms at 16t),
// added to s from term `start` on.  All threads return the same s.  Ends with
// the workgroup synchronised.  A thread's terms in [start, nc) are the set
// bits of its mask `act`; the per-term checks are branch-free selects.
__device__ __forceinline__ float exact_scan_chunk(const float* __restrict__ val, int nc,
                                                  int start, float s, ExactLds& x) {
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), wid = tid / kWave;
  const int lo = tid * kScanPT;
  float v[kScanPT];
  {
    const float4* v4 = reinterpret_cast<const float4*>(val + lo);
#pragma unroll
    for (int k = 0; k < kScanPT / 4; k++) {
      const float4 t = v4[k];
      v[4 * k] = t.x;
      v[4 * k + 1] = t.y;
      v[4 * k + 2] = t.z;
      v[4 * k + 3] = t.w;
    }
  }
  const uint32_t below_nc = (1u << min(max(nc - lo, 0), kScanPT)) - 1u;
  uint32_t negs = 0u, nonzero = 0u;  // per-term bits, computed once
#pragma unroll
  for (int k = 0; k < kScanPT; k++) {
    const uint32_t b = __float_as_uint(v[k]);
    negs |= (b >> 31) << k;
    nonzero |= (uint32_t)((b << 1) != 0u) << k;
  }
  start = uniform(start);
  s = uniform(s);
  while (start < nc) {
    const uint32_t sb = __float_as_uint(s);
    const int eb = (int)((sb >> 23) & 0xffu);
    if (eb <= 1 || eb == 0xff) {  // zero, subnormal, [2^-126, 2^-125), inf / NaN
      if (sb == 0u) {
        // +0 + (+-0) == +0: jump to the first nonzero term (+0 + v == v)
        const uint32_t cand =
            nonzero & below_nc & ~((1u << min(max(start - lo, 0), kScanPT)) - 1u);
        const int nz = cand ? lo + __builtin_ctz(cand) : kNoBad;
        const int jz = block_first(nz, x);
        if (jz == kNoBad) {
          start = nc;
        } else {
          s = uniform(val[jz]);
          start = max(jz, start) + 1;  // every pass advances: the loop ends
        }
        __syncthreads();  // wbad is reused
        continue;
      }
      s = uniform(s + val[start]);
      start++;
      continue;
    }
    const int ex = eb - 127;
    const int32_t P = (int32_t)(sb & 0x7fffffu) | 0x800000;
    const uint32_t Ps = (uint32_t)((sb >> 31) ? -P : P);
    const uint32_t act = below_nc & ~((1u << min(max(start - lo, 0), kScanPT)) - 1u);
    // this thread's steps (identity outside act) and their composition
    uint32_t q[kScanPT];
    uint32_t ties = 0u, irrs = 0u;
    TieMap L{0u, 0u, 0u};
    if (act) {  // whole waves before a restart skip this
#pragma unroll
      for (int k = 0; k < kScanPT; k++) {
        const uint32_t a = (act >> k) & 1u;  // integer masks: no lane-mask registers
        uint32_t irr;
        const TieMap mk = step_map(v[k], (negs >> k) & 1u, ex, irr);
        q[k] = mk.a & (0u - a);
        const uint32_t tk = mk.t & a;
        ties |= tk << k;
        irrs |= (irr & a) << k;
        L = tie_compose(L, TieMap{tk, q[k], 0u});
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanPT; k++) q[k] = 0u;
    }
    // workgroup prefix of the maps: waves in order, lanes in order
    const TieMap inc = wave_incl_scan_map(L);
    if (lane == kWave - 1) {
      x.wt[wid] = inc.t;
      x.wa[wid] = inc.a;
      x.wb[wid] = inc.b;
    }
    TieMap ex_lane;
    ex_lane.t = (uint32_t)__shfl_up((int)inc.t, 1, kWave);
    ex_lane.a = (uint32_t)__shfl_up((int)inc.a, 1, kWave);
    ex_lane.b = (uint32_t)__shfl_up((int)inc.b, 1, kWave);
    if (lane == 0) ex_lane = TieMap{0u, 0u, 0u};
    __syncthreads();
    TieMap before{0u, 0u, 0u}, total{0u, 0u, 0u};
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      const TieMap wm{x.wt[w], x.wa[w], x.wb[w]};
      if (w < wid) before = tie_compose(before, wm);
      total = tie_compose(total, wm);
    }
    // running values of this thread's terms, in order, up to the first one
    // whose step is irregular or leaves the binade: |r| in (2^23, 2^24), or
    // |r| == 2^23 (|s| == 2^k exactly) left unchanged by a zero term or one of
    // its own sign, whose exact sum cannot have come from below 2^k
    uint32_t run = tie_apply(tie_compose(before, ex_lane), Ps);
    int bad = kNoBad;
    uint32_t run_bad = 0u, alive = 1u;
#pragma unroll
    for (int k = 0; k < kScanPT; k++) {
      const uint32_t y = run + q[k];
      const uint32_t r = y + (y & (ties >> k) & 1u);  // a tie rounds to even
      const uint32_t ar = (uint32_t)abs((int32_t)r);
      const uint32_t away = ((nonzero >> k) ^ 1u) | ((negs >> k) ^ (run >> 31) ^ 1u);
      const uint32_t stay = (uint32_t)(r == run) & (uint32_t)(ar == 0x800000u) & away;
      const uint32_t ok = ((uint32_t)(ar - 0x800001u < 0x7fffffu) | stay) & ~(irrs >> k) & 1u;
      const uint32_t lv = (act >> k) & alive;
      const uint32_t fail = lv & (ok ^ 1u);
      run_bad = fail ? run : run_bad;
      bad = fail ? lo + k : bad;
      alive &= fail ^ 1u;
      run = (lv & ok) ? r : run;
    }
    const int jb = block_first(bad, x);
    if (jb == kNoBad) {
      // exact: an integer below 2^24 in magnitude times a power of two
      s = uniform(ldexpf((float)(int32_t)tie_apply(total, Ps), ex - 23));
      start = nc;
      __syncthreads();  // wt/wa/wb and wbad are reused
    } else {
      if (bad == jb) x.s = ldexpf((float)(int32_t)run_bad, ex - 23) + val[jb];  // FP32 step
      __syncthreads();
      s = uniform(x.s);
      start = max(jb, start) + 1;
      __syncthreads();
    }
  }
  return s;
}

// s0 + the gathered values of edges [e0, e0+ne) in edge order, rounded as the
// sequential FP32 loop starting from s0 (all threads return it).  `buf0`,
// `buf1`: two 16-byte aligned LDS buffers of kScanChunk floats.  Chunk c is
// gathered as thread t's terms t + 256 j (coalesced index loads), stored at
// their position in the chunk, and scanned as thread t's terms 16t..16t+15;
// chunk c+1's gathers and chunk c+2's index loads are in flight meanwhile.
// The first kExactHead terms are one lane's chain (lane_seq_sum): there the
// running sum crosses binades every few terms, and each crossing would cost
// the scan a restart.  STAGED: the row's values were gathered beforehand into
// a buffer (`g` reads it by edge, the column index is not loaded).
template <bool NT, class G, bool STAGED = false>
__device__ float exact_row_sum(const int32_t* __restrict__ col, int e_max, int e0, int ne,
                               const G& g, float s0, float* __restrict__ buf0,
                               float* __restrict__ buf1, ExactLds& x) {
  const int tid = threadIdx.x;
  const int nchunks = (ne + kScanChunk - 1) / kScanChunk;
  int idx[kScanPT];
  float vv[kScanPT];
  auto elem = [&](int c, int j) {  // edge offset in the row, clamped into it
    return min(c * kScanChunk + tid + j * kBlock, ne - 1);
  };
  auto load_idx = [&](int c) {
#pragma unroll
    for (int j = 0; j < kScanPT; j++)
      idx[j] = STAGED ? 0 : ld_stream<NT>(col + min(e0 + elem(c, j), e_max));
  };
  auto gather = [&](int c) {
#pragma unroll
    for (int j = 0; j < kScanPT; j++) vv[j] = g.value(e0 + elem(c, j), idx[j]);
  };
  float s = s0;
  if (nchunks == 0) return s;
  load_idx(0);
  gather(0);
  if (nchunks > 1) load_idx(1);
  for (int c = 0; c < nchunks; c++) {
    float* cur = (c & 1) ? buf1 : buf0;
    const int nc = min(kScanChunk, ne - c * kScanChunk);
#pragma unroll
    for (int j = 0; j < kScanPT; j++)
      if (tid + j * kBlock < nc) cur[tid + j * kBlock] = vv[j];
    if (c + 1 < nchunks) {
      gather(c + 1);                         // in flight during the scan of chunk c
      if (c + 2 < nchunks) load_idx(c + 2);  // likewise
    }
    __syncthreads();
    int start = 0;
    if (c == 0) {
      // the first terms, where s crosses a binade every few steps: one lane's
      // pipelined chain, in order
      const int nh = min(nc, kExactHead);
      if (tid == 0) x.s = lane_seq_sum<8>(s, cur, nh);
      __syncthreads();
      s = x.s;
      start = nh;
      __syncthreads();
    }
    s = exact_scan_chunk(cur, nc, start, s, x);
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
