// pull_iter.h -- device side of the fixed-point pull iteration, generic over a
// policy (plain / personalised PageRank, weighted PageRank, its adjoint).
//
// One iteration = the pull launches over the tile schedule (row tiles, and
// the exact hub rows beside them: gathers, per-row epilogue, per-workgroup L1
// partial) + one finalize launch (hub rows from their segment partials,
// deterministic L1 reduction, convergence flag); small iterations are one
// launch whose last workgroup does the finalize's work (pull_small_kernel,
// fused_finish).  The iteration counter and the done flag live on the device
// (IterState), so every iteration's launches are identical; launches after
// `done` return at once.
//
// A policy P provides
//   struct P::Args                      kernel arguments (pointers, scalars)
//   static constexpr bool kExactHubs    hub rows (whole-row tiles) summed in
//                                       edge order with sequential FP32
//                                       rounding (exact_hub_kernel), else hub
//                                       segments tree-summed + finalize
//   static constexpr bool kSeqAll       every row of a row tile summed by its
//                                       owner thread in edge order (else rows
//                                       of > 64 edges by a wavefront tree)
//   template <int NT> static G gather(const Args&, int odd)
//   template <int NT> static E epilogue(const Args&, int odd)
// with G / E as described in pull_engine.h (value(e, u); Pre, pre, init,
// finish).  `odd` is the parity of the iteration: the gathered vector is
// double-buffered.
#ifndef GD_PULL_ITER_H_
#define GD_PULL_ITER_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gd_common.h"
#include "pull_engine.h"

namespace gd {

constexpr int kFinBlocks = 128;  // finalize workgroups (<= kBlock: combined by one block)

struct IterState {
  int iter;         // iterations completed
  int done;         // set once converged or max_iter reached
  int max_iter;
  unsigned ticket;  // finalize workgroups arrived in this iteration
  double tol;
};

struct IterArgs {
  const Tile* __restrict__ tiles;  // row tiles (+ hub segments of tree-summing policies)
  int n_tiles;
  const Tile* __restrict__ hub_tiles;  // whole-row exact hubs (exact_hub_kernel)
  int n_hub_tiles;
  int hub_part_base;  // exact hub h writes its L1 partial to err_part[hub_part_base + h]
  const int32_t* __restrict__ off;
  const int32_t* __restrict__ col;
  int e_max;  // last valid index of col (a 1-element dummy when nnz == 0)
  const uint16_t* __restrict__ pos;  // sorted tiles: each edge's offset in its tile
  float* __restrict__ hub_part;
  double* __restrict__ err_part;
  IterState* st;
  // > 0: small iteration finished by its last pull workgroup (fused_finish)
  // over this many L1 partials, instead of a finalize launch
  int fuse_parts;
  double* __restrict__ errors;  // per-iteration L1 (fused_finish)
  // staged plans: exact hub h's values gathered beforehand (hub_stage_kernel)
  // at stage[hub_soff[h]...]; nullptr when the plan does not stage
  const int32_t* __restrict__ hub_soff;
  const float* __restrict__ stage;
};

// Values of a staged hub row: the row's first edge e0 is stage[0].
struct StagedGather {
  const float* __restrict__ s;
  int e0;
  __device__ __forceinline__ float value(int e, int /*u*/) const { return s[e - e0]; }
};

struct FinArgs {
  const int32_t* __restrict__ hub_row;
  const int32_t* __restrict__ hub_slot;
  const int32_t* __restrict__ hub_nseg;
  int n_hubs;
  int n_part;  // per-workgroup L1 partials written by the pull launch
  double* __restrict__ errors;
  double* __restrict__ fin_part;
};

static __global__ void iter_state_init_kernel(IterState* st, int max_iter, double tol) {
  st->iter = 0;
  st->done = max_iter <= 0;
  st->max_iter = max_iter;
  st->ticket = 0;
  st->tol = tol;
}

// Finish of a small iteration (it.fuse_parts > 0: fewer than 2*kBlock L1
// partials, no hub segments to close): every workgroup of the iteration's
// pull launches (row tiles and exact hubs) takes a ticket after writing its
// partial, and the last to arrive adds the partials -- in the grouping
// finalize_kernel uses with one workgroup, so the L1 values are the same bits
// -- and advances the iteration state.  Saves the finalize launch and its
// dependency, which dominate the iteration time of small graphs.
__device__ __forceinline__ void fused_finish(const IterArgs& it, double* s_err, int* s_last) {
  IterState* st = it.st;
  if (threadIdx.x == 0) {
    __threadfence();  // release this workgroup's err_part before taking a ticket
    const unsigned arrived = atomicAdd(&st->ticket, 1u);
    *s_last = (arrived == (unsigned)it.fuse_parts - 1);
  }
  __syncthreads();
  if (!*s_last) return;
  __threadfence();  // acquire: every partial of the iteration is visible
  const int n = it.fuse_parts;
  double err = 0.0;
  for (int i = threadIdx.x; i < n; i += 4 * kBlock) {
    const double a0 = it.err_part[i];
    const double a1 = (i + kBlock < n) ? it.err_part[i + kBlock] : 0.0;
    const double a2 = (i + 2 * kBlock < n) ? it.err_part[i + 2 * kBlock] : 0.0;
    const double a3 = (i + 3 * kBlock < n) ? it.err_part[i + 3 * kBlock] : 0.0;
    err += ((a0 + a1) + a2) + a3;
  }
  const double total = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) {
    const int i = st->iter;
    it.errors[i] = total;
    st->iter = i + 1;
    st->ticket = 0;
    if (total < st->tol || i + 1 >= st->max_iter) st->done = 1;
  }
}

// One workgroup of NB threads per tile (NB*IPT edges, 4*NB rows).  SORTED:
// each tile's edges are stored in source order with their tile offsets
// (tile_sort.h); sums are still in CSR order.  Larger workgroups sort larger
// tiles, so more lanes of a gather instruction share a cache line.
// tile_body: row tile b of the iteration (the caller checked st->done).
template <class P, int IPT, int NT, bool SORTED, int NB = kBlock>
__device__ __forceinline__ void tile_body(const typename P::Args& a, const IterArgs& it, int b,
                                          RowTileLds<IPT, NB>& lds) {
  static_assert(NB == kBlock || P::kExactHubs, "hub segments need kBlock-thread tiles");
  const int odd = it.st->iter & 1;
  __shared__ double s_err[NB / kWave];
  const Tile t = it.tiles[b];
  const auto g = P::template gather<NT>(a, odd);
  double err = 0.0;
  if (t.y >= 0) {
    const auto e = P::template epilogue<NT>(a, odd);
    err = row_tile<IPT, (NT > 0), SORTED, decltype(g), decltype(e), P::kSeqAll, NB>(
        it.off, it.col, it.pos, it.e_max, t.x, t.y, t.z, t.w - t.z, g, e, lds);
  } else if constexpr (!P::kExactHubs) {
    const float p = segment_sum<IPT, (NT > 0), SORTED>(it.col, it.pos, t.z, t.w - t.z, g, lds.val);
    if (threadIdx.x == 0) it.hub_part[-t.y - 1] = p;
  }
  const double tot = block_sum<NB / kWave>(err, s_err);
  if (threadIdx.x == 0) it.err_part[b] = tot;
  if constexpr (NB == kBlock) {
    if (it.fuse_parts > 0) {
      __shared__ int s_last;
      fused_finish(it, s_err, &s_last);
    }
  }
}

template <class P, int IPT, int NT, bool SORTED, int NB = kBlock>
__global__ __launch_bounds__(NB) void pull_tile_kernel(const typename P::Args a,
                                                       const IterArgs it) {
  if (it.st->done) return;
  __shared__ RowTileLds<IPT, NB> lds;
  tile_body<P, IPT, NT, SORTED, NB>(a, it, blockIdx.x, lds);
}

// Persistent, software-pipelined form: workgroup b walks tiles b, b+G, ...;
// the next tile's index / offset / epilogue loads are issued right after the
// current tile's gathers, so their latency overlaps the current tile's LDS
// reduction and epilogue (vector loads return in issue order, so waiting for
// the gathers never waits for the prefetch).  Same arithmetic and per-row
// order as pull_tile_kernel.
template <class P, int IPT, int NT, bool SORTED>
__global__ __launch_bounds__(kBlock) void pull_persistent_kernel(const typename P::Args a,
                                                                 const IterArgs it) {
  if (it.st->done) return;
  const int odd = it.st->iter & 1;
  __shared__ RowTileLds<IPT> lds;
  __shared__ double s_err[kWaves];
  const auto g = P::template gather<NT>(a, odd);
  const auto e = P::template epilogue<NT>(a, odd);
  using E = decltype(e);
  using Regs = TileRegs<IPT, typename E::Pre>;
  double err = 0.0;
  int t = blockIdx.x;
  if (t < it.n_tiles) {
    Tile cur = it.tiles[t];
    Regs R;
    issue_tile_loads<IPT, (NT > 0), SORTED>(cur, it.off, it.col, it.pos, it.e_max, e, R);
    for (;;) {
      const int tn = t + gridDim.x;
      const bool has_next = tn < it.n_tiles;
      const Tile nxt = it.tiles[has_next ? tn : t];
      const int ne = cur.w - cur.z;
      float v[IPT];
      if (ne > 0) gather_values<IPT, SORTED>(g, cur.z, ne, R.idx, R.pos, v);
      Regs Rn;
      // unconditional (clamped) so the compiler's wait counting stays exact
      issue_tile_loads<IPT, (NT > 0), SORTED>(nxt, it.off, it.col, it.pos, it.e_max, e, Rn);
      if (cur.y >= 0) {
        stage_row_tile<IPT, SORTED>(cur, R, v, lds);
        __syncthreads();
        err += reduce_row_tile<IPT, decltype(e), P::kSeqAll>(cur.x, cur.y - cur.x, e, R.pre, lds);
      } else if constexpr (!P::kExactHubs) {
        float sacc = 0.0f;
#pragma unroll
        for (int j = 0; j < IPT; j++) sacc += ((int)threadIdx.x + j * kBlock < ne) ? v[j] : 0.0f;
        const float p = block_sum<kWaves>(sacc, lds.val);
        if (threadIdx.x == 0) it.hub_part[-cur.y - 1] = p;
      }
      __syncthreads();  // LDS is reused by the next tile
      if (!has_next) break;
      cur = nxt;
      R = Rn;
      t = tn;
    }
  }
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) it.err_part[blockIdx.x] = tot;
}

// Whole-row hubs of an exact-hub policy: one workgroup per hub row adds the
// row in edge order with the reference's sequential rounding (exact_row_sum:
// the workgroup's binade scan, the next chunk's gathers in flight during the
// scan), then runs the row's epilogue.  A kernel of its own, launched before
// and beside the row-tile launch, longest row first.  Hub rows are stored in
// edge order.  hub_body: exact hub h of the iteration (the caller checked
// st->done).
struct HubLds {
  alignas(16) float buf[2 * kScanChunk];  // two chunks of the row, double-buffered
  ExactLds xs;
};

template <class P, int IPT, int NT, bool SORTED, bool STAGED = false>
__device__ __forceinline__ void hub_body(const typename P::Args& a, const IterArgs& it, int h,
                                         HubLds& lds) {
  const int odd = it.st->iter & 1;
  __shared__ double s_err[kWaves];
  float* const buf = lds.buf;
  ExactLds& xs = lds.xs;
  const Tile t = it.hub_tiles[h];
  const auto g = P::template gather<NT>(a, odd);
  const auto e = P::template epilogue<NT>(a, odd);
  const auto pre = e.pre(t.x);
  const int ne = t.w - t.z;
  float sum;
  if constexpr (STAGED)
    sum = exact_row_sum<(NT > 0), StagedGather, true>(
        it.col, it.e_max, t.z, ne, StagedGather{it.stage + it.hub_soff[h], t.z}, e.init(pre), buf,
        buf + kScanChunk, xs);
  else
    sum = exact_row_sum<(NT > 0)>(it.col, it.e_max, t.z, ne, g, e.init(pre), buf,
                                  buf + kScanChunk, xs);
  double err = 0.0;
  if (threadIdx.x == 0) err = e.finish(t.x, sum, pre);
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) it.err_part[it.hub_part_base + h] = tot;
  if (it.fuse_parts > 0) {
    __shared__ int s_last;
    fused_finish(it, s_err, &s_last);
  }
}

// Gathers the values of the long exact hub rows into `stage` ahead of
// exact_hub_kernel (same stream): the hub kernel then scans values read
// coalesced from L2 instead of waiting on each chunk's random gathers, so a
// long row's workgroup is bound by its scan.  Workgroup b gathers block b:
// x = first edge, y = edges (<= kScanChunk), z = offset in stage.
template <class P, int NT>
__global__ __launch_bounds__(kBlock) void hub_stage_kernel(const typename P::Args a,
                                                           const IterArgs it,
                                                           const int4* __restrict__ blocks,
                                                           float* __restrict__ stage) {
  if (it.st->done) return;
  const int4 b = blocks[blockIdx.x];
  const auto g = P::template gather<NT>(a, it.st->iter & 1);
  int idx[kScanPT];
  float v[kScanPT];
#pragma unroll
  for (int j = 0; j < kScanPT; j++)
    idx[j] = ld_stream<(NT > 0)>(it.col + b.x + min((int)threadIdx.x + j * kBlock, b.y - 1));
#pragma unroll
  for (int j = 0; j < kScanPT; j++)
    v[j] = g.value(b.x + min((int)threadIdx.x + j * kBlock, b.y - 1), idx[j]);
#pragma unroll
  for (int j = 0; j < kScanPT; j++)
    if ((int)threadIdx.x + j * kBlock < b.y) stage[b.z + threadIdx.x + j * kBlock] = v[j];
}

// Held to 3 waves per SIMD (<= 168 VGPRs): the hub workgroups share the CUs
// with the row tiles running beside them.  The gathering form spills three
// registers once per 4096-edge chunk, outside the scan's loops; the staged
// form spills none.
template <class P, int IPT, int NT, bool SORTED, bool STAGED>
__global__ __launch_bounds__(kBlock, 3) void exact_hub_kernel(const typename P::Args a,
                                                              const IterArgs it) {
  if (it.st->done) return;
  __shared__ HubLds lds;
  hub_body<P, IPT, NT, SORTED, STAGED>(a, it, blockIdx.x, lds);
}

// Small iterations (it.fuse_parts > 0) in one launch: workgroups [0,
// n_hub_tiles) take the exact hubs (longest first), the rest the row tiles,
// and the last workgroup to finish closes the iteration (fused_finish) --
// one kernel per iteration, no fork, join or finalize.
// The two bodies share one LDS allocation.
template <class P, int IPT, int NT, bool SORTED>
__global__ __launch_bounds__(kBlock, 3) void pull_small_kernel(const typename P::Args a,
                                                               const IterArgs it) {
  if (it.st->done) return;
  __shared__ union {
    RowTileLds<IPT> tile;
    HubLds hub;
  } lds;
  if constexpr (P::kExactHubs) {
    if ((int)blockIdx.x < it.n_hub_tiles) {
      hub_body<P, IPT, NT, SORTED>(a, it, blockIdx.x, lds.hub);
      return;
    }
  }
  tile_body<P, IPT, NT, SORTED>(a, it, blockIdx.x - it.n_hub_tiles, lds.tile);
}

// Hub-row epilogues (segment partials summed in segment order) and the L1
// reduction: each workgroup sums a contiguous slice of the per-workgroup
// partials, the last workgroup to arrive adds the workgroup sums in index
// order, so the result does not depend on scheduling.
template <class P>
__global__ __launch_bounds__(kBlock) void finalize_kernel(const typename P::Args a,
                                                          const IterArgs it, const FinArgs f) {
  IterState* st = it.st;
  if (st->done) return;
  __shared__ double s_err[kWaves];
  __shared__ int s_last;
  const int odd = st->iter & 1;
  const auto e = P::template epilogue<0>(a, odd);
  double err = 0.0;
  for (int h = blockIdx.x * kBlock + threadIdx.x; h < f.n_hubs; h += gridDim.x * kBlock) {
    const int v = f.hub_row[h];
    const auto p = e.pre(v);
    const float* part = it.hub_part + f.hub_slot[h];
    float s = e.init(p);
    for (int k = 0; k < f.hub_nseg[h]; k++) s += part[k];
    err += e.finish(v, s, p);
  }
  const int per = (f.n_part + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(lo + per, f.n_part);
  for (int i = lo + threadIdx.x; i < hi; i += 4 * kBlock) {
    const double a0 = it.err_part[i];
    const double a1 = (i + kBlock < hi) ? it.err_part[i + kBlock] : 0.0;
    const double a2 = (i + 2 * kBlock < hi) ? it.err_part[i + 2 * kBlock] : 0.0;
    const double a3 = (i + 3 * kBlock < hi) ? it.err_part[i + 3 * kBlock] : 0.0;
    err += ((a0 + a1) + a2) + a3;
  }
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) {
    f.fin_part[blockIdx.x] = tot;
    __threadfence();  // release fin_part before taking a ticket
    const unsigned arrived = atomicAdd(&st->ticket, 1u);
    s_last = (arrived == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();  // acquire: every other workgroup's fin_part is visible
  const double v = (threadIdx.x < gridDim.x) ? f.fin_part[threadIdx.x] : 0.0;
  const double total = block_sum<kWaves>(v, s_err);
  if (threadIdx.x == 0) {
    const int i = st->iter;
    f.errors[i] = total;
    st->iter = i + 1;
    st->ticket = 0;
    if (total < st->tol || i + 1 >= st->max_iter) st->done = 1;
  }
}

}  // namespace gd

#endif  // GD_PULL_ITER_H_
