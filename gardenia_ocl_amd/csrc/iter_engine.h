// iter_engine.h -- host side of the fixed-point pull iteration (pull_iter.h):
// tile schedule, scratch, persistent grid size, and the chunked launch loop.
#ifndef GD_ITER_ENGINE_H_
#define GD_ITER_ENGINE_H_

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "pull_iter.h"
#include "tile_sort.h"

namespace gd {

// Policies that may run row tiles on workgroups wider than kBlock threads
// (static constexpr bool kWideTiles = true); the others always use kBlock.
template <class P, class = void>
struct wide_tiles : std::false_type {};
template <class P>
struct wide_tiles<P, std::void_t<decltype(P::kWideTiles)>> : std::bool_constant<P::kWideTiles> {};

struct IterEngine {
  int m = 0;
  int64_t nnz = 0;
  hipStream_t stream = nullptr;
  int ipt = 8;               // edges per thread per tile (8 or 16)
  int block = kBlock;        // threads per row-tile workgroup: kBlock, or 1024 (wide_tiles
                             // policies): tiles of block*ipt edges and 4*block rows
  int nt = 1;                // cache-hint level (see gd_pr_plan_config.nontemporal)
  bool persistent = true;
  int exact_len = 0;         // > 0: rows longer than this are whole-row exact hub tiles
  bool sort_tiles = false;   // gather each tile's edges in source order (tile_sort.h)
  int32_t* col_owned = nullptr;  // the caller's private copy of col: sorted in place
  int pull_grid = 0;
  const int32_t* off = nullptr;
  const int32_t* col = nullptr;
  int e_max = 0;
  DeviceBuffer dummy_col, col_sorted, pos;
  Schedule sched;
  int n_tiles = 0, n_hubs = 0, n_hub_tiles = 0, fin_blocks = 1;
  bool fused = false;  // the pull launches finish each iteration (no finalize launch)
  // Exact hub rows gathered by hub_stage_kernel ahead of exact_hub_kernel
  // (every hub of the plan, or none): -1 auto (when the longest hub's chain would
  // outlast the row tiles: longest row * 500 > edges in row tiles), 0 off, 1 on.
  // Not used by the one-launch small iterations.
  int stage_hubs = -1;
  int n_stage_blocks = 0, n_staged_hubs = 0;
  DeviceBuffer stage, stage_blocks, hub_soff;
  DeviceBuffer tiles, hub_tiles, hub_row, hub_slot, hub_nseg, hub_part, err_part, fin_part, state,
      errors;
  int errors_cap = 0;
  std::vector<hipEvent_t> ev;
  // exact hubs run on a second stream, forked from / joined to `stream`
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // A chunk of iterations (identical launches) captured into a hipGraph when it
  // will run again (later in the run, or in a later run with the same
  // arguments), then replayed: one graph launch per chunk
  // instead of 3-4 launches and a fork/join per iteration.  Keyed by the launch
  // arguments and the chunk length (a run has at most two lengths: the chunk
  // and the tail).
  struct ChunkGraph {
    std::vector<unsigned char> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<ChunkGraph> graphs;
  std::vector<std::vector<unsigned char>> seen;      // keys of earlier runs
  std::vector<std::vector<unsigned char>> this_run;  // keys of the current run
  bool use_graphs = true;  // cleared for good if a capture ever fails
  // Capture runs on a private stream (the caller's may be the legacy default
  // stream, which cannot be captured); the graph is launched on the caller's.
  hipStream_t cap_stream = nullptr;

  void drop_graphs() {
    for (ChunkGraph& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs.clear();
  }

  ~IterEngine() {
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side) (void)hipStreamDestroy(side);
  }

  // Schedule and scratch over a validated CSR already in HBM (offsets are
  // copied to the host for the schedule).
  int build(int m_, int64_t nnz_, const int32_t* d_off, const int32_t* d_col, hipStream_t s) {
    m = m_;
    nnz = nnz_;
    stream = s;
    off = d_off;
    col = d_col;
    int rc = GD_OK;
    if (nnz == 0) {
      rc = dummy_col.alloc(4);
      if (rc) return rc;
      GD_HIP_TRY(hipMemsetAsync(dummy_col.ptr, 0, 4, stream));
      col = dummy_col.as<int32_t>();
    }
    e_max = (int)std::max<int64_t>(nnz - 1, 0);
    std::vector<int32_t> h_off(m + 1);
    GD_HIP_TRY(hipMemcpyAsync(h_off.data(), off, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost,
                              stream));
    GD_HIP_TRY(hipStreamSynchronize(stream));
    rc = build_schedule(m, nnz, h_off.data(), block * ipt, kRowsPerThread * block, &sched,
                        exact_len);
    if (rc) return rc;
    n_hubs = (int)sched.hub_row.size();
    rc = pos.alloc(sort_tiles ? 2 * (size_t)std::max<int64_t>(nnz, 1) : 2);
    if (rc) return rc;
    if (sort_tiles && nnz > 0) {
      int32_t* dst = col_owned;
      if (!dst) {
        rc = col_sorted.alloc(4 * (size_t)nnz);
        if (rc) return rc;
        dst = col_sorted.as<int32_t>();
      }
      // exact hub rows are kept in edge order
      rc = build_tile_sort(sched, block * ipt, col, dst, pos.as<uint16_t>(), stream,
                           exact_len > 0);
      if (rc) return rc;
      col = dst;
    }
    // exact hubs: their own tile list (longest first) and kernel
    std::vector<Tile> row_tiles, hubs;
    for (const Tile& t : sched.tiles) ((exact_len > 0 && t.y < 0) ? hubs : row_tiles).push_back(t);
    std::stable_sort(hubs.begin(), hubs.end(),
                     [](const Tile& a, const Tile& b) { return a.w - a.z > b.w - b.z; });
    sched.tiles = row_tiles;
    n_tiles = (int)row_tiles.size();
    n_hub_tiles = (int)hubs.size();
    if (n_hub_tiles) {
      rc = hub_tiles.alloc(sizeof(Tile) * n_hub_tiles);
      if (rc) return rc;
      GD_HIP_TRY(hipMemcpyAsync(hub_tiles.ptr, hubs.data(), sizeof(Tile) * n_hub_tiles,
                                hipMemcpyHostToDevice, stream));
      // staged: every hub, in hub order
      int64_t hub_edges = 0;
      for (const Tile& t : hubs) hub_edges += t.w - t.z;
      const int64_t longest = hubs[0].w - hubs[0].z;
      const bool want = stage_hubs == 1 ||
                        (stage_hubs < 0 && longest * 500 > (int64_t)nnz - hub_edges);
      std::vector<int32_t> soff(n_hub_tiles, -1);
      std::vector<int4> blocks;
      int64_t staged = 0;
      for (int h = 0; want && h < n_hub_tiles; h++) {
        const int ne = hubs[h].w - hubs[h].z;
        soff[h] = (int32_t)staged;
        n_staged_hubs++;
        for (int o = 0; o < ne; o += kScanChunk)
          blocks.push_back(make_int4(hubs[h].z + o, std::min(kScanChunk, ne - o),
                                     (int)(staged + o), 0));
        staged += ne;
      }
      n_stage_blocks = (int)blocks.size();
      if (n_stage_blocks) {
        rc = stage.alloc(4 * (size_t)staged);
        if (!rc) rc = stage_blocks.alloc(sizeof(int4) * blocks.size());
        if (!rc) rc = hub_soff.alloc(4 * (size_t)n_hub_tiles);
        if (rc) return rc;
        GD_HIP_TRY(hipMemcpyAsync(stage_blocks.ptr, blocks.data(), sizeof(int4) * blocks.size(),
                                  hipMemcpyHostToDevice, stream));
        GD_HIP_TRY(hipMemcpyAsync(hub_soff.ptr, soff.data(), 4 * (size_t)n_hub_tiles,
                                  hipMemcpyHostToDevice, stream));
      }
      GD_HIP_TRY(hipStreamSynchronize(stream));  // host vectors go out of scope
      GD_HIP_TRY(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      GD_HIP_TRY(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
      GD_HIP_TRY(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
    }
    rc = tiles.alloc(sizeof(Tile) * std::max(1, n_tiles));
    if (!rc) rc = hub_row.alloc(4 * (size_t)std::max(1, n_hubs));
    if (!rc) rc = hub_slot.alloc(4 * (size_t)std::max(1, n_hubs));
    if (!rc) rc = hub_nseg.alloc(4 * (size_t)std::max(1, n_hubs));
    if (!rc) rc = hub_part.alloc(4 * (size_t)std::max(1, sched.n_slots));
    if (!rc) rc = err_part.alloc(8 * ((size_t)std::max(1, n_tiles) + n_hub_tiles));
    if (!rc) rc = fin_part.alloc(8 * (size_t)kFinBlocks);
    if (!rc) rc = state.alloc(sizeof(IterState));
    if (rc) return rc;
    auto up = [&](DeviceBuffer& b, const void* src, size_t n) -> hipError_t {
      return n ? hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, stream) : hipSuccess;
    };
    GD_HIP_TRY(up(tiles, sched.tiles.data(), sizeof(Tile) * n_tiles));
    GD_HIP_TRY(up(hub_row, sched.hub_row.data(), 4 * (size_t)n_hubs));
    GD_HIP_TRY(up(hub_slot, sched.hub_slot.data(), 4 * (size_t)n_hubs));
    GD_HIP_TRY(up(hub_nseg, sched.hub_nseg.data(), 4 * (size_t)n_hubs));
    GD_HIP_TRY(hipStreamSynchronize(stream));
    return GD_OK;
  }

  template <class P>
  int setup_grid() {
    if (persistent) {
      int dev = 0, cus = 0, per_cu = 0;
      GD_HIP_TRY(hipGetDevice(&dev));
      GD_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const void* fn = (ipt == 16)
                           ? reinterpret_cast<const void*>(&pull_persistent_kernel<P, 16, 1, true>)
                           : reinterpret_cast<const void*>(&pull_persistent_kernel<P, 8, 1, true>);
      GD_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
      pull_grid = std::max(1, std::min(n_tiles, cus * std::max(1, per_cu)));
    } else {
      pull_grid = n_tiles;
    }
    fin_blocks = std::max(
        1, std::min(kFinBlocks, std::max(pull_grid + n_hub_tiles, n_hubs) / kBlock));
    // small iterations end in their last pull workgroup (fused_finish): no hub
    // segments to close, one finalize workgroup's worth of partials
    const int parts = (n_tiles > 0 ? pull_grid : 0) + n_hub_tiles;
    fused = !persistent && block == kBlock && (exact_len > 0 || n_hubs == 0) && parts > 0 &&
            parts < 2 * kBlock;
    if (fused && n_stage_blocks) {  // one-launch iterations gather their hubs themselves
      stage.release();
      stage_blocks.release();
      hub_soff.release();
      n_stage_blocks = n_staged_hubs = 0;
    }
    return GD_OK;
  }

  IterArgs iter_args() {
    const int parts = (n_tiles > 0 ? pull_grid : 0) + n_hub_tiles;
    const bool staged = n_stage_blocks > 0 && !fused;
    return IterArgs{tiles.as<Tile>(), n_tiles, hub_tiles.as<Tile>(), n_hub_tiles,
                    n_tiles > 0 ? pull_grid : 0, off, col, e_max, pos.as<uint16_t>(),
                    hub_part.as<float>(), err_part.as<double>(), state.as<IterState>(),
                    fused ? parts : 0, errors.as<double>(),
                    staged ? hub_soff.as<int32_t>() : nullptr, stage.as<float>()};
  }

  // exact hubs finish their rows in the pull launch: no hub epilogues here
  // L1 partials: the row-tile launch's (pull_grid of them) then the exact hubs'
  FinArgs fin_args() {
    return FinArgs{hub_row.as<int32_t>(), hub_slot.as<int32_t>(), hub_nseg.as<int32_t>(),
                   exact_len > 0 ? 0 : n_hubs, (n_tiles > 0 ? pull_grid : 0) + n_hub_tiles,
                   errors.as<double>(), fin_part.as<double>()};
  }

  template <class P, int IPT, int NT, bool SORTED>
  void launch_wide(const typename P::Args& a, const IterArgs& it, hipStream_t ts) {
    if constexpr (wide_tiles<P>::value) {
      if (block == 512)
        hipLaunchKernelGGL((pull_tile_kernel<P, IPT, NT, SORTED, 512>), dim3(n_tiles), dim3(512),
                           0, ts, a, it);
      else
        hipLaunchKernelGGL((pull_tile_kernel<P, IPT, NT, SORTED, 1024>), dim3(n_tiles),
                           dim3(1024), 0, ts, a, it);
    }
  }

  template <class P, int IPT, int NT, bool SORTED>
  void launch_pull_s(const typename P::Args& a, const IterArgs& it) {
    // Exact hubs go first, on the plan's stream, so that their workgroups
    // (longest row first) are placed before the row tiles fill the CUs; the
    // row tiles fork onto the side stream (after the same prior work) and
    // join back.
    if (fused) {  // small iteration: one launch (hubs first, then row tiles)
      hipLaunchKernelGGL((pull_small_kernel<P, IPT, NT, SORTED>),
                         dim3(n_hub_tiles + (n_tiles > 0 ? pull_grid : 0)), dim3(kBlock), 0,
                         stream, a, it);
      return;
    }
    const bool fork = n_hub_tiles > 0 && n_tiles > 0;
    const bool staged = P::kExactHubs && n_hub_tiles > 0 && it.hub_soff;
    // staged hubs: their values are gathered first, by the whole GPU, before
    // the fork (the row tiles would otherwise fill the CUs first and hold the
    // hub rows back until they drain)
    if (staged)
      hipLaunchKernelGGL((hub_stage_kernel<P, NT>), dim3(n_stage_blocks), dim3(kBlock), 0, stream,
                         a, it, stage_blocks.as<int4>(), stage.as<float>());
    hipStream_t ts = stream;
    if (fork) {
      (void)hipEventRecord(fork_ev, stream);
      (void)hipStreamWaitEvent(side, fork_ev, 0);
      ts = side;
    }
    if (n_hub_tiles) {
      if constexpr (P::kExactHubs) {
        if (staged) {
          hipLaunchKernelGGL((exact_hub_kernel<P, IPT, NT, SORTED, true>), dim3(n_hub_tiles),
                             dim3(kBlock), 0, stream, a, it);
        } else {
          hipLaunchKernelGGL((exact_hub_kernel<P, IPT, NT, SORTED, false>), dim3(n_hub_tiles),
                             dim3(kBlock), 0, stream, a, it);
        }
      }
    }
    if (n_tiles > 0) {
      if (persistent)
        hipLaunchKernelGGL((pull_persistent_kernel<P, IPT, NT, SORTED>), dim3(pull_grid),
                           dim3(kBlock), 0, ts, a, it);
      else if (wide_tiles<P>::value && block > kBlock)
        launch_wide<P, IPT, NT, SORTED>(a, it, ts);
      else
        hipLaunchKernelGGL((pull_tile_kernel<P, IPT, NT, SORTED>), dim3(n_tiles), dim3(kBlock), 0,
                           ts, a, it);
    }
    if (fork) {
      (void)hipEventRecord(join_ev, side);
      (void)hipStreamWaitEvent(stream, join_ev, 0);  // join
    }
  }

  template <class P, int IPT, int NT>
  void launch_pull_t(const typename P::Args& a, const IterArgs& it) {
    if (sort_tiles)
      launch_pull_s<P, IPT, NT, true>(a, it);
    else
      launch_pull_s<P, IPT, NT, false>(a, it);
  }

  template <class P>
  void launch_pull(const typename P::Args& a, const IterArgs& it) {
    switch (ipt * 4 + nt) {
      case 8 * 4 + 0: launch_pull_t<P, 8, 0>(a, it); break;
      case 8 * 4 + 1: launch_pull_t<P, 8, 1>(a, it); break;
      case 8 * 4 + 2: launch_pull_t<P, 8, 2>(a, it); break;
      case 16 * 4 + 0: launch_pull_t<P, 16, 0>(a, it); break;
      case 16 * 4 + 1: launch_pull_t<P, 16, 1>(a, it); break;
      default: launch_pull_t<P, 16, 2>(a, it); break;
    }
  }

  // Runs up to max_iter iterations (stopping after the first with L1 < tol).
  // `make_args(odd)` is not needed: the kernels pick buffers by the device
  // iteration parity.  Fills iterations / converged / final_error / pull
  // timing / near_tie of *s (solve_ms is the caller's).
  // The executable graph of n iterations with these arguments, captured on
  // first use; nullptr if capture is unavailable (then launches are direct).
  template <class P>
  hipGraphExec_t chunk_graph(const typename P::Args& a, const IterArgs& it, const FinArgs& f,
                             int n, bool repeats) {
    if (!use_graphs) return nullptr;
    std::vector<unsigned char> key(sizeof(a) + sizeof(it) + sizeof(f) + sizeof(n));
    unsigned char* k = key.data();
    memcpy(k, &a, sizeof(a));
    memcpy(k + sizeof(a), &it, sizeof(it));
    memcpy(k + sizeof(a) + sizeof(it), &f, sizeof(f));
    memcpy(k + sizeof(a) + sizeof(it) + sizeof(f), &n, sizeof(n));
    for (ChunkGraph& g : graphs)
      if (g.key == key) return g.exec;
    // capture when the chunk will be launched again in this run, or once the
    // arguments come back in a later run: single-chunk runs whose arguments
    // change every time launch directly instead of paying a capture per run
    bool again = repeats;
    for (const auto& k2 : seen) again = again || k2 == key;
    if (!again) {
      bool noted = false;
      for (const auto& k2 : this_run) noted = noted || k2 == key;
      if (!noted) this_run.push_back(key);
      return nullptr;
    }
    if (graphs.size() >= 4) drop_graphs();
    ChunkGraph g;
    g.key = std::move(key);
    if (!cap_stream && hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      use_graphs = false;
      return nullptr;
    }
    if (hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      use_graphs = false;
      return nullptr;
    }
    const hipStream_t user = stream;
    stream = cap_stream;  // launch_pull forks from / joins to `stream`
    for (int i = 0; i < n; i++) {
      launch_pull<P>(a, it);
      if (!fused)
        hipLaunchKernelGGL(finalize_kernel<P>, dim3(fin_blocks), dim3(kBlock), 0, stream, a, it,
                           f);
    }
    stream = user;
    hipError_t e = hipStreamEndCapture(cap_stream, &g.graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (g.graph) (void)hipGraphDestroy(g.graph);
      use_graphs = false;
      return nullptr;
    }
    graphs.push_back(std::move(g));
    return graphs.back().exec;
  }

  template <class P>
  int run(const typename P::Args& a, int max_iter, double tol, int chunk_opt, bool profile,
          gd_pr_stats* s, double* errors_out) {
    if (errors_cap < std::max(1, max_iter)) {
      int rc = errors.alloc(sizeof(double) * std::max(1, max_iter));
      if (rc) return rc;
      errors_cap = std::max(1, max_iter);
    }
    const int chunk = chunk_opt > 0 ? chunk_opt : (tol > 0 ? 8 : std::max(1, max_iter));
    if (profile) {
      while ((int)ev.size() < 2 * chunk) {
        hipEvent_t e;
        GD_HIP_TRY(hipEventCreate(&e));
        ev.push_back(e);
      }
    }
    const IterArgs it = iter_args();
    const FinArgs f = fin_args();
    for (auto& k2 : this_run) {  // keys of the previous run become capturable
      if (seen.size() >= 8) seen.erase(seen.begin());
      seen.push_back(std::move(k2));
    }
    this_run.clear();
    hipLaunchKernelGGL(iter_state_init_kernel, dim3(1), dim3(1), 0, stream, it.st, max_iter, tol);
    GD_HIP_TRY(hipGetLastError());
    int launched = 0;
    IterState h{};
    float pull_ms = 0.0f;
    int pull_n = 0;
    while (launched < max_iter) {
      const int n = std::min(chunk, max_iter - launched);
      hipGraphExec_t exec =
          profile ? nullptr : chunk_graph<P>(a, it, f, n, max_iter - launched >= 2 * n);
      if (exec) {
        GD_HIP_TRY(hipGraphLaunch(exec, stream));
      } else {
        for (int k = 0; k < n; k++) {
          if (profile) GD_HIP_TRY(hipEventRecord(ev[2 * k], stream));
          launch_pull<P>(a, it);
          if (profile) GD_HIP_TRY(hipEventRecord(ev[2 * k + 1], stream));
          if (!fused)
            hipLaunchKernelGGL(finalize_kernel<P>, dim3(fin_blocks), dim3(kBlock), 0, stream, a,
                               it, f);
        }
      }
      GD_HIP_TRY(hipGetLastError());
      launched += n;
      GD_HIP_TRY(hipMemcpyAsync(&h, it.st, sizeof(IterState), hipMemcpyDeviceToHost, stream));
      GD_HIP_TRY(hipStreamSynchronize(stream));
      if (profile) {
        // only launches that did work count; later ones in the chunk exit at once
        const int worked = std::min(n, h.iter - (launched - n));
        for (int k = 0; k < worked; k++) {
          float ms = 0.0f;
          GD_HIP_TRY(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
          pull_ms += ms;
          pull_n++;
        }
      }
      if (h.done) break;
    }
    s->iterations = h.iter;
    s->pull_ms = pull_ms;
    s->pull_launches = pull_n;
    std::vector<double> errs(std::max(1, h.iter));
    if (h.iter > 0) {
      GD_HIP_TRY(hipMemcpy(errs.data(), errors.ptr, sizeof(double) * h.iter,
                           hipMemcpyDeviceToHost));
      s->final_error = errs[h.iter - 1];
      s->converged = s->final_error < tol;
      // A stop decision within the band of tol where this library's L1 value
      // and the reference's can differ could fall one iteration apart from
      // the reference's: flag it.  Policies that sum every row in the
      // reference's order (kSeqAll + kExactHubs: plain, personalised and
      // weighted PageRank) give the reference's scores bit for bit, so their
      // L1 values differ only in the FP64 summation order of |change| (a
      // few ulps); the others tree-sum rows of 65-4096 edges, ~6e-8 apart
      // (measured on RMAT-12 to RMAT-24 from uniform scores).
      const double band = (P::kSeqAll && P::kExactHubs) ? 1e-12 + 1e-12 * tol
                                                          : 1e-7 + 1e-6 * tol;
      for (int i = std::max(0, h.iter - 2); i < h.iter; i++)
        if (tol > 0 && std::fabs(errs[i] - tol) <= band) s->near_tie = 1;
      if (errors_out)
        for (int i = 0; i < h.iter; i++) errors_out[i] = errs[i];
    }
    return GD_OK;
  }
};

}  // namespace gd

#endif  // GD_ITER_ENGINE_H_
