// multigraph.hip -- PageRank on many independent graphs in one launch per
// iteration, iterations captured in a hipGraph.
//
// The graphs are packed into one block-diagonal union (vertex and edge ids
// shifted by per-graph bases) at plan time; every graph keeps its own tile
// schedule (the same tiles a single-graph plan would build, so results are
// bit-identical to solving each graph alone), its own base score (1-d)/m_g,
// its own iteration counter / parity / convergence flag.  One pull launch
// covers the tiles of every graph (tiles of converged graphs exit at once;
// rows of > 2048 edges are whole-row tiles summed in the reference's order),
// then one finalize workgroup per graph adds its tiles' L1 partials in a fixed
// order and updates that graph's state -- no cross-workgroup synchronisation.  A chunk of iterations (pull + finalize
// launches with identical arguments) is captured once into a hipGraph and
// replayed until every graph has converged; small graphs are launch-bound, so
// this is where the time goes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "pr_policy.h"
#include "pull_engine.h"

namespace gd {
namespace {

constexpr int kMgIPT = 16;  // as the single-graph plan: same tiles, same per-row sums

struct GraphState {
  int iter;
  int done;
  int tile_lo, tile_hi;  // this graph's tiles
  int hub_lo, hub_hi;    // this graph's hub rows
  float base;            // (1 - d) / m_g
  int pad;
};

struct MgArgs {
  const Tile* __restrict__ tiles;
  const int32_t* __restrict__ tile_graph;
  const int32_t* __restrict__ off;
  const int32_t* __restrict__ col;
  int e_max;
  const int32_t* __restrict__ outdeg;
  float* __restrict__ scores;
  float* __restrict__ c0;
  float* __restrict__ c1;
  float* __restrict__ hub_part;
  double* __restrict__ err_part;
  GraphState* gs;
  const int32_t* __restrict__ hub_row;
  const int32_t* __restrict__ hub_slot;
  const int32_t* __restrict__ hub_nseg;
  double* __restrict__ errors;  // [graph][max_iter]
  int max_iter;
  double tol;
  float damp;
};

__global__ __launch_bounds__(kBlock) void mg_pull_kernel(const MgArgs a) {
  const int g = a.tile_graph[blockIdx.x];
  if (a.gs[g].done) return;
  const int odd = a.gs[g].iter & 1;
  // a row tile's LDS, or a hub row's two chunk buffers + the broadcast sum
  static_assert(sizeof(RowTileLds<kMgIPT>) >= (2 * kHubChunk + 1) * sizeof(float), "LDS union");
  __shared__ __align__(16) unsigned char smem[sizeof(RowTileLds<kMgIPT>)];
  RowTileLds<kMgIPT>& lds = *reinterpret_cast<RowTileLds<kMgIPT>*>(smem);
  __shared__ double s_err[kWaves];
  const Tile t = a.tiles[blockIdx.x];
  const PrGather gat{odd ? a.c1 : a.c0};
  double err = 0.0;
  const PrEpilogue<0> e{a.scores, odd ? a.c0 : a.c1, a.outdeg, a.gs[g].base, a.damp, -1};
  if (t.y >= 0) {
    err = row_tile<kMgIPT, false, false, decltype(gat), decltype(e), true>(
        a.off, a.col, nullptr, a.e_max, t.x, t.y, t.z, t.w - t.z, gat, e, lds);
  } else {  // whole hub row in the reference's order, as the single-graph plan
    const auto pre = e.pre(t.x);
    float* const buf = reinterpret_cast<float*>(smem);
    const float sum = hub_chain_sum<8, false>(a.col, nullptr, a.e_max, t.z, t.w - t.z, gat,
                                              e.init(pre), buf, buf + kHubChunk, buf + 2 * kHubChunk);
    if (threadIdx.x == 0) err = e.finish(t.x, sum, pre);
  }
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) a.err_part[blockIdx.x] = tot;
}

// One workgroup per graph: L1 over the graph's tiles (fixed order; hub rows
// were finished by their pull workgroups), state update.
__global__ __launch_bounds__(kBlock) void mg_finalize_kernel(const MgArgs a) {
  const int g = blockIdx.x;
  GraphState& st = a.gs[g];
  if (st.done) return;
  __shared__ double s_err[kWaves];
  double err = 0.0;
  for (int t = st.tile_lo + threadIdx.x; t < st.tile_hi; t += kBlock) err += a.err_part[t];
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) {
    const int i = st.iter;
    a.errors[(int64_t)g * a.max_iter + i] = tot;
    st.iter = i + 1;
    if (tot < a.tol || i + 1 >= a.max_iter) st.done = 1;
  }
}

__global__ void mg_init_kernel(const MgArgs a, int n_graphs, int total_m, int max_iter) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < total_m; v += gridDim.x * blockDim.x)
    a.c0[v] = a.scores[v] / (float)a.outdeg[v];
  if (blockIdx.x == 0)
    for (int g = threadIdx.x; g < n_graphs; g += blockDim.x) {
      a.gs[g].iter = 0;
      a.gs[g].done = max_iter <= 0;
    }
}

// Copies m_g floats per graph between the caller's arrays and the union.
__global__ void mg_copy_scores_kernel(int n_graphs, float* const* __restrict__ user,
                                      const int32_t* __restrict__ vbase, float* __restrict__ uni,
                                      int to_union) {
  const int g = blockIdx.y;
  if (g >= n_graphs) return;
  const int b = vbase[g], n = vbase[g + 1] - b;
  float* u = user[g];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (to_union)
      uni[b + i] = u[i];
    else
      u[i] = uni[b + i];
  }
}

// Union CSR: offsets/columns of graph g shifted by its edge / vertex base.
__global__ void mg_pack_kernel(const int32_t* __restrict__ src_off,
                               const int32_t* __restrict__ src_col,
                               const int32_t* __restrict__ src_deg, int m, int64_t nnz, int vbase,
                               int64_t ebase, int32_t* __restrict__ off, int32_t* __restrict__ col,
                               int32_t* __restrict__ deg) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride)
    col[ebase + i] = src_col[i] + vbase;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    off[vbase + i] = (int32_t)(src_off[i] + ebase);
    deg[vbase + i] = src_deg[i];
  }
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

}  // namespace
}  // namespace gd

using namespace gd;

struct gd_multi_pr_plan {
  int n_graphs = 0;
  int total_m = 0;
  int64_t total_nnz = 0;
  hipStream_t user_stream = nullptr;
  hipStream_t stream = nullptr;  // private stream (graph capture needs a non-null one)
  std::vector<int32_t> m, vbase;
  DeviceBuffer off, col, deg, scores, c0, c1, tiles, tile_graph, hub_row, hub_slot, hub_nseg,
      hub_part, err_part, gs, errors, d_vbase, d_user_ptrs;
  int n_tiles = 0;
  int e_max = 0;
  int errors_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_chunk = 0, graph_max_iter = -1;
  double graph_tol = -1;
  float graph_damp = -1;
  std::vector<GraphState> h_state;
  ~gd_multi_pr_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" {

void gd_multi_pr_default_config(gd_multi_pr_config* c) {
  c->use_graph = 1;
  c->chunk = 0;
}

int gd_multi_pr_plan_create(int n_graphs, const int32_t* m, const int64_t* nnz,
                            const int32_t* const* d_off, const int32_t* const* d_col,
                            const int32_t* const* d_outdeg, void* stream,
                            gd_multi_pr_plan** out) {
  GD_CHECK_ARG(out != nullptr, "plan output pointer is NULL");
  *out = nullptr;
  GD_CHECK_ARG(n_graphs > 0 && m && nnz && d_off && d_col && d_outdeg, "NULL argument");
  int64_t tm = 0, te = 0;
  for (int g = 0; g < n_graphs; g++) {
    GD_CHECK_ARG(m[g] > 0 && nnz[g] >= 0, "every graph needs m > 0 and nnz >= 0");
    GD_CHECK_ARG(d_off[g] && d_outdeg[g] && (nnz[g] == 0 || d_col[g]), "NULL graph array");
    tm += m[g];
    te += nnz[g];
  }
  GD_CHECK_ARG(tm < (1ll << 31) && te < (1ll << 31), "union too large for int32 ids");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  std::unique_ptr<gd_multi_pr_plan> owner(new gd_multi_pr_plan());
  gd_multi_pr_plan* p = owner.get();
  auto fail = [](int rc) { return rc; };  // owner frees the plan
  p->n_graphs = n_graphs;
  p->total_m = (int)tm;
  p->total_nnz = te;
  p->user_stream = (hipStream_t)stream;
  p->m.assign(m, m + n_graphs);
  GD_HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  GD_HIP_TRY(hipStreamSynchronize(p->user_stream));
  int rc = p->off.alloc(4 * ((size_t)tm + 1));
  if (!rc) rc = p->col.alloc(4 * (size_t)std::max<int64_t>(te, 1));
  if (!rc) rc = p->deg.alloc(4 * (size_t)tm);
  if (!rc) rc = p->scores.alloc(4 * (size_t)tm);
  if (!rc) rc = p->c0.alloc(4 * (size_t)tm);
  if (!rc) rc = p->c1.alloc(4 * (size_t)tm);
  if (rc) return fail(rc);
  std::vector<Tile> tiles;
  std::vector<int32_t> tile_graph, hub_row, hub_slot, hub_nseg;
  std::vector<GraphState> gs(n_graphs);
  p->vbase.assign(n_graphs + 1, 0);
  int64_t ebase = 0;
  int vb = 0, slots = 0;
  for (int g = 0; g < n_graphs; g++) {
    rc = check_columns(d_col[g], nnz[g], m[g], p->stream);
    if (rc) return fail(rc);
    std::vector<int32_t> h_off(m[g] + 1);
    GD_HIP_TRY(hipMemcpyAsync(h_off.data(), d_off[g], 4 * ((size_t)m[g] + 1),
                              hipMemcpyDeviceToHost, p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    Schedule s;
    rc = build_schedule(m[g], nnz[g], h_off.data(), kBlock * kMgIPT, kTileRows, &s,
                        kBlock * kMgIPT);  // rows longer than a tile: whole-row exact hubs
    if (rc) return fail(rc);
    hipLaunchKernelGGL(mg_pack_kernel, dim3(grid_for(std::max<int64_t>(nnz[g], m[g]))),
                       dim3(256), 0, p->stream, d_off[g], d_col[g], d_outdeg[g], m[g], nnz[g], vb,
                       ebase, p->off.as<int32_t>(), p->col.as<int32_t>(), p->deg.as<int32_t>());
    GD_HIP_TRY(hipGetLastError());
    gs[g].tile_lo = (int)tiles.size();
    for (const Tile& t : s.tiles) {
      if (t.y >= 0)
        tiles.push_back(Tile{t.x + vb, t.y + vb, (int32_t)(t.z + ebase), (int32_t)(t.w + ebase)});
      else
        tiles.push_back(Tile{t.x + vb, -(-t.y - 1 + slots) - 1, (int32_t)(t.z + ebase),
                             (int32_t)(t.w + ebase)});
      tile_graph.push_back(g);
    }
    gs[g].tile_hi = (int)tiles.size();
    gs[g].hub_lo = (int)hub_row.size();
    for (size_t h = 0; h < s.hub_row.size(); h++) {
      hub_row.push_back(s.hub_row[h] + vb);
      hub_slot.push_back(s.hub_slot[h] + slots);
      hub_nseg.push_back(s.hub_nseg[h]);
    }
    gs[g].hub_hi = (int)hub_row.size();
    slots += s.n_slots;
    vb += m[g];
    ebase += nnz[g];
    p->vbase[g + 1] = vb;
  }
  const int32_t total_off = (int32_t)te;
  GD_HIP_TRY(hipMemcpyAsync(p->off.as<int32_t>() + tm, &total_off, 4, hipMemcpyHostToDevice,
                            p->stream));
  p->n_tiles = (int)tiles.size();
  p->e_max = (int)std::max<int64_t>(te - 1, 0);
  if (te == 0) GD_HIP_TRY(hipMemsetAsync(p->col.ptr, 0, 4, p->stream));
  p->h_state = gs;
  const int nh = (int)hub_row.size();
  rc = p->tiles.alloc(sizeof(Tile) * p->n_tiles);
  if (!rc) rc = p->tile_graph.alloc(4 * (size_t)p->n_tiles);
  if (!rc) rc = p->hub_row.alloc(4 * (size_t)std::max(1, nh));
  if (!rc) rc = p->hub_slot.alloc(4 * (size_t)std::max(1, nh));
  if (!rc) rc = p->hub_nseg.alloc(4 * (size_t)std::max(1, nh));
  if (!rc) rc = p->hub_part.alloc(4 * (size_t)std::max(1, slots));
  if (!rc) rc = p->err_part.alloc(8 * (size_t)p->n_tiles);
  if (!rc) rc = p->gs.alloc(sizeof(GraphState) * n_graphs);
  if (!rc) rc = p->d_vbase.alloc(4 * ((size_t)n_graphs + 1));
  if (!rc) rc = p->d_user_ptrs.alloc(sizeof(float*) * n_graphs);
  if (rc) return fail(rc);
  auto up = [&](DeviceBuffer& b, const void* src, size_t n) -> hipError_t {
    return n ? hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, p->stream) : hipSuccess;
  };
  GD_HIP_TRY(up(p->tiles, tiles.data(), sizeof(Tile) * p->n_tiles));
  GD_HIP_TRY(up(p->tile_graph, tile_graph.data(), 4 * (size_t)p->n_tiles));
  GD_HIP_TRY(up(p->hub_row, hub_row.data(), 4 * (size_t)nh));
  GD_HIP_TRY(up(p->hub_slot, hub_slot.data(), 4 * (size_t)nh));
  GD_HIP_TRY(up(p->hub_nseg, hub_nseg.data(), 4 * (size_t)nh));
  GD_HIP_TRY(up(p->d_vbase, p->vbase.data(), 4 * ((size_t)n_graphs + 1)));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  *out = owner.release();
  return GD_OK;
}

int gd_multi_pr_plan_destroy(gd_multi_pr_plan* p) {
  delete p;
  return GD_OK;
}

int gd_multi_pr_run(gd_multi_pr_plan* p, float* const* d_scores, const gd_pr_options* opts,
                    const gd_multi_pr_config* cfg, int* iterations, double* final_errors,
                    float* solve_ms) {
  GD_CHECK_ARG(p != nullptr && d_scores != nullptr, "NULL plan / scores");
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  gd_multi_pr_config c;
  if (cfg)
    c = *cfg;
  else
    gd_multi_pr_default_config(&c);
  GD_CHECK_ARG(o.max_iter >= 0, "max_iter must be >= 0");
  for (int g = 0; g < p->n_graphs; g++) GD_CHECK_ARG(d_scores[g] != nullptr, "NULL scores");
  const int G = p->n_graphs;
  if (p->errors_cap < std::max(1, o.max_iter)) {
    int rc = p->errors.alloc(8 * (size_t)G * std::max(1, o.max_iter));
    if (rc) return rc;
    p->errors_cap = std::max(1, o.max_iter);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
  }
  for (int g = 0; g < G; g++) p->h_state[g].base = (1.0f - o.damping) / (float)p->m[g];
  GD_HIP_TRY(hipStreamSynchronize(p->user_stream));  // caller's scores are ready
  GD_HIP_TRY(hipMemcpyAsync(p->gs.ptr, p->h_state.data(), sizeof(GraphState) * G,
                            hipMemcpyHostToDevice, p->stream));
  GD_HIP_TRY(hipMemcpyAsync(p->d_user_ptrs.ptr, d_scores, sizeof(float*) * G,
                            hipMemcpyHostToDevice, p->stream));
  const MgArgs a{p->tiles.as<Tile>(), p->tile_graph.as<int32_t>(), p->off.as<int32_t>(),
                 p->col.as<int32_t>(), p->e_max, p->deg.as<int32_t>(), p->scores.as<float>(),
                 p->c0.as<float>(), p->c1.as<float>(), p->hub_part.as<float>(),
                 p->err_part.as<double>(), p->gs.as<GraphState>(), p->hub_row.as<int32_t>(),
                 p->hub_slot.as<int32_t>(), p->hub_nseg.as<int32_t>(), p->errors.as<double>(),
                 std::max(1, o.max_iter), o.tol, o.damping};
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  GD_HIP_TRY(hipEventRecord(t0, p->stream));
  int maxm = 0;
  for (int g = 0; g < G; g++) maxm = std::max(maxm, p->m[g]);
  hipLaunchKernelGGL(mg_copy_scores_kernel, dim3(grid_for(maxm), G), dim3(256), 0, p->stream, G,
                     p->d_user_ptrs.as<float*>(), p->d_vbase.as<int32_t>(), p->scores.as<float>(),
                     1);
  hipLaunchKernelGGL(mg_init_kernel, dim3(grid_for(p->total_m)), dim3(256), 0, p->stream, a, G,
                     p->total_m, o.max_iter);
  GD_HIP_TRY(hipGetLastError());
  const int chunk = c.chunk > 0 ? c.chunk : 8;
  auto launch_chunk = [&]() {
    for (int k = 0; k < chunk; k++) {
      hipLaunchKernelGGL(mg_pull_kernel, dim3(p->n_tiles), dim3(kBlock), 0, p->stream, a);
      hipLaunchKernelGGL(mg_finalize_kernel, dim3(G), dim3(kBlock), 0, p->stream, a);
    }
  };
  if (c.use_graph && (!p->exec || p->graph_chunk != chunk || p->graph_max_iter != o.max_iter ||
                      p->graph_tol != o.tol || p->graph_damp != o.damping)) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    GD_HIP_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    launch_chunk();
    GD_HIP_TRY(hipStreamEndCapture(p->stream, &p->graph));
    GD_HIP_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    p->graph_chunk = chunk;
    p->graph_max_iter = o.max_iter;
    p->graph_tol = o.tol;
    p->graph_damp = o.damping;
  }
  std::vector<GraphState> h(G);
  int launched = 0;
  while (launched < o.max_iter) {
    if (c.use_graph)
      GD_HIP_TRY(hipGraphLaunch(p->exec, p->stream));
    else
      launch_chunk();
    GD_HIP_TRY(hipGetLastError());
    launched += chunk;
    GD_HIP_TRY(hipMemcpyAsync(h.data(), p->gs.ptr, sizeof(GraphState) * G, hipMemcpyDeviceToHost,
                              p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    bool all = true;
    for (int g = 0; g < G; g++) all = all && h[g].done;
    if (all) break;
  }
  hipLaunchKernelGGL(mg_copy_scores_kernel, dim3(grid_for(maxm), G), dim3(256), 0, p->stream, G,
                     p->d_user_ptrs.as<float*>(), p->d_vbase.as<int32_t>(), p->scores.as<float>(),
                     0);
  GD_HIP_TRY(hipEventRecord(t1, p->stream));
  GD_HIP_TRY(hipEventSynchronize(t1));
  float ms = 0.0f;
  GD_HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  if (solve_ms) *solve_ms = ms;
  std::vector<double> errs((size_t)G * std::max(1, o.max_iter));
  GD_HIP_TRY(hipMemcpy(errs.data(), p->errors.ptr, 8 * errs.size(), hipMemcpyDeviceToHost));
  for (int g = 0; g < G; g++) {
    if (iterations) iterations[g] = h[g].iter;
    if (final_errors)
      final_errors[g] = h[g].iter > 0 ? errs[(size_t)g * std::max(1, o.max_iter) + h[g].iter - 1]
                                      : 0.0;
  }
  return GD_OK;
}

}  // extern "C"
