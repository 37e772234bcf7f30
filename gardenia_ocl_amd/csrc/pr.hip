// pr.hip -- pull PageRank for gfx950 (src/pr/omp_base.cc:9-49 semantics).
//
// Per iteration ONE fused pull kernel (gather of contrib over the in-CSR,
// damping, new contrib = score / out-degree for the next iteration, per-tile
// L1 partial) plus a single-workgroup finalize kernel (hub-row epilogues from
// segment partials, deterministic L1 reduction, convergence flag).  contrib is
// double-buffered so the in-place score update never races the gather, and the
// iteration counter lives on the device, so every iteration's launches are
// identical and the host only checks convergence once per chunk.
//
// Arithmetic matches the reference bit for bit on rows of <= 64 in-edges:
// base + kDamp*sum is rounded twice (this file is built with
// -ffp-contract=off), the division is IEEE (correctly rounded), and the row sum
// is accumulated sequentially in CSR order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "pull_engine.h"
#include "relabel.h"

namespace gd {

constexpr int kFinBlocks = 128;  // finalize workgroups (<= kBlock: combined by one block)

struct PrState {
  int iter;         // iterations completed
  int done;         // set once converged or max_iter reached
  int max_iter;
  unsigned ticket;  // finalize workgroups arrived in this iteration
  double tol;
};

struct PrGather {
  const float* __restrict__ contrib;
  __device__ __forceinline__ float value(int /*e*/, int u) const { return contrib[u]; }
};

struct PrEpilogue {
  float* __restrict__ scores;
  float* __restrict__ contrib_next;
  const int32_t* __restrict__ outdeg;
  float base;
  float damp;
  int seed;
  struct Pre {
    float old;
    int deg;
  };
  __device__ __forceinline__ Pre pre(int v) const { return Pre{scores[v], outdeg[v]}; }
  __device__ __forceinline__ float init(const Pre&) const { return 0.0f; }
  // scores[v] = base + kDamp * sum (two roundings), next contrib, |change|
  // (src/pr/omp_base.cc:38-40, :25-27)
  __device__ __forceinline__ double finish(int v, float sum, const Pre& p) const {
    const float b = (seed < 0 || v == seed) ? base : 0.0f;
    const float nv = b + damp * sum;
    scores[v] = nv;
    contrib_next[v] = nv / (float)p.deg;
    return (double)fabsf(nv - p.old);
  }
};

__global__ __launch_bounds__(kBlock) void pr_init_kernel(int m, const float* __restrict__ scores,
                                                         const int32_t* __restrict__ outdeg,
                                                         float* __restrict__ contrib0,
                                                         PrState* __restrict__ st, int max_iter,
                                                         double tol) {
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < m; v += gridDim.x * kBlock)
    contrib0[v] = scores[v] / (float)outdeg[v];  // calc_contrib, src/pr/omp_base.cc:25-27
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter = 0;
    st->done = max_iter <= 0;
    st->max_iter = max_iter;
    st->ticket = 0;
    st->tol = tol;
  }
}

template <int IPT>
__global__ __launch_bounds__(kBlock) void pr_pull_kernel(
    const Tile* __restrict__ tiles, const int32_t* __restrict__ off,
    const int32_t* __restrict__ col, const int32_t* __restrict__ outdeg,
    float* __restrict__ scores, float* __restrict__ contrib0, float* __restrict__ contrib1,
    float* __restrict__ hub_part, double* __restrict__ err_part, const PrState* __restrict__ st,
    float base, float damp, int seed) {
  if (st->done) return;
  const int odd = st->iter & 1;
  __shared__ RowTileLds<IPT> lds;
  __shared__ double s_err[kWaves];
  const Tile t = tiles[blockIdx.x];
  const PrGather g{odd ? contrib1 : contrib0};
  double err = 0.0;
  if (t.y >= 0) {
    const PrEpilogue e{scores, odd ? contrib0 : contrib1, outdeg, base, damp, seed};
    err = row_tile<IPT>(off, col, t.x, t.y, t.z, t.w - t.z, g, e, lds);
  } else {
    const float p = segment_sum<IPT>(col, t.z, t.w - t.z, g, lds.val);
    if (threadIdx.x == 0) hub_part[-t.y - 1] = p;
  }
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) err_part[blockIdx.x] = tot;
}

// Hub-row epilogues (segment partials summed in segment order) and the L1
// reduction: each workgroup sums a contiguous slice of the per-tile partials,
// the last workgroup to arrive adds the workgroup sums in index order, so the
// result does not depend on scheduling.
__global__ __launch_bounds__(kBlock) void pr_finalize_kernel(
    const int32_t* __restrict__ hub_row, const int32_t* __restrict__ hub_slot,
    const int32_t* __restrict__ hub_nseg, int n_hubs, const float* __restrict__ hub_part,
    const double* __restrict__ err_part, int n_tiles, const int32_t* __restrict__ outdeg,
    float* __restrict__ scores, float* __restrict__ contrib0, float* __restrict__ contrib1,
    PrState* __restrict__ st, double* __restrict__ errors, double* __restrict__ fin_part,
    float base, float damp, int seed) {
  if (st->done) return;
  __shared__ double s_err[kWaves];
  __shared__ int s_last;
  const int odd = st->iter & 1;
  const PrEpilogue e{scores, odd ? contrib0 : contrib1, outdeg, base, damp, seed};
  double err = 0.0;
  for (int h = blockIdx.x * kBlock + threadIdx.x; h < n_hubs; h += gridDim.x * kBlock) {
    const int v = hub_row[h];
    const PrEpilogue::Pre p = e.pre(v);
    const float* part = hub_part + hub_slot[h];
    float s = 0.0f;
    for (int k = 0; k < hub_nseg[h]; k++) s += part[k];
    err += e.finish(v, s, p);
  }
  const int per = (n_tiles + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(lo + per, n_tiles);
  for (int i = lo + threadIdx.x; i < hi; i += 4 * kBlock) {
    const double a0 = err_part[i];
    const double a1 = (i + kBlock < hi) ? err_part[i + kBlock] : 0.0;
    const double a2 = (i + 2 * kBlock < hi) ? err_part[i + 2 * kBlock] : 0.0;
    const double a3 = (i + 3 * kBlock < hi) ? err_part[i + 3 * kBlock] : 0.0;
    err += ((a0 + a1) + a2) + a3;
  }
  const double tot = block_sum<kWaves>(err, s_err);
  if (threadIdx.x == 0) {
    fin_part[blockIdx.x] = tot;
    __threadfence();  // release fin_part before taking a ticket
    const unsigned arrived = atomicAdd(&st->ticket, 1u);
    s_last = (arrived == gridDim.x - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();  // acquire: every other workgroup's fin_part is visible
  const double v = (threadIdx.x < gridDim.x) ? fin_part[threadIdx.x] : 0.0;
  const double total = block_sum<kWaves>(v, s_err);
  if (threadIdx.x == 0) {
    const int it = st->iter;
    errors[it] = total;
    st->iter = it + 1;
    st->ticket = 0;
    if (total < st->tol || it + 1 >= st->max_iter) st->done = 1;
  }
}

}  // namespace gd

using namespace gd;

struct gd_pr_plan {
  int m = 0;
  int64_t nnz = 0;
  hipStream_t stream = nullptr;
  gd_pr_plan_config cfg{};
  bool relabel = false;
  int ipt = 8;
  // arrays the kernels read: the caller's, or the relabelled copies
  const int32_t* off = nullptr;
  const int32_t* col = nullptr;
  const int32_t* outdeg = nullptr;
  Relabel rl;
  DeviceBuffer s_int;  // scores in the internal order (relabel only)
  Schedule sched;
  int n_tiles = 0, n_hubs = 0;
  DeviceBuffer tiles, hub_row, hub_slot, hub_nseg, hub_part, err_part, fin_part, contrib0,
      contrib1, state, errors;
  int fin_blocks = 1;
  int errors_cap = 0;
  std::vector<hipEvent_t> ev;
  float plan_ms = 0.0f;
};

namespace {

int upload(DeviceBuffer& b, const void* src, size_t n, hipStream_t s) {
  if (!n) return GD_OK;
  GD_HIP_TRY(hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, s));
  return GD_OK;
}

int build_plan(gd_pr_plan* p, const int32_t* d_off, const int32_t* d_col,
               const int32_t* d_outdeg) {
  const int m = p->m;
  int rc = check_columns(d_col, p->nnz, m, p->stream);
  if (rc) return rc;
  p->off = d_off;
  p->col = d_col;
  p->outdeg = d_outdeg;
  if (p->relabel) {
    // offsets must be valid before the device walks them
    std::vector<int32_t> h_off(m + 1);
    GD_HIP_TRY(hipMemcpyAsync(h_off.data(), d_off, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost,
                              p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    Schedule tmp;
    rc = build_schedule(m, p->nnz, h_off.data(), kBlock * p->ipt, kTileRows, &tmp);
    if (rc) return rc;
    rc = build_relabel(m, p->nnz, d_off, d_col, d_outdeg, p->stream, &p->rl);
    if (rc) return rc;
    rc = p->s_int.alloc(4 * (size_t)m);
    if (rc) return rc;
    p->off = p->rl.off.as<int32_t>();
    p->col = p->rl.col.as<int32_t>();
    p->outdeg = p->rl.outdeg.as<int32_t>();
  }
  std::vector<int32_t> h_off(m + 1);
  GD_HIP_TRY(hipMemcpyAsync(h_off.data(), p->off, 4 * ((size_t)m + 1), hipMemcpyDeviceToHost,
                            p->stream));
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  rc = build_schedule(m, p->nnz, h_off.data(), kBlock * p->ipt, kTileRows, &p->sched);
  if (rc) return rc;
  p->n_tiles = (int)p->sched.tiles.size();
  p->n_hubs = (int)p->sched.hub_row.size();
  rc = p->tiles.alloc(sizeof(Tile) * p->n_tiles);
  if (!rc) rc = p->hub_row.alloc(4 * (size_t)std::max(1, p->n_hubs));
  if (!rc) rc = p->hub_slot.alloc(4 * (size_t)std::max(1, p->n_hubs));
  if (!rc) rc = p->hub_nseg.alloc(4 * (size_t)std::max(1, p->n_hubs));
  if (!rc) rc = p->hub_part.alloc(4 * (size_t)std::max(1, p->sched.n_slots));
  if (!rc) rc = p->err_part.alloc(8 * (size_t)std::max(1, p->n_tiles));
  if (!rc) rc = p->contrib0.alloc(4 * (size_t)m);
  if (!rc) rc = p->contrib1.alloc(4 * (size_t)m);
  if (!rc) rc = p->state.alloc(sizeof(PrState));
  if (!rc) rc = p->fin_part.alloc(8 * (size_t)kFinBlocks);
  if (rc) return rc;
  p->fin_blocks = std::max(1, std::min(kFinBlocks, std::max(p->n_tiles, p->n_hubs) / kBlock));
  rc = upload(p->tiles, p->sched.tiles.data(), sizeof(Tile) * p->n_tiles, p->stream);
  if (!rc) rc = upload(p->hub_row, p->sched.hub_row.data(), 4 * (size_t)p->n_hubs, p->stream);
  if (!rc) rc = upload(p->hub_slot, p->sched.hub_slot.data(), 4 * (size_t)p->n_hubs, p->stream);
  if (!rc) rc = upload(p->hub_nseg, p->sched.hub_nseg.data(), 4 * (size_t)p->n_hubs, p->stream);
  if (rc) return rc;
  GD_HIP_TRY(hipStreamSynchronize(p->stream));
  return GD_OK;
}

template <int IPT>
void launch_pull(gd_pr_plan* p, float* scores, float* c0, float* c1, PrState* st, float base,
                 float damp, int seed) {
  hipLaunchKernelGGL(pr_pull_kernel<IPT>, dim3(p->n_tiles), dim3(kBlock), 0, p->stream,
                     p->tiles.as<Tile>(), p->off, p->col, p->outdeg, scores, c0, c1,
                     p->hub_part.as<float>(), p->err_part.as<double>(), st, base, damp, seed);
}

}  // namespace

extern "C" {

int gd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void gd_pr_default_options(gd_pr_options* o) {
  o->damping = 0.85f;
  o->tol = 1e-3;
  o->max_iter = 30;
  o->seed_vertex = -1;
  o->chunk = 0;
  o->profile = 0;
}

void gd_pr_default_plan_config(gd_pr_plan_config* c) {
  c->relabel = -1;
  c->items_per_thread = 0;
}

int gd_pr_plan_create_ex(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                         const int32_t* d_outdeg, void* stream, const gd_pr_plan_config* cfg,
                         gd_pr_plan** out) {
  GD_CHECK_ARG(out != nullptr, "plan output pointer is NULL");
  *out = nullptr;
  GD_CHECK_ARG(m >= 0 && nnz >= 0 && nnz < (1ll << 31), "m/nnz out of range");
  GD_CHECK_ARG(m == 0 || (d_off && d_outdeg), "NULL device array");
  GD_CHECK_ARG(nnz == 0 || d_col, "NULL column indices");
  gd_pr_plan_config c;
  if (cfg)
    c = *cfg;
  else
    gd_pr_default_plan_config(&c);
  GD_CHECK_ARG(c.items_per_thread == 0 || c.items_per_thread == 8 || c.items_per_thread == 16,
               "items_per_thread must be 0, 8 or 16");
  GD_CHECK_ARG(c.relabel >= -1 && c.relabel <= 1, "relabel must be -1, 0 or 1");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  gd_pr_plan* p = new gd_pr_plan();
  p->m = m;
  p->nnz = nnz;
  p->stream = (hipStream_t)stream;
  p->cfg = c;
  p->ipt = c.items_per_thread ? c.items_per_thread : 8;
  p->relabel = (c.relabel == 1) || (c.relabel == -1 && nnz >= (1 << 20));
  int rc = GD_OK;
  if (m > 0) {
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess)
      (void)hipEventRecord(t0, p->stream);
    rc = build_plan(p, d_off, d_col, d_outdeg);
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

int gd_pr_plan_create(int m, int64_t nnz, const int32_t* d_off, const int32_t* d_col,
                      const int32_t* d_outdeg, void* stream, gd_pr_plan** out) {
  return gd_pr_plan_create_ex(m, nnz, d_off, d_col, d_outdeg, stream, nullptr, out);
}

int gd_pr_plan_info(const gd_pr_plan* p, gd_pr_plan_info_t* info) {
  GD_CHECK_ARG(p != nullptr && info != nullptr, "NULL plan / info");
  info->n_tiles = p->n_tiles;
  info->n_hubs = p->n_hubs;
  info->n_hub_segments = p->sched.n_slots;
  info->max_row_len = p->sched.max_row_len;
  info->relabel = p->relabel ? 1 : 0;
  info->items_per_thread = p->ipt;
  info->plan_ms = p->plan_ms;
  return GD_OK;
}

int gd_pr_plan_destroy(gd_pr_plan* p) {
  if (!p) return GD_OK;
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  delete p;
  return GD_OK;
}

int gd_pr_plan_run(gd_pr_plan* p, float* d_scores, const gd_pr_options* opts, gd_pr_stats* stats,
                   double* errors) {
  GD_CHECK_ARG(p != nullptr, "NULL plan");
  gd_pr_options o;
  if (opts)
    o = *opts;
  else
    gd_pr_default_options(&o);
  GD_CHECK_ARG(o.max_iter >= 0, "max_iter must be >= 0");
  GD_CHECK_ARG(o.seed_vertex >= -1 && o.seed_vertex < p->m, "seed_vertex out of range");
  GD_CHECK_ARG(p->m == 0 || d_scores != nullptr, "NULL scores");
  gd_pr_stats s;
  memset(&s, 0, sizeof(s));
  const int m = p->m;
  if (m == 0) {
    // Empty graph: the reference's loop body does nothing, error 0 < tol.
    if (o.max_iter > 0) {
      s.iterations = 1;
      s.converged = 0.0 < o.tol;
      if (errors) errors[0] = 0.0;
    }
    if (stats) *stats = s;
    return GD_OK;
  }
  if (p->errors_cap < std::max(1, o.max_iter)) {
    int rc = p->errors.alloc(sizeof(double) * std::max(1, o.max_iter));
    if (rc) return rc;
    p->errors_cap = std::max(1, o.max_iter);
  }
  const float damp = o.damping;
  const float base = (o.seed_vertex < 0) ? (1.0f - damp) / (float)m : (1.0f - damp);
  int seed = o.seed_vertex;
  if (p->relabel && seed >= 0) {
    GD_HIP_TRY(hipMemcpyAsync(&seed, p->rl.inv.as<int32_t>() + o.seed_vertex, 4,
                              hipMemcpyDeviceToHost, p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
  }
  const int chunk = o.chunk > 0 ? o.chunk : (o.tol > 0 ? 8 : std::max(1, o.max_iter));
  if (o.profile) {
    while ((int)p->ev.size() < 2 * chunk) {
      hipEvent_t e;
      GD_HIP_TRY(hipEventCreate(&e));
      p->ev.push_back(e);
    }
  }
  hipEvent_t t0, t1;
  GD_HIP_TRY(hipEventCreate(&t0));
  GD_HIP_TRY(hipEventCreate(&t1));
  PrState* st = p->state.as<PrState>();
  float* c0 = p->contrib0.as<float>();
  float* c1 = p->contrib1.as<float>();
  float* work = p->relabel ? p->s_int.as<float>() : d_scores;
  const int init_blocks = std::min((m + kBlock - 1) / kBlock, 8192);
  GD_HIP_TRY(hipEventRecord(t0, p->stream));
  if (p->relabel) {
    int rc = relabel_to_internal(p->rl, m, d_scores, work, p->stream);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(pr_init_kernel, dim3(init_blocks), dim3(kBlock), 0, p->stream, m, work,
                     p->outdeg, c0, st, o.max_iter, o.tol);
  GD_HIP_TRY(hipGetLastError());
  int launched = 0;
  PrState h{};
  float pull_ms = 0.0f;
  int pull_n = 0;
  while (launched < o.max_iter) {
    const int n = std::min(chunk, o.max_iter - launched);
    for (int k = 0; k < n; k++) {
      if (o.profile) GD_HIP_TRY(hipEventRecord(p->ev[2 * k], p->stream));
      if (p->ipt == 16)
        launch_pull<16>(p, work, c0, c1, st, base, damp, seed);
      else
        launch_pull<8>(p, work, c0, c1, st, base, damp, seed);
      if (o.profile) GD_HIP_TRY(hipEventRecord(p->ev[2 * k + 1], p->stream));
      hipLaunchKernelGGL(pr_finalize_kernel, dim3(p->fin_blocks), dim3(kBlock), 0, p->stream,
                         p->hub_row.as<int32_t>(), p->hub_slot.as<int32_t>(),
                         p->hub_nseg.as<int32_t>(), p->n_hubs, p->hub_part.as<float>(),
                         p->err_part.as<double>(), p->n_tiles, p->outdeg, work, c0, c1, st,
                         p->errors.as<double>(), p->fin_part.as<double>(), base, damp, seed);
    }
    GD_HIP_TRY(hipGetLastError());
    launched += n;
    GD_HIP_TRY(hipMemcpyAsync(&h, st, sizeof(PrState), hipMemcpyDeviceToHost, p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
    if (o.profile) {
      // Only launches that did work count; later ones in the chunk exit at once.
      const int worked = std::min(n, h.iter - (launched - n));
      for (int k = 0; k < worked; k++) {
        float ms = 0.0f;
        GD_HIP_TRY(hipEventElapsedTime(&ms, p->ev[2 * k], p->ev[2 * k + 1]));
        pull_ms += ms;
        pull_n++;
      }
    }
    if (h.done) break;
  }
  if (p->relabel) {
    int rc = relabel_to_user(p->rl, m, work, d_scores, p->stream);
    if (rc) return rc;
  }
  GD_HIP_TRY(hipEventRecord(t1, p->stream));
  GD_HIP_TRY(hipEventSynchronize(t1));
  GD_HIP_TRY(hipEventElapsedTime(&s.solve_ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  s.iterations = h.iter;
  s.pull_ms = pull_ms;
  s.pull_launches = pull_n;
  std::vector<double> errs(std::max(1, h.iter));
  if (h.iter > 0)
    GD_HIP_TRY(hipMemcpy(errs.data(), p->errors.ptr, sizeof(double) * h.iter,
                         hipMemcpyDeviceToHost));
  if (h.iter > 0) {
    s.final_error = errs[h.iter - 1];
    s.converged = s.final_error < o.tol;
    // Two correct FP32 evaluations of an iteration differ in L1 by up to about
    // one ulp per vertex, i.e. ~2^-22 * sum(scores) (sum(scores) <= 1 for
    // scores started from a distribution); flag a tie within that band.
    const double band = 1e-5 * o.tol + 2.4e-7;
    for (int i = std::max(0, h.iter - 2); i < h.iter; i++)
      if (o.tol > 0 && std::fabs(errs[i] - o.tol) <= band) s.near_tie = 1;
  }
  if (errors)
    for (int i = 0; i < h.iter; i++) errors[i] = errs[i];
  if (stats) *stats = s;
  return GD_OK;
}

int gd_pr_solve(int m, int nnz, const int32_t* in_off, const int32_t* in_col,
                const int32_t* /*out_off*/, const int32_t* /*out_col*/, const int32_t* degree,
                float* scores, const gd_pr_options* opts, gd_pr_stats* stats, double* errors) {
  GD_CHECK_ARG(m >= 0 && nnz >= 0, "negative size");
  GD_CHECK_ARG(m == 0 || (in_off && degree && scores), "NULL host array");
  GD_CHECK_ARG(nnz == 0 || in_col, "NULL column indices");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  DeviceBuffer d_off, d_col, d_deg, d_sc;
  int rc = d_off.alloc(sizeof(int32_t) * (m + 1));
  if (!rc) rc = d_col.alloc(sizeof(int32_t) * (size_t)nnz);
  if (!rc) rc = d_deg.alloc(sizeof(int32_t) * (size_t)m);
  if (!rc) rc = d_sc.alloc(sizeof(float) * (size_t)m);
  if (rc) return rc;
  if (m > 0) {
    GD_HIP_TRY(hipMemcpy(d_off.ptr, in_off, sizeof(int32_t) * (m + 1), hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(d_deg.ptr, degree, sizeof(int32_t) * m, hipMemcpyHostToDevice));
    GD_HIP_TRY(hipMemcpy(d_sc.ptr, scores, sizeof(float) * m, hipMemcpyHostToDevice));
  }
  if (nnz > 0)
    GD_HIP_TRY(hipMemcpy(d_col.ptr, in_col, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice));
  gd_pr_plan* plan = nullptr;
  rc = gd_pr_plan_create(m, nnz, d_off.as<int32_t>(), d_col.as<int32_t>(), d_deg.as<int32_t>(),
                         nullptr, &plan);
  if (rc) return rc;
  rc = gd_pr_plan_run(plan, d_sc.as<float>(), opts, stats, errors);
  gd_pr_plan_destroy(plan);
  if (rc) return rc;
  if (m > 0) GD_HIP_TRY(hipMemcpy(scores, d_sc.ptr, sizeof(float) * m, hipMemcpyDeviceToHost));
  return GD_OK;
}

}  // extern "C"
