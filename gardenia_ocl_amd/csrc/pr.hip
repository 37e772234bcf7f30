// pr.hip -- pull PageRank for gfx950 (src/pr/omp_base.cc:9-49 semantics).
//
// Per iteration one fused pull kernel (gather of contrib over the in-CSR,
// damping, new contrib = score / out-degree for the next iteration, per-tile
// L1 partial) and one finalize kernel (hub rows, deterministic L1 reduction,
// convergence flag); see pull_iter.h / iter_engine.h.  contrib is
// double-buffered so the in-place score update never races the gather.
//
// Arithmetic matches the reference bit for bit on every row: base + kDamp*sum
// is rounded twice (this file is built with -ffp-contract=off), the division
// is IEEE (correctly rounded), and every row sum is accumulated sequentially
// in CSR order (row tiles: the owner thread; longer rows: exact_hub_kernel).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gardenia_hip.h"
#include "gd_common.h"
#include "iter_engine.h"
#include "pr_policy.h"
#include "pull_engine.h"
#include "relabel.h"

namespace gd {

}  // namespace gd

namespace gd {

__global__ __launch_bounds__(kBlock) void pr_init_kernel(int m, const float* __restrict__ scores,
                                                         const int32_t* __restrict__ outdeg,
                                                         float* __restrict__ contrib0) {
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < m; v += gridDim.x * kBlock)
    contrib0[v] = scores[v] / (float)outdeg[v];  // calc_contrib, src/pr/omp_base.cc:25-27
}

}  // namespace gd

using namespace gd;

struct gd_pr_plan {
  int m = 0;
  int64_t nnz = 0;
  hipStream_t stream = nullptr;
  bool relabel = false;
  // arrays the kernels read: the caller's, or the relabelled copies
  const int32_t* outdeg = nullptr;
  Relabel rl;
  DeviceBuffer s_int;  // scores in the internal order (relabel only)
  DeviceBuffer contrib0, contrib1;
  IterEngine eng;
  float plan_ms = 0.0f;
};

namespace {

int build_plan(gd_pr_plan* p, const int32_t* d_off, const int32_t* d_col,
               const int32_t* d_outdeg) {
  const int m = p->m;
  int rc = check_columns(d_col, p->nnz, m, p->stream);
  if (rc) return rc;
  // offsets must be valid before any device walk (relabel, schedule)
  rc = check_offsets(d_off, m, p->nnz, p->stream);
  if (rc) return rc;
  const int32_t* off = d_off;
  const int32_t* col = d_col;
  p->outdeg = d_outdeg;
  if (p->relabel) {
    rc = build_relabel(m, p->nnz, d_off, d_col, d_outdeg, p->stream, &p->rl);
    if (!rc) rc = p->s_int.alloc(4 * (size_t)m);
    if (rc) return rc;
    off = p->rl.off.as<int32_t>();
    col = p->rl.col.as<int32_t>();
    p->outdeg = p->rl.outdeg.as<int32_t>();
  }
  // the relabelled copy of col is the plan's own: tiles are sorted in place
  p->eng.col_owned = p->relabel ? p->rl.col.as<int32_t>() : nullptr;
  rc = p->contrib0.alloc(4 * (size_t)m);
  if (!rc) rc = p->contrib1.alloc(4 * (size_t)m);
  if (!rc) rc = p->eng.build(m, p->nnz, off, col, p->stream);
  if (!rc) rc = p->eng.setup_grid<PrPolicy>();
  return rc;
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
  c->nontemporal = -1;
  c->persistent = -1;
  c->sort_tiles = -1;
  c->block = 0;
  c->hub_threshold = 0;
  c->stage_hubs = -1;
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
  GD_CHECK_ARG(c.nontemporal >= -1 && c.nontemporal <= 2, "nontemporal must be -1, 0, 1 or 2");
  GD_CHECK_ARG(c.persistent >= -1 && c.persistent <= 1, "persistent must be -1, 0 or 1");
  GD_CHECK_ARG(c.block == 0 || c.block == 256 || c.block == 512 || c.block == 1024,
               "block must be 0, 256, 512 or 1024");
  GD_CHECK_ARG(c.stage_hubs >= -1 && c.stage_hubs <= 1, "stage_hubs must be -1, 0 or 1");
  const int ipt = c.items_per_thread ? c.items_per_thread : 16;
  GD_CHECK_ARG(c.hub_threshold == 0 || (c.hub_threshold >= 64 && c.hub_threshold <= kBlock * ipt),
               "hub_threshold must be 0 or in [64, 256*items_per_thread]");
  if (gd_device_count() <= 0) {
    set_error("no HIP device visible");
    return GD_ERR_NO_DEVICE;
  }
  gd_pr_plan* p = new gd_pr_plan();
  p->m = m;
  p->nnz = nnz;
  p->stream = (hipStream_t)stream;
  p->relabel = (c.relabel == 1) || (c.relabel == -1 && nnz >= (1 << 20));
  // Defaults from RMAT-24 measurements on MI355X: 16 edges per thread, no
  // cache hints, one workgroup per tile (the pipelined persistent form needs
  // 86 VGPRs, which cuts occupancy to 5 workgroups per CU and runs slower).
  p->eng.ipt = ipt;
  p->eng.nt = c.nontemporal < 0 ? 0 : c.nontemporal;
  p->eng.persistent = c.persistent == 1;
  p->eng.sort_tiles = c.sort_tiles != 0;
  p->eng.block = (c.block > kBlock && !p->eng.persistent) ? c.block : kBlock;
  // Rows longer than hub_threshold (256*ipt by default) edges are whole-row
  // hubs summed in the reference's order and rounding (exact_hub_kernel);
  // shorter rows by their owner thread in a row tile: every score is the
  // reference's bit for bit (its sequential FP32 sum drifts by 3e-4 relative
  // on the 238,668-edge RMAT-24 row, so matching its rounding matters).
  p->eng.exact_len = c.hub_threshold ? c.hub_threshold : kBlock * ipt;
  p->eng.stage_hubs = c.stage_hubs;
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
  info->n_tiles = p->eng.n_tiles + p->eng.n_hub_tiles;
  info->n_hubs = p->eng.n_hubs;
  info->n_hub_segments = p->eng.sched.n_slots;
  info->max_row_len = p->eng.sched.max_row_len;
  info->relabel = p->relabel ? 1 : 0;
  info->items_per_thread = p->eng.ipt;
  info->nontemporal = p->eng.nt;
  info->persistent = p->eng.persistent ? 1 : 0;
  info->pull_grid = p->eng.pull_grid;
  info->plan_ms = p->plan_ms;
  info->sort_tiles = p->eng.sort_tiles ? 1 : 0;
  info->block = p->eng.block;
  info->hub_threshold = p->eng.exact_len;
  info->staged_hubs = p->eng.fused ? 0 : p->eng.n_staged_hubs;
  return GD_OK;
}

int gd_pr_plan_destroy(gd_pr_plan* p) {
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
    // Empty graph: every iteration of the reference's loop has error 0, so it
    // stops after the first when 0 < tol and otherwise runs max_iter of them.
    if (o.max_iter > 0) {
      s.converged = 0.0 < o.tol;
      s.iterations = s.converged ? 1 : o.max_iter;
      if (errors)
        for (int i = 0; i < s.iterations; i++) errors[i] = 0.0;
    }
    if (stats) *stats = s;
    return GD_OK;
  }
  const float damp = o.damping;
  const float base = (o.seed_vertex < 0) ? (1.0f - damp) / (float)m : (1.0f - damp);
  int seed = o.seed_vertex;
  if (p->relabel && seed >= 0) {
    GD_HIP_TRY(hipMemcpyAsync(&seed, p->rl.inv.as<int32_t>() + o.seed_vertex, 4,
                              hipMemcpyDeviceToHost, p->stream));
    GD_HIP_TRY(hipStreamSynchronize(p->stream));
  }
  EventPair evp;
  GD_HIP_TRY(evp.create());
  hipEvent_t t0 = evp.a, t1 = evp.b;
  float* work = p->relabel ? p->s_int.as<float>() : d_scores;
  GD_HIP_TRY(hipEventRecord(t0, p->stream));
  if (p->relabel) {
    int rc = relabel_to_internal(p->rl, m, d_scores, work, p->stream);
    if (rc) return rc;
  }
  const int init_blocks = std::min((m + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(pr_init_kernel, dim3(init_blocks), dim3(kBlock), 0, p->stream, m, work,
                     p->outdeg, p->contrib0.as<float>());
  GD_HIP_TRY(hipGetLastError());
  const PrPolicy::Args a{work, p->contrib0.as<float>(), p->contrib1.as<float>(), p->outdeg, base,
                         damp, seed};
  int rc = p->eng.run<PrPolicy>(a, o.max_iter, o.tol, o.chunk, o.profile != 0, &s, errors);
  if (rc) return rc;
  if (p->relabel) {
    rc = relabel_to_user(p->rl, m, work, d_scores, p->stream);
    if (rc) return rc;
  }
  GD_HIP_TRY(hipEventRecord(t1, p->stream));
  GD_HIP_TRY(hipEventSynchronize(t1));
  GD_HIP_TRY(hipEventElapsedTime(&s.solve_ms, t0, t1));
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
