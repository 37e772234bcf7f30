/* gardenia_hip.h -- C ABI of libgardenia_hip.so (MI355X / gfx950).
 *
 * Drop-in replacement for the device side of the reference's two solver
 * plug-ins:
 *   void PRSolver(int m, int nnz, int *in_row_offsets, int *in_column_indices,
 *                 int *out_row_offsets, int *out_column_indices, int *degree,
 *                 ScoreT *scores);                               src/pr/pr.h:23
 *   void SpmvSolver(int m, int nnz, int *Ap, int *Aj, ValueType *Ax,
 *                   ValueType *x, ValueType *y);                 src/spmv/spmv.h:29
 * Plain pointers and sizes only.  Every entry point returns a GD_* status
 * (gardenia_graph.h); gd_last_error() gives the text of the last failure on
 * the calling thread.  "d_" pointers are device (HBM) pointers, others host.
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 */
#ifndef GARDENIA_HIP_H_
#define GARDENIA_HIP_H_

#include <stdint.h>

#include "gardenia_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *gd_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int gd_device_count(void);

/* ------------------------------------------------------------ PageRank */

/* Run parameters.  Defaults (gd_pr_default_options) are the reference's
 * compile-time constants: kDamp 0.85f (include/common.h:18), EPSILON 1e-3 and
 * MAX_ITER 30 (src/pr/pr.h:4-5). */
typedef struct gd_pr_options {
  float damping;
  double tol;      /* stop after the first iteration whose L1 change < tol */
  int max_iter;
  int seed_vertex; /* -1: uniform teleport (1-d)/m as in the reference;
                      v >= 0: personalised, teleport mass (1-d) to v only */
  int chunk;       /* iterations launched between host convergence checks (0: auto) */
  int profile;     /* nonzero: time every pull-kernel launch with HIP events */
} gd_pr_options;

typedef struct gd_pr_stats {
  int iterations;       /* iterations executed (== reference's count when it converges) */
  int converged;        /* 1 if an iteration's L1 change fell below tol */
  double final_error;   /* L1 change of the last iteration */
  float solve_ms;       /* HIP-event time of the whole iteration loop */
  float pull_ms;        /* summed pull-kernel time (profile != 0), else 0 */
  int pull_launches;    /* pull kernels timed */
  int near_tie;         /* 1 if one of the last two L1 changes lies within the band
                           of tol where this library's and the reference's L1 values
                           can differ, so a stop decision could fall one iteration
                           apart: 1e-12 + 1e-12*tol for PageRank (every row summed in
                           the reference's order: the L1 values differ only in their
                           FP64 summation order), 1e-7 + 1e-6*tol for the adjoint
                           (tree-summed rows) */
} gd_pr_stats;

void gd_pr_default_options(gd_pr_options *opts);

typedef struct gd_pr_plan gd_pr_plan;

/* Plan-time choices (the autotuner's knobs). */
typedef struct gd_pr_plan_config {
  int relabel;          /* 1: relabel vertices by decreasing out-degree octave (original
                           order inside an octave) inside the plan
                           (hot sources become an L2-resident prefix of contrib; rows keep
                           their edge order; scores permuted in/out per run), 0: never,
                           -1: auto (when nnz >= 2^20) */
  int items_per_thread; /* edges per thread per tile: 8 (2048-edge tiles) or 16 (4096);
                           0: auto (16) */
  int nontemporal;      /* 1: streamed arrays (indices, offsets, scores, out-degree) use
                           non-temporal loads/stores so they do not evict the gathered
                           vector from L2; 2: the new contrib vector is written
                           non-temporally too; 0: plain; -1: auto (0) */
  int persistent;       /* 1: a persistent grid walks the tiles and prefetches the next
                           tile's loads during the current tile's reduction; 0: one
                           workgroup per tile; -1: auto (0) */
  int sort_tiles;       /* 1: each tile's edges are gathered in increasing source order
                           (lanes reading neighbouring sources share cache lines), each
                           value placed back at its CSR position in LDS: same results,
                           fewer L2 requests; costs 2 bytes per edge of plan memory and of
                           index stream; 0: CSR order; -1: auto (1) */
  int block;            /* threads per row-tile workgroup: 256 (tiles of 256*items_per_thread
                           edges), 512 or 1024 (two or four times larger tiles: with
                           sort_tiles the gathers of a tile are sorted over more edges, so
                           more lanes of a gather instruction share a cache line, at fewer
                           workgroups per CU); 0: auto (256); the persistent grid always
                           uses 256 */
  int hub_threshold;    /* rows with more in-edges than this are whole-row hubs summed by a
                           workgroup of their own (exact_hub_kernel), the rest by their
                           owner thread inside a row tile; 64 .. 256*items_per_thread;
                           0: auto (256*items_per_thread, the longest row a 256-thread
                           tile holds) */
  int stage_hubs;       /* 1: the hub rows are gathered by a kernel of their own, spread
                           over the GPU, before their
                           workgroups sum them (each hub workgroup is then bound by its scan
                           instead of its gathers); same results; 0: never; -1: auto (when
                           the longest hub row * 500 exceeds the edges in row tiles: the
                           longest row's chain would outlast the row tiles).  Iterations
                           small enough to run as one launch never stage */
} gd_pr_plan_config;

void gd_pr_default_plan_config(gd_pr_plan_config *cfg);

typedef struct gd_pr_plan_info_t {
  int64_t n_tiles;        /* workgroups per iteration: row tiles + whole-row hub tiles */
  int64_t n_hubs;         /* rows summed as hubs: longer than hub_threshold edges
                             (whole-row tiles in the reference's summation order:
                             one lane's chain over the first 4096 edges, the
                             workgroup's exact scan beyond) */
  int64_t n_hub_segments; /* tiles spent on hub rows */
  int32_t max_row_len;
  int32_t relabel;        /* 1 if the plan relabelled the vertices */
  int32_t items_per_thread;
  int32_t nontemporal;
  int32_t persistent;
  int32_t pull_grid;      /* workgroups per pull launch */
  float plan_ms;          /* plan construction time (validation, relabel, schedule) */
  int32_t sort_tiles;     /* 1 if tiles are gathered in source order */
  int32_t block;          /* threads per row-tile workgroup */
  int32_t hub_threshold;  /* rows longer than this are whole-row hubs */
  int32_t staged_hubs;    /* hub rows gathered ahead by the staging kernel (0: none) */
} gd_pr_plan_info_t;

/* Plan over a device-resident in-CSR (rows = destinations, columns = sources)
 * and the out-degree array, all borrowed (must outlive the plan).  Validates
 * the CSR (offsets, column range) and builds the tile schedule plus scratch
 * (2 contrib buffers, partials).  cfg may be NULL (defaults). */
int gd_pr_plan_create_ex(int m, int64_t nnz, const int32_t *d_in_row_offsets,
                         const int32_t *d_in_column_indices, const int32_t *d_out_degree,
                         void *stream, const gd_pr_plan_config *cfg, gd_pr_plan **plan);

/* gd_pr_plan_create_ex with the default configuration. */
int gd_pr_plan_create(int m, int64_t nnz, const int32_t *d_in_row_offsets,
                      const int32_t *d_in_column_indices, const int32_t *d_out_degree, void *stream,
                      gd_pr_plan **plan);

/* Pull PageRank on d_scores (in: initial scores, out: result), the loop of
 * src/pr/omp_base.cc:23-44 with one fused pull/damping/contrib/L1 kernel per
 * iteration.  errors (host, may be NULL) receives max_iter L1 values. */
int gd_pr_plan_run(gd_pr_plan *plan, float *d_scores, const gd_pr_options *opts,
                   gd_pr_stats *stats, double *errors);

int gd_pr_plan_info(const gd_pr_plan *plan, gd_pr_plan_info_t *info);

int gd_pr_plan_destroy(gd_pr_plan *plan);

/* fill_data() on the device (include/graph_io.h:21-130): CSR from an edge list
 * already in HBM (0-based src/dst, optional weights; file order), the flags of
 * gd_csr_from_edges.  Same result as the host builder, except that with
 * `sorted` duplicates of one edge are kept in push order (std::sort leaves
 * their order unspecified), which can only change the weight kept for them.
 * The library allocates the output arrays; copy them out with
 * gd_device_csr_copy and release them with gd_device_csr_free. */
typedef struct gd_device_csr {
  int32_t m;
  int64_t nnz;
  int32_t *row_offsets;    /* m+1, device */
  int32_t *column_indices; /* nnz, device */
  float *weight;           /* nnz, device (1.0 when no weights were given) */
  int32_t *degree;         /* m, device: row lengths */
} gd_device_csr;

int gd_csr_from_edges_device(int m, int64_t n, const int32_t *d_src, const int32_t *d_dst,
                             const float *d_weight, int symmetrize, int transpose, int sorted,
                             int remove_selfloops, int remove_redundents, int mirror_selfloops,
                             void *stream, gd_device_csr *out);
/* Device-to-device copies into caller arrays (any may be NULL), on stream. */
int gd_device_csr_copy(const gd_device_csr *c, int32_t *d_row_offsets, int32_t *d_column_indices,
                       float *d_weight, int32_t *d_degree, void *stream);
int gd_device_csr_free(gd_device_csr *c);

/* Push-style PageRank over the OUT-CSR: the reference's scatter variant
 * (src/pr/scatter.cu:11-41 and its host loop :57-78), a cross-check of the pull
 * path.  Every iteration sets next[v] = (1-d)/m, adds kDamp*score[u]/outdeg[u]
 * (outdeg[u] = u's out-CSR row length) over every out-edge (u, v), sums
 * |next - score| in FP64 in a fixed order, and sets score = next; it continues
 * while that change is > tol (at least one iteration, the reference's
 * do-while).  The additions are propagation-blocked: each edge's value is
 * binned by destination window (32,768 vertices) and summed per window in LDS
 * with FP32 atomic adds (addition order: the hardware's, as with the
 * reference's atomicAdd), so scores agree with the pull path to FP32
 * summation-order noise.  Each call first lays out the edges by window (a
 * stable radix sort; not in stats->solve_ms) and holds device memory of about
 * 36 bytes per edge while it does so, 16 during the iterations.  d_scores (float[m], HBM) holds the initial
 * scores and receives the result; opts->seed_vertex must be -1; errors (host,
 * room for max(1, max_iter) values, may be NULL) receives each iteration's
 * change.  The out-CSR is validated on the device first. */
int gd_pr_push_solve(int m, int64_t nnz, const int32_t *d_out_row_offsets,
                     const int32_t *d_out_column_indices, float *d_scores,
                     const gd_pr_options *opts, void *stream, gd_pr_stats *stats,
                     double *errors);

/* The reference PRVerifier's residual (src/pr/verifier.cc:10-69, the push part
 * :50-68) on the device: one push iteration of the tested scores over the
 * OUT-CSR (d_degree = out-degree), residual = sum_v |base + d*sums_v -
 * scores_v| in FP64.  Each per-vertex term is bit-identical to the
 * reference's (sources added in increasing order with sequential FP32
 * rounding); only the FP64 sum runs in a different fixed order.  The
 * reference prints "Correct" when residual < EPSILON.  Synchronises stream. */
int gd_pr_verify(int m, int64_t nnz, const int32_t *d_out_row_offsets,
                 const int32_t *d_out_column_indices, const int32_t *d_degree,
                 const float *d_scores, float damping, void *stream, double *residual);

/* One-shot form of PRSolver with host arrays (src/pr/pr.h:23): copies the
 * in-CSR, degree and scores to HBM, runs, copies scores back, frees.  The
 * out-CSR arguments are accepted and unused, as in every pull variant of the
 * reference. */
int gd_pr_solve(int m, int nnz, const int32_t *in_row_offsets, const int32_t *in_column_indices,
                const int32_t *out_row_offsets, const int32_t *out_column_indices,
                const int32_t *degree, float *scores, const gd_pr_options *opts,
                gd_pr_stats *stats, double *errors);

/* Pull PageRank that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * iteration, options, stats and L1 log as gd_pr_plan_run over the same
 * device-resident in-CSR, out-degree and scores, with every row summed in
 * CSR order (the reference's sequential FP32 sum) and no plan: the contrib
 * vectors, L1 partials, L1 log and iteration state live in a static device
 * array of about 16 MB per GPU, held until the process ends.  Limits:
 * m <= GD_NOALLOC_MAX_VERTICES, max_iter <= GD_NOALLOC_MAX_ITER (else
 * GD_ERR_INVALID, before any device work).  The CSR is validated on the device
 * first (GD_ERR_INVALID for bad offsets or columns).  Calls that overlap on
 * one GPU share the array and race.  Synchronises `stream`; stats->solve_ms
 * is host wall time of the call. */
#define GD_NOALLOC_MAX_VERTICES (1 << 21)
#define GD_NOALLOC_MAX_ITER 4096
int gd_pr_run_noalloc(int m, int64_t nnz, const int32_t *d_in_row_offsets,
                      const int32_t *d_in_column_indices, const int32_t *d_out_degree,
                      float *d_scores, const gd_pr_options *opts, void *stream,
                      gd_pr_stats *stats, double *errors);

/* ------------------------------------------------------------ batched personalised PageRank */

/* k personalised PageRank vectors at once (no counterpart in the reference,
 * which has no PPR): column j teleports its (1-d) mass to seeds[j].  Scores
 * are vertex-major, X[v*k + j], so each gathered source index feeds k
 * additions (an SpMM-style pull).  Same iteration as gd_pr_plan_run with a
 * per-column teleport; a run stops when every column's L1 change is < tol. */
typedef struct gd_ppr_plan gd_ppr_plan;

typedef struct gd_ppr_plan_config {
  int relabel; /* as gd_pr_plan_config.relabel */
  int k_chunk; /* columns per wavefront pass: 64, 128 or 256 (one 256*k_chunk/64-byte
                  gather per edge); 0: the largest that divides k */
  int column_stop; /* 1 (default, -1): each column stops after its first iteration whose
                      L1 change is below tol, as k independent runs would; 0: every
                      column iterates until the largest L1 change is below tol */
  int compact;     /* 1 (default, -1; needs k <= 4096): stopped columns are moved behind
                      the running ones between host checks, so their lanes stop gathering;
                      0: stopped columns stay in place (masked) */
  int hub_cols;    /* columns per workgroup of a hub row (> 4096 in-edges, added in edge
                      order per column): 16, 32 or 64; narrower groups keep more of the
                      row's edges in flight per workgroup, wider ones gather each
                      edge's columns in fewer requests; 0: auto (64, the fastest of
                      the three for RMAT-22 at k = 256) */
} gd_ppr_plan_config;

void gd_ppr_default_plan_config(gd_ppr_plan_config *cfg);

typedef struct gd_ppr_plan_info_t {
  int64_t n_tiles;
  int64_t n_hubs;
  int32_t k;
  int32_t k_chunk;
  int32_t relabel;
  int32_t grid; /* persistent workgroups per column chunk */
  float plan_ms;
  int32_t hub_cols;
} gd_ppr_plan_info_t;

/* k must be a positive multiple of 64.  Arrays are borrowed, as for gd_pr_plan_create. */
int gd_ppr_plan_create(int m, int64_t nnz, const int32_t *d_in_row_offsets,
                       const int32_t *d_in_column_indices, const int32_t *d_out_degree, int k,
                       void *stream, const gd_ppr_plan_config *cfg, gd_ppr_plan **plan);

/* d_seeds: k device int32 vertex ids; d_X: m*k device floats, in: initial
 * scores (typically one-hot at the seeds), out: result.  errors (host, may be
 * NULL) gets, per iteration, the max L1 change over the columns still running;
 * col_errors (host, may be NULL) each column's L1 change in its last iteration
 * (k values).  stats->iterations is the longest column's count.
 * opts->seed_vertex is ignored. */
int gd_ppr_plan_run(gd_ppr_plan *plan, const int32_t *d_seeds, float *d_X,
                    const gd_pr_options *opts, gd_pr_stats *stats, double *errors,
                    double *col_errors);

int gd_ppr_plan_info(const gd_ppr_plan *plan, gd_ppr_plan_info_t *info);

/* Iterations each column ran in the last gd_ppr_plan_run (k host int32, user
 * column order). */
int gd_ppr_plan_column_iterations(const gd_ppr_plan *plan, int32_t *iterations);

int gd_ppr_plan_destroy(gd_ppr_plan *plan);

/* ------------------------------------------------------------ SpMV */

typedef struct gd_spmv_plan gd_spmv_plan;

/* Plan over a device-resident CSR matrix (num_rows x num_cols), borrowed;
 * the tile mode (see gd_spmv_plan_create_ex for the others). */
int gd_spmv_plan_create(int num_rows, int num_cols, int64_t nnz, const int32_t *d_Ap,
                        const int32_t *d_Aj, const float *d_Ax, void *stream, gd_spmv_plan **plan);

/* y <- y + A x (src/spmv/omp_base.cc:17-28).  Every row is summed
 * sequentially in CSR order from y[row], bit-identical to the reference: rows
 * of up to 2048 nonzeros by one thread of an LDS row tile, longer rows by one
 * lane of a workgroup that gathers the row chunk by chunk. */
int gd_spmv_plan_run(gd_spmv_plan *plan, const float *d_x, float *d_y);

/* n_tiles: LDS row tiles; n_hubs: rows of more than 2048 nonzeros (one
 * workgroup each). */
int gd_spmv_plan_info(const gd_spmv_plan *plan, int64_t *n_tiles, int64_t *n_hubs,
                      int32_t *max_row_len);

int gd_spmv_plan_destroy(gd_spmv_plan *plan);

/* SpMV execution modes.
 *   GD_SPMV_TILES   the tile engine over the caller's CSR (any matrix, any
 *                   row order; bit-identical to the reference).
 *   GD_SPMV_SLICED  columns split into 8 slices over a slice-major copy of the
 *                   matrix; one tile-engine pass per slice, slice 0 first,
 *                   adds each row's slice entries onto y in column order (the
 *                   pass's x gathers stay in one slice, which every XCD's L2
 *                   holds).  Needs sorted columns per row; bit-identical to
 *                   the reference for every row.  The copy holds the values:
 *                   call gd_spmv_plan_refresh_values after changing d_Ax.
 *   GD_SPMV_AUTO    SLICED when x spans more than one XCD's L2 (cols*4 > 4 MB),
 *                   every row has <= 64 nonzeros and the columns are sorted,
 *                   else TILES. */
enum { GD_SPMV_AUTO = 0, GD_SPMV_TILES = 1, GD_SPMV_SLICED = 2 };

int gd_spmv_plan_create_ex(int num_rows, int num_cols, int64_t nnz, const int32_t *d_Ap,
                           const int32_t *d_Aj, const float *d_Ax, void *stream, int mode,
                           gd_spmv_plan **plan);

/* Mode the plan runs in (GD_SPMV_TILES or GD_SPMV_SLICED), -1 for NULL. */
int gd_spmv_plan_mode(const gd_spmv_plan *plan);

/* Re-read d_Ax into the plan's sliced copy (no-op for the tile mode). */
int gd_spmv_plan_refresh_values(gd_spmv_plan *plan);

/* One-shot form of SpmvSolver with host arrays (src/spmv/spmv.h:29); mode AUTO. */
int gd_spmv_solve(int m, int nnz, const int32_t *Ap, const int32_t *Aj, const float *Ax,
                  const float *x, float *y, float *kernel_ms);

/* ------------------------------------------------------------ many graphs, one launch */

/* PageRank on many independent graphs at once (no counterpart in the
 * reference, which solves one graph per process).  The graphs are packed into
 * one block-diagonal union at plan time; each keeps its own tiles, base score,
 * iteration count and convergence test, so every graph's scores and iteration
 * count equal a single-graph plan's (same tiles, no relabel).  One pull launch
 * per iteration covers all graphs; a chunk of iterations is captured into a
 * hipGraph once and replayed until every graph has converged.  Work runs on a
 * private stream; the calls are synchronous. */
typedef struct gd_multi_pr_plan gd_multi_pr_plan;

typedef struct gd_multi_pr_config {
  int use_graph; /* 1: replay a captured hipGraph per chunk; 0: launch eagerly */
  int chunk;     /* iterations per chunk / captured graph (0: 8) */
} gd_multi_pr_config;

void gd_multi_pr_default_config(gd_multi_pr_config *cfg);

/* Per graph g: m[g] vertices, nnz[g] edges, device in-CSR and out-degree
 * (borrowed only during the call: the plan keeps its own packed copy). */
int gd_multi_pr_plan_create(int n_graphs, const int32_t *m, const int64_t *nnz,
                            const int32_t *const *d_in_row_offsets,
                            const int32_t *const *d_in_column_indices,
                            const int32_t *const *d_out_degree, void *stream,
                            gd_multi_pr_plan **plan);

/* d_scores[g]: device scores of graph g, in/out.  iterations / final_errors
 * (host, n_graphs each, may be NULL); solve_ms (may be NULL). */
int gd_multi_pr_run(gd_multi_pr_plan *plan, float *const *d_scores, const gd_pr_options *opts,
                    const gd_multi_pr_config *cfg, int *iterations, double *final_errors,
                    float *solve_ms);

int gd_multi_pr_plan_destroy(gd_multi_pr_plan *plan);

/* ------------------------------------------------------------ differentiation */

/* Weighted PageRank and its implicit adjoint (no counterpart in the reference,
 * whose PageRank is unweighted).  Raw edge weights a (in-CSR order) are
 * normalised per source: P[v,u] = a_uv / s_u, s_u = sum_v a_uv, and
 *     x = (1-d)/m + d P x            (a == 1 gives the reference iteration).
 * For a loss with dL/dx = w the backward pass iterates the adjoint
 *     lambda = w + d P^T lambda
 * with the same pull engine over a CSC built once per plan, then writes
 *     grad[e] = dL/da[e] = d (x_u / s_u) (lambda_v - (P^T lambda)_u)
 * for every in-CSR edge e = (u -> v).  O(V + E) memory, no stored iterates.
 * The weight array is borrowed and re-read by every forward / backward call
 * (it may change between calls, e.g. during training). */
typedef struct gd_wpr_plan gd_wpr_plan;

int gd_wpr_plan_create(int m, int64_t nnz, const int32_t *d_in_row_offsets,
                       const int32_t *d_in_column_indices, const float *d_weight, void *stream,
                       gd_wpr_plan **plan);

/* d_x: in initial scores (e.g. 1/m), out the fixed point. */
int gd_wpr_forward(gd_wpr_plan *plan, float *d_x, const gd_pr_options *opts, gd_pr_stats *stats,
                   double *errors);

/* d_x: the forward solution; d_w: dL/dx; out: d_lambda (m), d_grad (nnz, in-CSR order). */
int gd_wpr_backward(gd_wpr_plan *plan, const float *d_x, const float *d_w, float *d_lambda,
                    float *d_grad, const gd_pr_options *opts, gd_pr_stats *stats);

int gd_wpr_plan_destroy(gd_wpr_plan *plan);

/* Backward of y = A x (+ y0): dx = A^T g (the SpMV engine over a CSC built once
 * per plan, values re-permuted from d_Ax on each call) and
 * dA[k] = g[row(k)] * x[col(k)] in CSR order.  Either output may be NULL. */
typedef struct gd_spmv_bwd_plan gd_spmv_bwd_plan;

int gd_spmv_bwd_plan_create(int num_rows, int num_cols, int64_t nnz, const int32_t *d_Ap,
                            const int32_t *d_Aj, void *stream, gd_spmv_bwd_plan **plan);

int gd_spmv_bwd_run(gd_spmv_bwd_plan *plan, const float *d_Ax, const float *d_x, const float *d_g,
                    float *d_dx, float *d_dA);

int gd_spmv_bwd_plan_destroy(gd_spmv_bwd_plan *plan);

/* ---- SymGS: multicolor symmetric Gauss-Seidel sweep ----------------------
 * Replaces SymGSSolver (src/symgs/symgs.h:28; omp_base.cc:27-42).  One sweep
 * = colors 0..C-1 then C-1..0; each row r of color c (r = indices[k],
 * color_offsets[c] <= k < color_offsets[c+1]) gets
 *   x[r] = (b[r] - sum_{j != r} x[j] * A[r][j]) / A[r][r]   unless A[r][r] == 0,
 * the sum in CSR order with sequential FP32 rounding and A[r][r] the row's last
 * diagonal entry: bit-exact with the reference's sequential sweep
 * (src/symgs/verifier.cc:50-77).  The coloring must be proper (no off-diagonal
 * entry joins two rows of one color) and list each row at most once; the plan
 * checks both.  Device arrays: d_Ap[m+1], d_Aj/d_Ax[nnz], d_indices (the
 * entries color_offsets[0] .. color_offsets[n_colors]-1 are read); host
 * color_offsets[n_colors+1] within [0, m]. */
typedef struct gd_symgs_plan gd_symgs_plan;

typedef struct gd_symgs_plan_info_t {
  int32_t rows;         /* rows swept: color_offsets[C] - color_offsets[0] */
  int32_t n_colors;
  int64_t offdiag_nnz;  /* off-diagonal entries of the swept rows */
  int32_t n_tiles;      /* LDS row tiles over all colors */
  int32_t n_exact_rows; /* rows of more than 64 off-diagonal entries */
  int32_t max_row_len;
  double plan_ms;
} gd_symgs_plan_info_t;

int gd_symgs_plan_create(int m, int64_t nnz, const int32_t *d_Ap, const int32_t *d_Aj,
                         const float *d_Ax, const int32_t *d_indices,
                         const int32_t *color_offsets, int n_colors, void *stream,
                         gd_symgs_plan **plan);

/* `sweeps` symmetric sweeps in place on d_x (m floats), right-hand side d_b. */
int gd_symgs_plan_run(gd_symgs_plan *plan, float *d_x, const float *d_b, int sweeps, void *stream);

int gd_symgs_plan_info(const gd_symgs_plan *plan, gd_symgs_plan_info_t *info);

int gd_symgs_plan_destroy(gd_symgs_plan *plan);

/* SymGSSolver semantics on host arrays (one sweep, x updated in place);
 * kernel_ms (may be NULL) receives the device time of the sweep. */
int gd_symgs_solve(int m, int nnz, const int32_t *Ap, const int32_t *Aj, const int32_t *indices,
                   const float *Ax, float *x, const float *b, const int32_t *color_offsets,
                   int n_colors, float *kernel_ms);

#ifdef __cplusplus
}
#endif
#endif /* GARDENIA_HIP_H_ */
