// oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's OpenMP PageRank / SpMV path, used as the
// checker for the HIP library.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library; the product path never
// calls it and has no CPU fallback.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root).  Built with `g++ -O3 -fopenmp
// -ffp-contract=off` (the reference builds with `-O3 -Wall -fopenmp`,
// src/common.mk:15, on baseline x86-64 where no FMA contraction can happen;
// -ffp-contract=off pins that even if -march flags are ever added).
//
// Pins (see tests/test_oracle_*.py):
//   orc_csr_*         golden CSR arrays from the reference's own read_graph()
//                     compiled from source (tests/golden/), plus invariants.
//   orc_pr            golden scores / iteration counts / per-iteration L1 from
//                     the reference's own PRSolver compiled from source, plus
//                     the closed-form fixed point of datasets/test_pr.mtx.
//   orc_spmv          golden y from the reference's SpmvSolver, plus the
//                     hand-computed nvGRAPH 5x5 example.
//   orc_pr_residual   golden residual from the reference PRVerifier.
//   orc_pr_push       the pull oracle's scores and iteration counts to FP32
//                     summation-order noise, and the FP64 fixed point.
//   orc_max_rel_err   golden values on crafted vectors.
//   orc_pr_fp64       dense LAPACK solve of (I - d A^T D^-1) x = (1-d)/m.
//   orc_ppr_fp64      linearity + dense solve on small graphs.
//   orc_wpr_*         central finite differences + dense solve.
//   orc_spmv_bwd_fp64 dense matrix transpose on small inputs.
//   orc_symgs         golden x from the reference's SymGSSolver (omp_base.cc)
//                     and verdicts of its SymGSVerifier, compiled from source;
//                     hand-computed 3x3 sweeps; many sweeps on an SPD system
//                     converge to the dense solve.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>
#include <omp.h>

namespace {

struct Edge {  // include/graph_io.h:14-17
  int dst;
  float wt;
};

bool by_dst(Edge a, Edge b) { return a.dst < b.dst; }  // graph_io.h:19

struct Csr {
  int m = 0;
  int nnz = 0;
  std::vector<int> row_offsets;
  std::vector<int> column_indices;
  std::vector<float> weight;
  std::vector<int> degree;
};

// include/graph_io.h:21-130 (fill_data) + :289-293 (degree).
void fill_csr(int m, std::vector<std::vector<Edge>>& vertices, bool sorted,
              bool remove_selfloops, bool remove_redundents, Csr* out) {
  if (sorted)  // :23-29  (std::sort, unstable, same comparator)
    for (int i = 0; i < m; i++)
      std::sort(vertices[i].begin(), vertices[i].end(), by_dst);
  if (remove_selfloops)  // :32-45  (order-preserving erase)
    for (int i = 0; i < m; i++) {
      auto& v = vertices[i];
      v.erase(std::remove_if(v.begin(), v.end(), [i](const Edge& e) { return e.dst == i; }),
              v.end());
    }
  if (remove_redundents)  // :48-61  only ADJACENT duplicates, first kept
    for (int i = 0; i < m; i++) {
      auto& v = vertices[i];
      v.erase(std::unique(v.begin(), v.end(),
                          [](const Edge& a, const Edge& b) { return a.dst == b.dst; }),
              v.end());
    }
  out->m = m;
  out->row_offsets.assign(m + 1, 0);  // :72-78
  long long count = 0;
  for (int i = 0; i < m; i++) {
    out->row_offsets[i] = (int)count;
    count += (long long)vertices[i].size();
  }
  out->row_offsets[m] = (int)count;
  out->nnz = (int)count;
  out->column_indices.resize(count);  // :105-116
  out->weight.resize(count);
  long long k = 0;
  for (int i = 0; i < m; i++)
    for (const Edge& e : vertices[i]) {
      out->column_indices[k] = e.dst;
      out->weight[k] = e.wt;
      k++;
    }
  out->degree.resize(m);  // :289-293
  for (int i = 0; i < m; i++) out->degree[i] = out->row_offsets[i + 1] - out->row_offsets[i];
}

// Edge placement of mtx2csr (graph_io.h:238-258) and gr2csr/graph2csr
// (:152-172, :189-208).  Edges are 0-based, in file order.  mtx mirrors only
// non-self-loop edges when symmetrizing (:246); gr/graph mirror all (:160, :196).
// `transpose` is forced off by the first mirrored edge (:249, :163, :199).
void place_edges(int m, long long n, const int* src, const int* dst, const float* wt,
                 bool symmetrize, bool transpose, bool mirror_selfloops,
                 std::vector<std::vector<Edge>>* vertices) {
  vertices->assign(m, std::vector<Edge>());
  bool tr = transpose;
  for (long long i = 0; i < n; i++) {
    int s = src[i], d = dst[i];
    float w = wt ? wt[i] : 1.0f;
    if (symmetrize && (mirror_selfloops || s != d)) {
      (*vertices)[d].push_back(Edge{s, w});
      tr = false;
    }
    if (!tr)
      (*vertices)[s].push_back(Edge{d, w});
    else
      (*vertices)[d].push_back(Edge{s, w});
  }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- CSR build

void* orc_csr_from_edges(int m, long long n, const int* src, const int* dst, const float* wt,
                         int symmetrize, int transpose, int sorted, int remove_selfloops,
                         int remove_redundents, int mirror_selfloops) {
  std::vector<std::vector<Edge>> vertices;
  place_edges(m, n, src, dst, wt, symmetrize, transpose, mirror_selfloops, &vertices);
  Csr* c = new Csr();
  fill_csr(m, vertices, sorted, remove_selfloops, remove_redundents, c);
  return c;
}

// mtx2csr (graph_io.h:216-261): skip '%' lines, read "m n nnz", then nnz lines
// "src dst [wt]" 1-based; |wt| kept (:242); wt defaults to 1 (:241).
void* orc_csr_from_mtx(const char* path, int symmetrize, int transpose, int sorted,
                       int remove_selfloops, int remove_redundents) {
  std::ifstream f(path);
  if (!f) return nullptr;
  std::string line;
  std::getline(f, line);
  while (!line.empty() && line[0] == '%') std::getline(f, line);
  int m = 0, n = 0, nnz = 0;
  if (sscanf(line.c_str(), "%d %d %d", &m, &n, &nnz) != 3) return nullptr;
  std::vector<int> s(nnz), d(nnz);
  std::vector<float> w(nnz);
  float wt = 1.0f;
  for (int i = 0; i < nnz; i++) {
    std::getline(f, line);
    int a = 0, b = 0;
    int num = sscanf(line.c_str(), "%d %d %f", &a, &b, &wt);
    if (num == 2) wt = 1;
    if (wt < 0) wt = -wt;
    s[i] = a - 1;
    d[i] = b - 1;
    w[i] = wt;
  }
  return orc_csr_from_edges(m, nnz, s.data(), d.data(), w.data(), symmetrize, transpose, sorted,
                            remove_selfloops, remove_redundents, 0);
}

void orc_csr_info(void* h, int* m, int* nnz) {
  Csr* c = (Csr*)h;
  *m = c->m;
  *nnz = c->nnz;
}

void orc_csr_copy(void* h, int* row_offsets, int* column_indices, float* weight, int* degree) {
  Csr* c = (Csr*)h;
  if (row_offsets) memcpy(row_offsets, c->row_offsets.data(), sizeof(int) * (c->m + 1));
  if (column_indices) memcpy(column_indices, c->column_indices.data(), sizeof(int) * c->nnz);
  if (weight) memcpy(weight, c->weight.data(), sizeof(float) * c->nnz);
  if (degree) memcpy(degree, c->degree.data(), sizeof(int) * c->m);
}

void orc_csr_free(void* h) { delete (Csr*)h; }

// ---------------------------------------------------------------- PageRank

// src/pr/omp_base.cc:9-49.  Pull Jacobi PR, FP32 scores, FP64 L1 error,
// dangling mass dropped.  `damp` is kDamp (include/common.h:18, 0.85f).
// Returns the number of iterations executed; errors[i] gets iteration i's L1.
// (The reference prints iter+1, which equals this count except when
// max_iter is hit without convergence, where it prints max_iter+1.)
int orc_pr(int m, const int* row_offsets, const int* column_indices, const int* degree,
           float* scores, float damp, double tol, int max_iter, double* errors) {
  const float base_score = (1.0f - damp) / m;  // omp_base.cc:16
  std::vector<float> contrib(m);
  int iter;
  for (iter = 0; iter < max_iter; iter++) {  // :23
    double error = 0;
#pragma omp parallel for
    for (int n = 0; n < m; n++) contrib[n] = scores[n] / degree[n];  // :25-27
#pragma omp parallel for reduction(+ : error) schedule(dynamic, 64)
    for (int v = 0; v < m; v++) {  // :28-41
      float incoming_total = 0;
      for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++)
        incoming_total += contrib[column_indices[e]];
      float old_score = scores[v];
      scores[v] = base_score + damp * incoming_total;
      error += std::fabs(scores[v] - old_score);
    }
    if (errors) errors[iter] = error;
    if (error < tol) {  // :43
      iter++;
      return iter;
    }
  }
  return iter;
}

// Personalised variant of the same loop: teleport mass (1-d) goes to a single
// seed vertex instead of (1-d)/m to every vertex.  Not in the reference (no
// PPR there); it is the FP32 scalar counterpart of the batched kernel.
int orc_ppr(int m, const int* row_offsets, const int* column_indices, const int* degree,
            int seed, float* scores, float damp, double tol, int max_iter, double* errors) {
  const float teleport = 1.0f - damp;
  std::vector<float> contrib(m);
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    double error = 0;
#pragma omp parallel for
    for (int n = 0; n < m; n++) contrib[n] = scores[n] / degree[n];
#pragma omp parallel for reduction(+ : error) schedule(dynamic, 64)
    for (int v = 0; v < m; v++) {
      float incoming_total = 0;
      for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++)
        incoming_total += contrib[column_indices[e]];
      float old_score = scores[v];
      float base = (v == seed) ? teleport : 0.0f;
      scores[v] = base + damp * incoming_total;
      error += std::fabs(scores[v] - old_score);
    }
    if (errors) errors[iter] = error;
    if (error < tol) return iter + 1;
  }
  return iter;
}

// src/pr/verifier.cc:50-68: one push iteration over the OUT-CSR from the
// tested scores; returns the L1 residual  sum |base + d*in_sum - test|.
double orc_pr_residual(int m, const int* out_row_offsets, const int* out_column_indices,
                       const int* degree, const float* scores_to_test, float damp) {
  const float base_score = (1.0f - damp) / m;  // verifier.cc:12
  std::vector<float> sums(m, 0.0f);
  for (int s = 0; s < m; s++) {
    float oc = scores_to_test[s] / degree[s];
    for (int e = out_row_offsets[s]; e < out_row_offsets[s + 1]; e++) sums[out_column_indices[e]] += oc;
  }
  double error = 0;
  for (int i = 0; i < m; i++) {
    float new_score = base_score + damp * sums[i];
    error += std::fabs(new_score - scores_to_test[i]);
  }
  return error;
}

// src/pr/scatter.cu:11-41 and its host loop (:57-78): push PageRank over the
// OUT-CSR.  next starts at base; every source adds kDamp*score/degree (its
// out-CSR row length; float, (kDamp*score) first) to each out-neighbour --
// here in out-CSR order, one of the orders the reference's atomicAdds can
// take; then the L1 change, the swap and the reset of next to base.  The loop
// continues while the change is > tol (:78; at least one iteration).  The L1
// change is summed in FP64 (the reference: FP32 block sums and an FP32
// atomicAdd, :36-40).  Returns the iterations.
int orc_pr_push(int m, const int* out_row_offsets, const int* out_column_indices, float* scores,
                float damp, double tol, int max_iter, double* errors) {
  const float base_score = (1.0f - damp) / m;  // :57
  std::vector<float> next(m, base_score);     // initialize, :11-14
  int iter = 0;
  do {
    ++iter;
    for (int src = 0; src < m; src++) {  // scatter, :16-27
      const int b = out_row_offsets[src], e = out_row_offsets[src + 1];
      const float value = damp * scores[src] / (float)(e - b);
      for (int k = b; k < e; k++) next[out_column_indices[k]] += value;
    }
    double diff = 0;
    for (int v = 0; v < m; v++) {  // reduce, :30-41
      diff += std::fabs(next[v] - scores[v]);
      scores[v] = next[v];
      next[v] = base_score;
    }
    if (errors) errors[iter - 1] = diff;
    if (!(diff > tol)) break;
  } while (iter < max_iter);
  return iter;
}

// FP64 power iteration of the same operator (uniform teleport, dropped
// dangling mass): x <- (1-d)/m + d * sum_{u in in(v)} x_u / outdeg_u.
// Runs until the L1 change is < tol or max_iter; returns iterations.
int orc_pr_fp64(int m, const int* row_offsets, const int* column_indices, const int* degree,
                double* x, double damp, double tol, int max_iter) {
  std::vector<double> contrib(m);
  const double base = (1.0 - damp) / m;
  for (int i = 0; i < m; i++) x[i] = 1.0 / m;
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    double error = 0;
#pragma omp parallel for
    for (int n = 0; n < m; n++) contrib[n] = degree[n] ? x[n] / degree[n] : 0.0;
#pragma omp parallel for reduction(+ : error) schedule(dynamic, 64)
    for (int v = 0; v < m; v++) {
      double s = 0;
      for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++) s += contrib[column_indices[e]];
      double nv = base + damp * s;
      error += std::fabs(nv - x[v]);
      x[v] = nv;
    }
    if (error < tol) return iter + 1;
  }
  return iter;
}

// FP64 batched personalised PR: X is [m][k] row-major (k seeds contiguous per
// vertex); column j teleports to seeds[j].  Init X = one-hot seeds.
int orc_ppr_fp64(int m, const int* row_offsets, const int* column_indices, const int* degree,
                 int k, const int* seeds, double* X, double damp, double tol, int max_iter) {
  std::vector<double> contrib((size_t)m * k);
  for (size_t i = 0; i < (size_t)m * k; i++) X[i] = 0.0;
  for (int j = 0; j < k; j++) X[(size_t)seeds[j] * k + j] = 1.0;
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    double error_max = 0;
#pragma omp parallel for
    for (int n = 0; n < m; n++)
      for (int j = 0; j < k; j++)
        contrib[(size_t)n * k + j] = degree[n] ? X[(size_t)n * k + j] / degree[n] : 0.0;
    std::vector<double> err(k, 0.0);
#pragma omp parallel
    {
      std::vector<double> acc(k), lerr(k, 0.0);
#pragma omp for schedule(dynamic, 64)
      for (int v = 0; v < m; v++) {
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++) {
          const double* c = &contrib[(size_t)column_indices[e] * k];
          for (int j = 0; j < k; j++) acc[j] += c[j];
        }
        for (int j = 0; j < k; j++) {
          double nv = damp * acc[j] + ((seeds[j] == v) ? (1.0 - damp) : 0.0);
          lerr[j] += std::fabs(nv - X[(size_t)v * k + j]);
          X[(size_t)v * k + j] = nv;
        }
      }
#pragma omp critical
      for (int j = 0; j < k; j++) err[j] += lerr[j];
    }
    for (int j = 0; j < k; j++) error_max = std::max(error_max, err[j]);
    if (error_max < tol) return iter + 1;
  }
  return iter;
}

// ---------------------------------------------------------------- SpMV

// src/spmv/omp_base.cc:17-28: y[i] = y[i] + sum_k x[Aj[k]] * Ax[k], in CSR
// order, product rounded before the add.
void orc_spmv(int num_rows, const int* Ap, const int* Aj, const float* Ax, const float* x,
              float* y) {
#pragma omp parallel for
  for (int i = 0; i < num_rows; i++) {
    float sum = y[i];
    for (int jj = Ap[i]; jj < Ap[i + 1]; jj++) sum += x[Aj[jj]] * Ax[jj];
    y[i] = sum;
  }
}

// src/spmv/verifier.cc:9-22 (maximum_relative_error<float>).
float orc_max_rel_err(const float* A, const float* B, long long N) {
  float max_error = 0;
  const float eps = std::sqrt(std::numeric_limits<float>::epsilon());
  for (long long i = 0; i < N; i++) {
    const float a = A[i], b = B[i];
    const float error = std::fabs(a - b);
    if (error != 0) max_error = std::max(max_error, error / (std::fabs(a) + std::fabs(b) + eps));
  }
  return max_error;
}

// ---------------------------------------------------------------- SymGS

// src/symgs/verifier.cc:50-65 (gs_serial): rows indices[row_start..row_stop)
// in order; each row's off-diagonal products summed in CSR order (product
// rounded before the add), the diagonal the row's LAST entry with column ==
// row; x[row] = (b[row] - rsum) / diag unless diag == 0.
static void gs_serial(const int* Ap, const int* Aj, const int* indices, const float* Ax, float* x,
                      const float* b, int row_start, int row_stop) {
  for (int i = row_start; i != row_stop; i++) {
    const int inew = indices[i];
    float rsum = 0, diag = 0;
    for (int jj = Ap[inew]; jj < Ap[inew + 1]; jj++) {
      const int j = Aj[jj];
      if (inew == j)
        diag = Ax[jj];
      else
        rsum += x[j] * Ax[jj];
    }
    if (diag != 0) x[inew] = (b[inew] - rsum) / diag;
  }
}

// src/symgs/verifier.cc:74-77 (SymGSVerifier's sweep) = src/symgs/omp_base.cc:
// 35-38: colors 0..C-1 forward, then C-1..0 backward; `sweeps` repetitions.
void orc_symgs(const int* Ap, const int* Aj, const int* indices, const float* Ax, float* x,
               const float* b, const int* color_offsets, int n_colors, int sweeps) {
  for (int s = 0; s < sweeps; s++) {
    for (int c = 0; c < n_colors; c++)
      gs_serial(Ap, Aj, indices, Ax, x, b, color_offsets[c], color_offsets[c + 1]);
    for (int c = n_colors; c > 0; c--)
      gs_serial(Ap, Aj, indices, Ax, x, b, color_offsets[c - 1], color_offsets[c]);
  }
}

// ---------------------------------------------------------------- adjoints (FP64)

// SpMV backward for y = A x (+ y0): dx = A^T g, dA[k] = g[row(k)] * x[col(k)].
void orc_spmv_bwd_fp64(int num_rows, int num_cols, const int* Ap, const int* Aj, const float* Ax,
                       const float* x, const float* g, double* dx, double* dA) {
  for (int j = 0; j < num_cols; j++) dx[j] = 0.0;
  for (int i = 0; i < num_rows; i++)
    for (int k = Ap[i]; k < Ap[i + 1]; k++) {
      dx[Aj[k]] += (double)Ax[k] * (double)g[i];
      dA[k] = (double)g[i] * (double)x[Aj[k]];
    }
}

// Weighted pull PageRank in FP64.  Input is the IN-CSR (row v lists sources u)
// with raw edge weights a_uv; s_u = sum_v a_uv is the source's out-strength.
//   x = (1-d)/m + d * P x,   P[v,u] = a_uv / s_u   (dangling mass dropped).
// warm != 0: iterate from the x passed in (e.g. the unperturbed solution when
// differencing), else from 1/m.  Jacobi sweeps (c is a snapshot of x).
int orc_wpr_fp64(int m, const int* row_offsets, const int* column_indices, const double* a,
                 double* x, double damp, double tol, int max_iter, int warm) {
  std::vector<double> s(m, 0.0), c(m);
  long long nnz = row_offsets[m];
  for (long long e = 0; e < nnz; e++) s[column_indices[e]] += a[e];
  if (!warm)
    for (int i = 0; i < m; i++) x[i] = 1.0 / m;
  const double base = (1.0 - damp) / m;
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    double err = 0;
#pragma omp parallel for
    for (int u = 0; u < m; u++) c[u] = s[u] > 0 ? x[u] / s[u] : 0.0;
#pragma omp parallel for reduction(+ : err) schedule(dynamic, 256)
    for (int v = 0; v < m; v++) {
      double acc = 0;
      for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++) acc += a[e] * c[column_indices[e]];
      double nv = base + damp * acc;
      err += std::fabs(nv - x[v]);
      x[v] = nv;
    }
    if (err < tol) return iter + 1;
  }
  return iter;
}

// Implicit adjoint of the weighted PR fixed point for L = w^T x.
//   lambda = w + d * P^T lambda       (solved by fixed-point iteration)
//   dL/da_uv = d * (x_u / s_u) * (lambda_v - sum_t P[t,u] lambda_t)
// where the second term comes from d s_u / d a_uv = 1 (row normalisation).
// grad has the in-CSR layout of `a`.
int orc_wpr_grad_fp64(int m, const int* row_offsets, const int* column_indices, const double* a,
                      const double* x, const double* w, double damp, double tol, int max_iter,
                      double* lambda, double* grad) {
  std::vector<double> s(m, 0.0), t(m);
  long long nnz = row_offsets[m];
  for (long long e = 0; e < nnz; e++) s[column_indices[e]] += a[e];
  // transposed index (by source u, edges in CSR order) so P^T lambda is a
  // parallel, order-fixed gather: t_u = sum over u's out-edges of a_uv lambda_v
  std::vector<long long> toff(m + 1, 0);
  for (long long e = 0; e < nnz; e++) toff[column_indices[e] + 1]++;
  for (int u = 0; u < m; u++) toff[u + 1] += toff[u];
  std::vector<long long> fill(toff.begin(), toff.end() - 1), tedge(nnz);
  std::vector<int> tdst(nnz);
  for (int v = 0; v < m; v++)
    for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++) {
      const long long k = fill[column_indices[e]]++;
      tedge[k] = e;
      tdst[k] = v;
    }
  auto transpose_apply = [&]() {
#pragma omp parallel for schedule(dynamic, 256)
    for (int u = 0; u < m; u++) {
      double acc = 0.0;
      for (long long k = toff[u]; k < toff[u + 1]; k++) acc += a[tedge[k]] * lambda[tdst[k]];
      t[u] = acc;
    }
  };
  for (int i = 0; i < m; i++) lambda[i] = w[i];
  int iter;
  for (iter = 0; iter < max_iter; iter++) {
    // (P^T lambda)_u = sum_v P[v,u] lambda_v = sum_{v: u in in(v)} a_uv lambda_v / s_u
    transpose_apply();
    double err = 0;
#pragma omp parallel for reduction(+ : err)
    for (int u = 0; u < m; u++) {
      double nl = w[u] + damp * (s[u] > 0 ? t[u] / s[u] : 0.0);
      err += std::fabs(nl - lambda[u]);
      lambda[u] = nl;
    }
    if (err < tol) { iter++; break; }
  }
  // q_u = sum_t P[t,u] lambda_t
  transpose_apply();
#pragma omp parallel for schedule(dynamic, 256)
  for (int v = 0; v < m; v++)
    for (int e = row_offsets[v]; e < row_offsets[v + 1]; e++) {
      int u = column_indices[e];
      double q = s[u] > 0 ? t[u] / s[u] : 0.0;
      grad[e] = s[u] > 0 ? damp * (x[u] / s[u]) * (lambda[v] - q) : 0.0;
    }
  return iter;
}

int orc_num_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}

}  // extern "C"
