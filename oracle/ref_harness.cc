// ref_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// A small C-ABI harness linked against the reference's own OpenMP sources,
// compiled from where they lie (see oracle/Makefile, target `ref`).  It lets
// tests/golden/make_golden.py call the reference's read_graph(), PRSolver(),
// PRVerifier(), SpmvSolver() and SpmvVerifier() on chosen inputs and record
// their outputs as golden vectors.  The output goes to oracle/_ref/ only and is
// never committed or shipped with the product.
//
// The reference prints to stdout; every entry point here redirects fd 1 to a
// caller-named log file for the duration of the call so the per-iteration
// " %2d    %lf" lines (src/pr/omp_base.cc:42) can be parsed.

#include <fcntl.h>
#include <unistd.h>

#include "pr.h"      // src/pr/pr.h  (PRSolver, PRVerifier, EPSILON, MAX_ITER)
#include "graph_io.h"  // include/graph_io.h (read_graph, header-only)

// src/spmv/spmv.h:29-30 (declared here because spmv.h re-#defines ValueType)
void SpmvSolver(int m, int nnz, int *Ap, int *Aj, float *Ax, float *x, float *y);
void SpmvVerifier(int m, int nnz, int *Ap, int *Aj, float *Ax, float *x, float *y, float *test_y);

namespace {
struct Redirect {
  int saved = -1, fd = -1;
  explicit Redirect(const char *path) {
    fflush(stdout);
    if (!path) return;
    fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return;
    saved = dup(1);
    dup2(fd, 1);
  }
  ~Redirect() {
    fflush(stdout);
    if (saved >= 0) {
      dup2(saved, 1);
      close(saved);
    }
    if (fd >= 0) close(fd);
  }
};
}  // namespace

extern "C" {

int ref_max_iter(void) { return MAX_ITER; }
double ref_epsilon(void) { return EPSILON; }

// include/graph_io.h:280-295.  Arrays are malloc'd by the reference; free with ref_free.
int ref_read_graph(const char *path, int symmetrize, int transpose, int sorted, int remove_selfloops,
                   int remove_redundents, int *m, int *nnz, int **row_offsets, int **column_indices,
                   float **weight, int **degree, const char *log) {
  Redirect r(log);
  char prog[] = "ref";
  char *argv[2] = {prog, (char *)path};
  int n = 0;
  read_graph(2, argv, *m, n, *nnz, *row_offsets, *column_indices, *degree, *weight, symmetrize != 0,
             transpose != 0, sorted != 0, remove_selfloops != 0, remove_redundents != 0);
  return 0;
}

void ref_free(void *p) { free(p); }

void ref_pr_solve(int m, int nnz, int *in_off, int *in_col, int *out_off, int *out_col, int *degree,
                  float *scores, const char *log) {
  Redirect r(log);
  PRSolver(m, nnz, in_off, in_col, out_off, out_col, degree, scores);
}

void ref_pr_verify(int m, int *in_off, int *in_col, int *out_off, int *out_col, int *degree,
                   float *scores, double target, const char *log) {
  Redirect r(log);
  PRVerifier(m, in_off, in_col, out_off, out_col, degree, scores, target);
}

void ref_spmv_solve(int m, int nnz, int *Ap, int *Aj, float *Ax, float *x, float *y, const char *log) {
  Redirect r(log);
  SpmvSolver(m, nnz, Ap, Aj, Ax, x, y);
}

void ref_spmv_verify(int m, int nnz, int *Ap, int *Aj, float *Ax, float *x, float *y, float *test_y,
                     const char *log) {
  Redirect r(log);
  SpmvVerifier(m, nnz, Ap, Aj, Ax, x, y, test_y);
}

}  // extern "C"
