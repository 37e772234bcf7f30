// solver_shim.cpp -- PRSolver / SpmvSolver / SymGSSolver with the reference's
// signatures.
//
// The link-time plug-in boundary of the reference (one solver definition per
// binary, src/pr/Makefile:17-47, src/spmv/Makefile:16-50): these definitions
// forward to the C ABI and print the reference's stdout lines
// (src/pr/omp_base.cc:15,42,46-47; src/spmv/omp_base.cc:14,30;
// src/symgs/omp_base.cc:33,40).

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gardenia_hip.h"
#include "gardenia_solver.h"

namespace {

[[noreturn]] void fail(const char* what, int st) {
  fprintf(stderr, "%s failed (status %d): %s\n", what, st, gd_last_error());
  exit(1);
}

}  // namespace

void PRSolver(int m, int nnz, int* in_row_offsets, int* in_column_indices, int* out_row_offsets,
              int* out_column_indices, int* degree, float* scores) {
  gd_pr_options o;
  gd_pr_default_options(&o);
  if (const char* t = getenv("GD_PR_TOL")) o.tol = atof(t);
  if (const char* t = getenv("GD_PR_MAX_ITER")) o.max_iter = atoi(t);
  printf("Launching HIP PR solver (gfx950 fused pull) ...\n");
  std::vector<double> errors(o.max_iter > 0 ? o.max_iter : 1);
  gd_pr_stats st;
  int rc = gd_pr_solve(m, nnz, in_row_offsets, in_column_indices, out_row_offsets,
                       out_column_indices, degree, scores, &o, &st, errors.data());
  if (rc != GD_OK) fail("PRSolver", rc);
  for (int i = 0; i < st.iterations; i++) printf(" %2d    %lf\n", i + 1, errors[i]);
  // The reference prints iter+1 after its for-loop (src/pr/omp_base.cc:23,46):
  // the iterations run when one converged, MAX_ITER+1 otherwise (1 for MAX_ITER 0).
  printf("\titerations = %d.\n", st.iterations + (st.converged ? 0 : 1));
  printf("\truntime [%s] = %f ms.\n", "hip_pull", st.solve_ms);
}

void SpmvSolver(int m, int nnz, int* Ap, int* Aj, float* Ax, float* x, float* y) {
  printf("Launching HIP SpMV solver (gfx950 LDS tiles) ...\n");
  float ms = 0.0f;
  int rc = gd_spmv_solve(m, nnz, Ap, Aj, Ax, x, y, &ms);
  if (rc != GD_OK) fail("SpmvSolver", rc);
  printf("\truntime [%s] = %f ms.\n", "hip_tile", ms);
}

void SymGSSolver(int num_rows, int nnz, int* Ap, int* Aj, int* indices, float* Ax, float* x,
                 float* b, std::vector<int> color_offsets) {
  printf("Launching HIP SymGS solver (gfx950 LDS tiles, exact row order) ...\n");
  if (color_offsets.empty()) fail("SymGSSolver", GD_ERR_INVALID);
  float ms = 0.0f;
  int rc = gd_symgs_solve(num_rows, nnz, Ap, Aj, indices, Ax, x, b, color_offsets.data(),
                          (int)color_offsets.size() - 1, &ms);
  if (rc != GD_OK) fail("SymGSSolver", rc);
  printf("\truntime [%s] = %f ms.\n", "hip_tile", ms);
}
