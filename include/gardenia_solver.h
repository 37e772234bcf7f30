/* gardenia_solver.h -- the reference's two solver plug-in signatures, C++ linkage.
 *
 * libgardenia_hip.so defines exactly these functions, so a driver written
 * against the reference (src/pr/main.cc + src/pr/verifier.cc,
 * src/spmv/main.cc + src/spmv/verifier.cc) links against the library
 * unchanged, in place of the reference's omp_base.o / gather.o / ...:
 *
 *   void PRSolver(...)    replaces src/pr/pr.h:23      (e.g. src/pr/omp_base.cc:9)
 *   void SpmvSolver(...)  replaces src/spmv/spmv.h:29  (e.g. src/spmv/omp_base.cc:8)
 *   void SymGSSolver(...) replaces src/symgs/symgs.h:28 (e.g. src/symgs/omp_base.cc:27)
 *
 * Behaviour mirrors the reference plug-ins: host arrays in, scores / y
 * updated in place, the same stdout lines, EPSILON 1e-3 / MAX_ITER 30 /
 * kDamp 0.85f, and on failure a message on stderr followed by exit(1)
 * (include/cutil_subset.h:4-12).  The environment variables GD_PR_TOL and
 * GD_PR_MAX_ITER override the two PageRank constants, which the reference
 * fixes at compile time.
 */
#ifndef GARDENIA_SOLVER_H_
#define GARDENIA_SOLVER_H_

#include <vector>

void PRSolver(int m, int nnz, int *in_row_offsets, int *in_column_indices, int *out_row_offsets,
              int *out_column_indices, int *degree, float *scores);

void SpmvSolver(int m, int nnz, int *Ap, int *Aj, float *Ax, float *x, float *y);

void SymGSSolver(int num_rows, int nnz, int *Ap, int *Aj, int *indices, float *Ax, float *x,
                 float *b, std::vector<int> color_offsets);

#endif /* GARDENIA_SOLVER_H_ */
