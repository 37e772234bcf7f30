// ref_symgs_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// C-ABI harness linked against the reference's OpenMP SymGS sources
// (src/symgs/omp_base.cc, src/symgs/verifier.cc), compiled from where they lie
// (oracle/Makefile, target `ref`).  tests/golden/make_golden.py calls
// SymGSSolver() and SymGSVerifier() through it to record golden vectors.  A
// library of its own: the SymGS verifier's file-local templates share names
// with the SpMV verifier's.  Output in oracle/_ref/ only, never committed.

#include <fcntl.h>
#include <unistd.h>

#include <vector>

#include "symgs.h"  // src/symgs/symgs.h (SymGSSolver, SymGSVerifier)

namespace {
struct Redirect {  // fd 1 -> a caller-named log file for the duration of a call
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

// src/symgs/symgs.h:28 (SymGSSolver), x updated in place.
void ref_symgs_solve(int m, int nnz, int *Ap, int *Aj, int *indices, float *Ax, float *x, float *b,
                     const int *color_offsets, int n_colors, const char *log) {
  Redirect r(log);
  std::vector<int> co(color_offsets, color_offsets + n_colors + 1);
  SymGSSolver(m, nnz, Ap, Aj, indices, Ax, x, b, co);
}

// src/symgs/symgs.h:29 (SymGSVerifier): prints "Correct" or "POSSIBLE FAILURE".
void ref_symgs_verify(int m, int *Ap, int *Aj, int *indices, float *Ax, float *test_x,
                      float *x_host, float *b, const int *color_offsets, int n_colors,
                      const char *log) {
  Redirect r(log);
  std::vector<int> co(color_offsets, color_offsets + n_colors + 1);
  SymGSVerifier(m, Ap, Aj, indices, Ax, test_x, x_host, b, co);
}

}  // extern "C"
