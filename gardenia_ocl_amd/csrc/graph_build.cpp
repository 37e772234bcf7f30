// graph_build.cpp -- host-side graph layer of the library (C++17 + OpenMP).
//
// Replaces the reference's CSR builder (include/graph_io.h) for the PR/SpMV
// path:
//   * gd_csr_from_edges / gd_read_mtx / gd_read_gr / gd_read_dotgraph /
//     gd_read_graph: fill_data()/mtx2csr()/gr2csr()/graph2csr() semantics
//     (graph_io.h:21-130, :216-261, :280-295), including the unsorted-input
//     quirk where only ADJACENT duplicates are removed (graph_io.h:48-61).
//     Built with a stable counting sort by row instead of vector<vector<Edge>>.
//   * gd_rmat_graph / gd_er_matrix: seeded synthetic inputs (the reference has
//     no generator; its inputs are SNAP/UF downloads).  Counter-based hashing
//     makes every edge a pure function of (seed, index), so the edge list is
//     identical for any thread count and can be regenerated on demand.
//
// All functions are pure host code: they never touch the GPU.

#include <omp.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gardenia_graph.h"

namespace gd {
void set_error(const std::string& msg);  // gd_last_error() text (schedule.hip)
}

namespace {

int fail(int code, const std::string& msg) {
  gd::set_error(msg);
  return code;
}

inline uint64_t mix64(uint64_t x) {  // SplitMix64 finaliser (Steele et al. 2014)
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
T* xalloc(size_t n) {
  if (n == 0) n = 1;
  return static_cast<T*>(malloc(n * sizeof(T)));
}

// One RMAT edge: `scale` quadrant choices, MSB first, two 32-bit uniforms per
// 64-bit hash.  Quadrant a -> (0,0), b -> (0,1), c -> (1,0), d -> (1,1) for
// (src bit, dst bit).
inline void rmat_edge(uint64_t skey, int scale, double a, double ab, double abc, uint64_t idx,
                      uint32_t* s_out, uint32_t* d_out) {
  uint32_t s = 0, d = 0;
  uint64_t h = 0;
  for (int l = 0; l < scale; l++) {
    uint32_t r;
    if ((l & 1) == 0) {
      h = mix64(skey + idx * 16u + (uint64_t)(l >> 1));
      r = (uint32_t)(h >> 32);
    } else {
      r = (uint32_t)h;
    }
    const double u = (double)r * (1.0 / 4294967296.0);
    uint32_t sb, db;
    if (u < a) {
      sb = 0; db = 0;
    } else if (u < ab) {
      sb = 0; db = 1;
    } else if (u < abc) {
      sb = 1; db = 0;
    } else {
      sb = 1; db = 1;
    }
    s = (s << 1) | sb;
    d = (d << 1) | db;
  }
  *s_out = s;
  *d_out = d;
}

// Parallel stable LSD radix sort of 64-bit keys on the low `key_bits` bits.
// Returns the buffer holding the result (a or tmp).
uint64_t* radix_sort(uint64_t* a, uint64_t* tmp, size_t n, int key_bits) {
  const int kDigit = 8, kBuckets = 1 << kDigit;
  int nthreads = omp_get_max_threads();
  std::vector<size_t> hist((size_t)nthreads * kBuckets);
  uint64_t* src = a;
  uint64_t* dst = tmp;
  for (int shift = 0; shift < key_bits; shift += kDigit) {
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nthreads)
    {
      int t = omp_get_thread_num();
      size_t lo = n * t / nthreads, hi = n * (t + 1) / nthreads;
      size_t* h = &hist[(size_t)t * kBuckets];
      for (size_t i = lo; i < hi; i++) h[(src[i] >> shift) & (kBuckets - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        size_t sum = 0;
        for (int b = 0; b < kBuckets; b++)
          for (int tt = 0; tt < nthreads; tt++) {
            size_t c = hist[(size_t)tt * kBuckets + b];
            hist[(size_t)tt * kBuckets + b] = sum;
            sum += c;
          }
      }
      for (size_t i = lo; i < hi; i++) {
        uint64_t k = src[i];
        dst[h[(k >> shift) & (kBuckets - 1)]++] = k;
      }
    }
    std::swap(src, dst);
  }
  return src;
}

// Sorted, unique keys (row << bits | col) -> CSR (row_offsets, column_indices).
void csr_from_sorted_keys(const uint64_t* keys, size_t n, int m, int bits, int32_t* off,
                          int32_t* col) {
  const uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    int64_t r = (int64_t)(keys[i] >> bits);
    int64_t rp = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> bits);
    for (int64_t q = rp + 1; q <= r; q++) off[q] = (int32_t)i;
    col[i] = (int32_t)(keys[i] & mask);
  }
  int64_t last = n ? (int64_t)(keys[n - 1] >> bits) : -1;
  for (int64_t q = last + 1; q <= m; q++) off[q] = (int32_t)n;
}

void degrees_from_offsets(int m, const int32_t* off, int32_t* deg) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; i++) deg[i] = off[i + 1] - off[i];
}

struct EdgeRec {  // same layout and ordering predicate as the reference's Edge
  int dst;
  float wt;
};
inline bool edge_less(EdgeRec x, EdgeRec y) { return x.dst < y.dst; }

}  // namespace

extern "C" {

void gd_csr_free(gd_host_csr* c) {
  if (!c) return;
  free(c->row_offsets);
  free(c->column_indices);
  free(c->weight);
  free(c->degree);
  memset(c, 0, sizeof(*c));
}

int gd_rmat_edges(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                  int64_t first, int64_t count, int32_t* src, int32_t* dst) {
  if (scale < 1 || scale > 30 || edgefactor < 1 || first < 0 || count < 0) return GD_ERR_INVALID;
  const uint64_t skey = mix64(seed ^ 0x52A7C0DE5EEDull);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; i++) {
    uint32_t s, d;
    rmat_edge(skey, scale, a, ab, abc, (uint64_t)(first + i), &s, &d);
    src[i] = (int32_t)s;
    dst[i] = (int32_t)d;
  }
  return GD_OK;
}

int gd_rmat_graph(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                  gd_host_csr* out_csr, gd_host_csr* in_csr) {
  if (scale < 1 || scale > 30 || edgefactor < 1 || !(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return GD_ERR_INVALID;
  const int m = 1 << scale;
  const size_t n_raw = (size_t)edgefactor << scale;
  if (n_raw >= (1ull << 31)) return GD_ERR_INVALID;
  uint64_t* keys = xalloc<uint64_t>(n_raw);
  uint64_t* tmp = xalloc<uint64_t>(n_raw);
  if (!keys || !tmp) {
    free(keys);
    free(tmp);
    return GD_ERR_ALLOC;
  }
  const uint64_t skey = mix64(seed ^ 0x52A7C0DE5EEDull);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n_raw; i++) {
    uint32_t s, d;
    rmat_edge(skey, scale, a, ab, abc, (uint64_t)i, &s, &d);
    keys[i] = ((uint64_t)s << scale) | d;
  }
  uint64_t* sorted = radix_sort(keys, tmp, n_raw, 2 * scale);
  uint64_t* other = (sorted == keys) ? tmp : keys;
  // unique + drop self-loops (sequential compaction; bandwidth-bound)
  const uint64_t lo_mask = (1ull << scale) - 1;
  size_t n = 0;
  for (size_t i = 0; i < n_raw; i++) {
    uint64_t k = sorted[i];
    if ((k >> scale) == (k & lo_mask)) continue;
    if (n && sorted[n - 1] == k) continue;
    sorted[n++] = k;
  }
  int rc = GD_OK;
  if (out_csr) {
    out_csr->m = m;
    out_csr->nnz = (int64_t)n;
    out_csr->row_offsets = xalloc<int32_t>(m + 1);
    out_csr->column_indices = xalloc<int32_t>(n);
    out_csr->weight = nullptr;
    out_csr->degree = xalloc<int32_t>(m);
    if (!out_csr->row_offsets || !out_csr->column_indices || !out_csr->degree) rc = GD_ERR_ALLOC;
    else {
      csr_from_sorted_keys(sorted, n, m, scale, out_csr->row_offsets, out_csr->column_indices);
      degrees_from_offsets(m, out_csr->row_offsets, out_csr->degree);
    }
  }
  if (rc == GD_OK && in_csr) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
      uint64_t k = sorted[i];
      other[i] = ((k & lo_mask) << scale) | (k >> scale);
    }
    uint64_t* t2 = radix_sort(other, sorted, n, 2 * scale);
    in_csr->m = m;
    in_csr->nnz = (int64_t)n;
    in_csr->row_offsets = xalloc<int32_t>(m + 1);
    in_csr->column_indices = xalloc<int32_t>(n);
    in_csr->weight = nullptr;
    in_csr->degree = xalloc<int32_t>(m);
    if (!in_csr->row_offsets || !in_csr->column_indices || !in_csr->degree) rc = GD_ERR_ALLOC;
    else {
      csr_from_sorted_keys(t2, n, m, scale, in_csr->row_offsets, in_csr->column_indices);
      degrees_from_offsets(m, in_csr->row_offsets, in_csr->degree);
    }
  }
  free(keys);
  free(tmp);
  return rc;
}

int gd_er_matrix(int log2_rows, int nnz_per_row, uint64_t seed, gd_host_csr* out) {
  if (log2_rows < 1 || log2_rows > 30 || nnz_per_row < 0 || out == nullptr) return GD_ERR_INVALID;
  const int64_t m = 1ll << log2_rows;
  if (nnz_per_row > m) return GD_ERR_INVALID;
  const int64_t nnz = m * nnz_per_row;
  if (nnz >= (1ll << 31)) return GD_ERR_INVALID;
  out->m = (int32_t)m;
  out->nnz = nnz;
  out->row_offsets = xalloc<int32_t>(m + 1);
  out->column_indices = xalloc<int32_t>(nnz);
  out->weight = xalloc<float>(nnz);
  out->degree = xalloc<int32_t>(m);
  if (!out->row_offsets || !out->column_indices || !out->weight || !out->degree) {
    gd_csr_free(out);
    return GD_ERR_ALLOC;
  }
  const uint64_t ckey = mix64(seed ^ 0xE5C01u);
  const uint64_t vkey = mix64(seed ^ 0xA1C0FFEEu);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; i++) {
    int32_t* row = out->column_indices + i * nnz_per_row;
    uint64_t attempt = 0;
    int k = 0;
    while (k < nnz_per_row) {
      uint64_t h = mix64(ckey + (uint64_t)i * 0x100000000ull + attempt++);
      int32_t cidx = (int32_t)(((h >> 32) * (uint64_t)m) >> 32);
      bool dup = false;
      for (int t = 0; t < k; t++) dup |= (row[t] == cidx);
      if (!dup) row[k++] = cidx;
    }
    std::sort(row, row + nnz_per_row);
    out->row_offsets[i] = (int32_t)(i * nnz_per_row);
    out->degree[i] = nnz_per_row;
  }
  out->row_offsets[m] = (int32_t)nnz;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nnz; k++) out->weight[k] = gd_uniform_pm1(vkey, (uint64_t)k);
  return GD_OK;
}

float gd_uniform_pm1(uint64_t key, uint64_t idx) {
  // 24-bit uniform in [0,1) scaled to [-1,1): exact in FP32.
  uint64_t h = mix64(key + idx);
  float u = (float)(h >> 40) * (1.0f / 16777216.0f);
  return 2.0f * u - 1.0f;
}

int gd_fill_uniform_pm1(uint64_t seed, int64_t n, float* out) {
  if (n < 0 || (n > 0 && !out)) return GD_ERR_INVALID;
  const uint64_t key = mix64(seed ^ 0x5EC7012u);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) out[i] = gd_uniform_pm1(key, (uint64_t)i);
  return GD_OK;
}

// fill_data() semantics on an explicit edge list (0-based, file order).
int gd_csr_from_edges(int m, int64_t n, const int32_t* src, const int32_t* dst, const float* wt,
                      int symmetrize, int transpose, int sorted, int remove_selfloops,
                      int remove_redundents, int mirror_selfloops, gd_host_csr* out) {
  if (m < 0 || n < 0 || !out) return GD_ERR_INVALID;
  for (int64_t i = 0; i < n; i++)
    if (src[i] < 0 || src[i] >= m || dst[i] < 0 || dst[i] >= m)
      return fail(GD_ERR_INVALID, "edge " + std::to_string(i) + " has a vertex id outside [0, m)");
  // Placement (graph_io.h:245-257): the row that receives each pushed edge, in
  // push order.  A mirrored edge is pushed before the forward one.
  int64_t cap = symmetrize ? 2 * n : n;
  std::vector<int32_t> prow;
  std::vector<EdgeRec> pe;
  prow.reserve(cap);
  pe.reserve(cap);
  bool tr = transpose != 0;
  for (int64_t i = 0; i < n; i++) {
    int32_t s = src[i], d = dst[i];
    float w = wt ? wt[i] : 1.0f;
    if (symmetrize && (mirror_selfloops || s != d)) {
      prow.push_back(d);
      pe.push_back(EdgeRec{s, w});
      tr = false;
    }
    if (!tr) {
      prow.push_back(s);
      pe.push_back(EdgeRec{d, w});
    } else {
      prow.push_back(d);
      pe.push_back(EdgeRec{s, w});
    }
  }
  // Stable counting sort by row keeps each row's push order.
  std::vector<int64_t> start(m + 1, 0);
  for (int32_t r : prow) start[r + 1]++;
  for (int i = 0; i < m; i++) start[i + 1] += start[i];
  std::vector<EdgeRec> rows(pe.size());
  {
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (size_t i = 0; i < pe.size(); i++) rows[pos[prow[i]]++] = pe[i];
  }
  std::vector<int32_t> len(m);
  for (int i = 0; i < m; i++) {
    EdgeRec* b = rows.data() + start[i];
    EdgeRec* e = rows.data() + start[i + 1];
    if (sorted) std::sort(b, e, edge_less);
    if (remove_selfloops) e = std::remove_if(b, e, [i](const EdgeRec& x) { return x.dst == i; });
    if (remove_redundents)
      e = std::unique(b, e, [](const EdgeRec& x, const EdgeRec& y) { return x.dst == y.dst; });
    len[i] = (int32_t)(e - b);
  }
  int64_t total = 0;
  for (int i = 0; i < m; i++) total += len[i];
  if (total >= (1ll << 31)) return GD_ERR_INVALID;
  out->m = m;
  out->nnz = total;
  out->row_offsets = xalloc<int32_t>(m + 1);
  out->column_indices = xalloc<int32_t>(total);
  out->weight = xalloc<float>(total);
  out->degree = xalloc<int32_t>(m);
  if (!out->row_offsets || !out->column_indices || !out->weight || !out->degree) {
    gd_csr_free(out);
    return GD_ERR_ALLOC;
  }
  int64_t k = 0;
  for (int i = 0; i < m; i++) {
    out->row_offsets[i] = (int32_t)k;
    out->degree[i] = len[i];
    const EdgeRec* b = rows.data() + start[i];
    for (int j = 0; j < len[i]; j++, k++) {
      out->column_indices[k] = b[j].dst;
      out->weight[k] = b[j].wt;
    }
  }
  out->row_offsets[m] = (int32_t)k;
  return GD_OK;
}

// mtx2csr() (graph_io.h:216-261) + read_graph() degree (:289-293).
int gd_read_mtx(const char* path, int symmetrize, int transpose, int sorted, int remove_selfloops,
                int remove_redundents, gd_host_csr* out) {
  if (!path || !out) return GD_ERR_INVALID;
  FILE* f = fopen(path, "r");
  if (!f) return fail(GD_ERR_IO, std::string("cannot open ") + path);
  std::string line;
  char* buf = nullptr;
  size_t cap = 0;
  ssize_t got;
  int m = 0, ncols = 0;
  long long nnz = 0;
  bool have_header = false;
  while ((got = getline(&buf, &cap, f)) >= 0) {
    if (got > 0 && buf[0] == '%') continue;
    if (sscanf(buf, "%d %d %lld", &m, &ncols, &nnz) == 3) have_header = true;
    break;
  }
  if (!have_header || m < 0 || nnz < 0) {
    free(buf);
    fclose(f);
    return fail(GD_ERR_IO, "malformed .mtx header");
  }
  std::vector<int32_t> s(nnz), d(nnz);
  std::vector<float> w(nnz);
  float wt = 1.0f;
  for (long long i = 0; i < nnz; i++) {
    if (getline(&buf, &cap, f) < 0) {
      free(buf);
      fclose(f);
      return fail(GD_ERR_IO, "truncated .mtx edge list");
    }
    int a = 0, b = 0;
    int num = sscanf(buf, "%d %d %f", &a, &b, &wt);
    if (num < 2) {
      free(buf);
      fclose(f);
      return fail(GD_ERR_IO, "malformed .mtx edge line");
    }
    if (num == 2) wt = 1;
    if (wt < 0) wt = -wt;
    s[i] = a - 1;
    d[i] = b - 1;
    w[i] = wt;
  }
  free(buf);
  fclose(f);
  return gd_csr_from_edges(m, nnz, s.data(), d.data(), w.data(), symmetrize, transpose, sorted,
                           remove_selfloops, remove_redundents, 0, out);
}

// gr2csr() (graph_io.h:133-174): leading lines starting with 'c' skipped, a
// "p sp <m> <nnz>" header, then nnz arc lines "a <src> <dst> [w]" with 1-based
// ids; the weight column is ignored (every weight 1), symmetrize mirrors
// self-loops too (:160-164).  A malformed arc line or an id outside [1, m] is
// an error here (the reference reuses the previous line's ids or indexes out of
// bounds).
int gd_read_gr(const char* path, int symmetrize, int transpose, int sorted, int remove_selfloops,
               int remove_redundents, gd_host_csr* out) {
  if (!path || !out) return GD_ERR_INVALID;
  FILE* f = fopen(path, "r");
  if (!f) return fail(GD_ERR_IO, std::string("cannot open ") + path);
  char* buf = nullptr;
  size_t cap = 0;
  int m = -1;
  long long nnz = -1;
  bool header = false;
  while (getline(&buf, &cap, f) >= 0) {
    if (buf[0] == 'c' || buf[0] == '\n' || buf[0] == '\0') continue;
    char c = 0, sp[8] = {0};
    header = sscanf(buf, "%c %7s %d %lld", &c, sp, &m, &nnz) == 4;
    break;
  }
  if (!header || m < 0 || nnz < 0) {
    free(buf);
    fclose(f);
    return fail(GD_ERR_IO, "malformed .gr header (expected p sp <m> <nnz>)");
  }
  std::vector<int32_t> s(nnz), d(nnz);
  for (long long i = 0; i < nnz; i++) {
    char c = 0;
    int a = 0, b = 0;
    if (getline(&buf, &cap, f) < 0 || sscanf(buf, "%c %d %d", &c, &a, &b) != 3 || c != 'a') {
      free(buf);
      fclose(f);
      return fail(GD_ERR_IO, "truncated or malformed .gr arc line (expected a <src> <dst>)");
    }
    s[i] = a - 1;
    d[i] = b - 1;
  }
  free(buf);
  fclose(f);
  return gd_csr_from_edges(m, nnz, s.data(), d.data(), nullptr, symmetrize, transpose, sorted,
                           remove_selfloops, remove_redundents, 1, out);
}

// graph2csr() (graph_io.h:177-213), the METIS-style adjacency format: a
// "<m> <nnz> ..." header, then line v (v = 1..m) lists v's neighbours, 1-based.
// A line is read as integers up to its first non-integer token (istream >>
// int); lines missing at the end of the file are empty rows.  Every weight is
// 1 and symmetrize mirrors self-loops too.
int gd_read_dotgraph(const char* path, int symmetrize, int transpose, int sorted,
                     int remove_selfloops, int remove_redundents, gd_host_csr* out) {
  if (!path || !out) return GD_ERR_INVALID;
  FILE* f = fopen(path, "r");
  if (!f) return fail(GD_ERR_IO, std::string("cannot open ") + path);
  char* buf = nullptr;
  size_t cap = 0;
  int m = -1;
  long long nnz = 0;
  if (getline(&buf, &cap, f) < 0 || sscanf(buf, "%d %lld", &m, &nnz) != 2 || m < 0) {
    free(buf);
    fclose(f);
    return fail(GD_ERR_IO, "malformed .graph header (expected <m> <nnz>)");
  }
  std::vector<int32_t> s, d;
  s.reserve(nnz > 0 ? nnz : 0);
  d.reserve(nnz > 0 ? nnz : 0);
  for (int v = 0; v < m; v++) {
    if (getline(&buf, &cap, f) < 0) break;
    const char* p = buf;
    for (;;) {
      char* end = nullptr;
      errno = 0;
      const long x = strtol(p, &end, 10);
      if (end == p) break;
      if (errno == ERANGE || x < INT32_MIN || x > INT32_MAX) {
        free(buf);
        fclose(f);
        return fail(GD_ERR_IO, "integer out of range in .graph adjacency line");
      }
      s.push_back(v);
      d.push_back((int32_t)(x - 1));
      p = end;
    }
  }
  free(buf);
  fclose(f);
  return gd_csr_from_edges(m, (int64_t)s.size(), s.data(), d.data(), nullptr, symmetrize,
                           transpose, sorted, remove_selfloops, remove_redundents, 1, out);
}

// Sequential first-fit coloring in vertex order over the union of the CSR and
// its transpose (an off-diagonal entry in either direction separates two
// rows), self-loops ignored.  Deterministic; a proper coloring for the SymGS
// sweep whatever the symmetry of the input.  The reference colors with its VC
// benchmark (src/vc/omp_base.cc:9-47), whose parallel first-fit is racy.
int gd_color_first_fit(int m, const int32_t* row_offsets, const int32_t* column_indices,
                       int32_t* colors, int32_t* n_colors) {
  if (m < 0 || (m > 0 && (!row_offsets || !colors)) || !n_colors) return GD_ERR_INVALID;
  *n_colors = 0;
  if (m == 0) return GD_OK;
  const int64_t nnz = row_offsets[m];
  if (row_offsets[0] != 0 || nnz < 0 || (nnz > 0 && !column_indices))
    return fail(GD_ERR_INVALID, "bad row offsets");
  for (int i = 0; i < m; i++)
    if (row_offsets[i + 1] < row_offsets[i]) return fail(GD_ERR_INVALID, "bad row offsets");
  for (int64_t k = 0; k < nnz; k++)
    if (column_indices[k] < 0 || column_indices[k] >= m)
      return fail(GD_ERR_INVALID, "column index out of range [0, m)");
  // transpose: tstart/tcol list the rows that reference each column
  std::vector<int64_t> tstart(m + 1, 0);
  for (int64_t k = 0; k < nnz; k++) tstart[column_indices[k] + 1]++;
  for (int i = 0; i < m; i++) tstart[i + 1] += tstart[i];
  std::vector<int32_t> tcol(nnz);
  {
    std::vector<int64_t> at(tstart.begin(), tstart.end() - 1);
    for (int r = 0; r < m; r++)
      for (int64_t k = row_offsets[r]; k < row_offsets[r + 1]; k++)
        tcol[at[column_indices[k]]++] = r;
  }
  std::vector<int32_t> mark;  // mark[c] == v: color c is taken by a neighbour of v
  int32_t nc = 0;
  for (int v = 0; v < m; v++) colors[v] = -1;
  for (int v = 0; v < m; v++) {
    auto forbid = [&](int u) {
      if (u != v && colors[u] >= 0) mark[colors[u]] = v;
    };
    for (int64_t k = row_offsets[v]; k < row_offsets[v + 1]; k++) forbid(column_indices[k]);
    for (int64_t k = tstart[v]; k < tstart[v + 1]; k++) forbid(tcol[k]);
    int32_t c = 0;
    while (c < nc && mark[c] == v) c++;
    if (c == nc) {
      nc++;
      mark.push_back(-1);
    }
    colors[v] = c;
  }
  *n_colors = nc;
  return GD_OK;
}

// The driver's ordering (src/symgs/main.cc:38-49): vertices sorted by color
// (stable: increasing vertex id inside a color) and the color offsets.
int gd_color_order(int m, const int32_t* colors, int32_t n_colors, int32_t* indices,
                   int32_t* color_offsets) {
  if (m < 0 || n_colors < 0 || (m > 0 && (!colors || !indices)) || !color_offsets)
    return GD_ERR_INVALID;
  std::vector<int32_t> cnt(n_colors + 1, 0);
  for (int v = 0; v < m; v++) {
    if (colors[v] < 0 || colors[v] >= n_colors)
      return fail(GD_ERR_INVALID, "color outside [0, n_colors)");
    cnt[colors[v] + 1]++;
  }
  for (int c = 0; c < n_colors; c++) cnt[c + 1] += cnt[c];
  for (int c = 0; c <= n_colors; c++) color_offsets[c] = cnt[c];
  for (int v = 0; v < m; v++) indices[cnt[colors[v]]++] = v;
  return GD_OK;
}

// read_graph() (graph_io.h:280-295): the format chosen by the first of
// ".mtx", ".graph", ".gr" found in the path (strstr, in that order).
int gd_read_graph(const char* path, int symmetrize, int transpose, int sorted, int remove_selfloops,
                  int remove_redundents, gd_host_csr* out) {
  if (!path || !out) return GD_ERR_INVALID;
  if (strstr(path, ".mtx"))
    return gd_read_mtx(path, symmetrize, transpose, sorted, remove_selfloops, remove_redundents,
                       out);
  if (strstr(path, ".graph"))
    return gd_read_dotgraph(path, symmetrize, transpose, sorted, remove_selfloops,
                            remove_redundents, out);
  if (strstr(path, ".gr"))
    return gd_read_gr(path, symmetrize, transpose, sorted, remove_selfloops, remove_redundents,
                      out);
  return fail(GD_ERR_IO, "Unrecognizable input file format (not .mtx, .graph or .gr)");
}

}  // extern "C"
