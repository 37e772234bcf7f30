/* gardenia_graph.h -- host-side graph layer of libgardenia_hip.so (C ABI).
 *
 * Replaces the reference's CSR builder include/graph_io.h for the PageRank /
 * SpMV path.  Host memory only; nothing here touches the GPU.
 */
#ifndef GARDENIA_GRAPH_H_
#define GARDENIA_GRAPH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared by every gd_* entry point (the reference has none: it
 * prints and exit()s, include/cutil_subset.h:4-12). */
enum {
  GD_OK = 0,
  GD_ERR_INVALID = 1,  /* bad argument or malformed CSR */
  GD_ERR_ALLOC = 2,    /* host or device allocation failed */
  GD_ERR_DEVICE = 3,   /* HIP runtime error */
  GD_ERR_IO = 4,       /* file missing or malformed */
  GD_ERR_NO_DEVICE = 5 /* no HIP device visible */
};

/* A host CSR as produced by read_graph() (include/graph_io.h:280-295):
 * row_offsets[m+1], column_indices[nnz], weight[nnz] (may be NULL),
 * degree[m] = row length.  Arrays are malloc'd by the library; release with
 * gd_csr_free(). */
typedef struct gd_host_csr {
  int32_t m;
  int64_t nnz;
  int32_t *row_offsets;
  int32_t *column_indices;
  float *weight;
  int32_t *degree;
} gd_host_csr;

void gd_csr_free(gd_host_csr *csr);

/* fill_data() semantics (include/graph_io.h:21-130) on an explicit 0-based
 * edge list in file order: optional symmetrize (mirror_selfloops=0 is the .mtx
 * rule of graph_io.h:246, =1 the .gr/.graph rule of :160/:196), transpose,
 * per-row sort (std::sort by dst), self-loop removal, ADJACENT duplicate
 * removal. */
int gd_csr_from_edges(int m, int64_t n, const int32_t *src, const int32_t *dst, const float *wt,
                      int symmetrize, int transpose, int sorted, int remove_selfloops,
                      int remove_redundents, int mirror_selfloops, gd_host_csr *out);

/* mtx2csr() + degree (include/graph_io.h:216-261, :289-293).  The PR driver
 * uses (symmetrize=!directed, transpose=0/1, sorted=0, 1, 1)
 * (src/pr/main.cc:23-24); the SpMV driver (sym, 0, 1, 0, 1)
 * (src/spmv/main.cc:24). */
int gd_read_mtx(const char *path, int symmetrize, int transpose, int sorted, int remove_selfloops,
                int remove_redundents, gd_host_csr *out);

/* gr2csr() + degree (include/graph_io.h:133-174): DIMACS arcs "a src dst [w]"
 * (1-based, weight ignored), symmetrize mirrors self-loops as well. */
int gd_read_gr(const char *path, int symmetrize, int transpose, int sorted, int remove_selfloops,
               int remove_redundents, gd_host_csr *out);

/* graph2csr() + degree (include/graph_io.h:177-213): "m nnz" header, then one
 * line of 1-based neighbours per vertex. */
int gd_read_dotgraph(const char *path, int symmetrize, int transpose, int sorted,
                     int remove_selfloops, int remove_redundents, gd_host_csr *out);

/* read_graph() dispatch (include/graph_io.h:280-295): ".mtx", ".graph", ".gr"
 * tested in that order with strstr; anything else is GD_ERR_IO. */
int gd_read_graph(const char *path, int symmetrize, int transpose, int sorted, int remove_selfloops,
                  int remove_redundents, gd_host_csr *out);

/* Deterministic sequential first-fit coloring (vertex order) over the CSR and
 * its transpose, self-loops ignored: colors[v] in [0, *n_colors), no two rows
 * joined by an off-diagonal entry in either direction share a color.  Stands
 * in for the SymGS driver's VCSolver call (src/symgs/main.cc:41). */
int gd_color_first_fit(int m, const int32_t *row_offsets, const int32_t *column_indices,
                       int32_t *colors, int32_t *n_colors);

/* The SymGS driver's ordering (src/symgs/main.cc:38-49): indices[m] = vertices
 * sorted by color (increasing id inside a color), color_offsets[n_colors+1]. */
int gd_color_order(int m, const int32_t *colors, int32_t n_colors, int32_t *indices,
                   int32_t *color_offsets);

/* Seeded, unpermuted RMAT (a,b,c, d=1-a-b-c), 2^scale vertices,
 * edgefactor*2^scale raw edges; self-loops and duplicates removed.  Fills the
 * out-CSR (rows = sources) and/or the in-CSR (rows = destinations), both with
 * sorted columns -- identical to what read_graph() builds from the canonical
 * edge list written as .mtx in (src,dst) order. */
int gd_rmat_graph(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                  gd_host_csr *out_csr, gd_host_csr *in_csr);

/* The raw (pre-canonicalisation) RMAT edges [first, first+count). */
int gd_rmat_edges(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                  int64_t first, int64_t count, int32_t *src, int32_t *dst);

/* Erdos-Renyi matrix: 2^log2_rows square, exactly nnz_per_row distinct uniform
 * columns per row (sorted), weight[k] ~ U(-1,1) (24-bit, exact in FP32). */
int gd_er_matrix(int log2_rows, int nnz_per_row, uint64_t seed, gd_host_csr *out);

/* out[i] ~ U(-1,1), 24-bit, a pure function of (seed, i). */
int gd_fill_uniform_pm1(uint64_t seed, int64_t n, float *out);
float gd_uniform_pm1(uint64_t key, uint64_t idx);

#ifdef __cplusplus
}
#endif
#endif /* GARDENIA_GRAPH_H_ */
