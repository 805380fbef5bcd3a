/*
 * eg_oracle.h -- CPU oracle for the EverythingGraph data-parallel path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libeg, the drop-in
 * drivers) links, loads or calls this code.  It is used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only as the
 * checker or as the timed CPU baseline.
 *
 * Each function restates the reference algorithm it names (file:line in the
 * EverythingGraph tree).  Where the reference has a documented defect the
 * oracle follows the mathematical contract instead; the header comment of each
 * function says which.  How each function is pinned (closed forms, library
 * routines, brute force) is listed in tests/test_oracle.py.
 */
#ifndef EG_ORACLE_H
#define EG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based generator (no reference counterpart) ---- */
uint64_t ego_mix64(uint64_t z);
void ego_gen_rmat(int scale, uint64_t m_in, uint64_t seed, uint32_t ta,
                  uint32_t tab, uint32_t tabc, uint32_t *src, uint32_t *dst);
void ego_gen_er(uint32_t n, uint64_t m_in, uint64_t seed, uint32_t *src,
                uint32_t *dst);
void ego_gen_weights(uint64_t m_in, uint64_t seed, float *w);

/* ---- ingest: preload_graph, init_all.c:81-110 ---- */
void ego_symmetrize(uint64_t m_in, const uint32_t *src, const uint32_t *dst,
                    const float *w, uint32_t *osrc, uint32_t *odst, float *ow);

/* ---- partition geometry: get_partition_id / get_partition_range,
 *      init_all.c:29-41,54-60 ---- */
uint32_t ego_partition_id(uint32_t n, uint32_t P, uint32_t v);
uint32_t ego_default_P(uint32_t n); /* random.c:69-71, widened (R13) */

/* ---- layout builders ---- */
void ego_build_csr(uint32_t n, uint64_t m, const uint32_t *src,
                   const uint32_t *dst, const float *w, uint64_t *row_ptr,
                   uint32_t *col, float *wout);
void ego_build_csc(uint32_t n, uint64_t m, const uint64_t *row_ptr,
                   const uint32_t *col, const float *w, uint64_t *col_ptr,
                   uint32_t *row_idx, float *win);
void ego_build_grid(uint32_t n, uint32_t P, uint64_t m, const uint32_t *src,
                    const uint32_t *dst, const float *w, int presort_by_src,
                    uint64_t *cell_ptr, uint32_t *gsrc, uint32_t *gdst,
                    float *gw);
void ego_build_edgelist(uint32_t n, uint64_t m, const uint32_t *src,
                        const uint32_t *dst, const float *w, int mode,
                        uint32_t *osrc, uint32_t *odst, float *ow);

/* ---- algorithms ---- */
void ego_pagerank_pull(uint32_t n, const uint64_t *in_ptr,
                       const uint32_t *in_idx, const uint32_t *out_deg,
                       double damping, int iterations, double init,
                       int dangling_uniform, double *rank,
                       double *dangling_hist, double *l1_hist);
void ego_pagerank_grid(uint32_t n, uint32_t P, const uint64_t *cell_ptr,
                       const uint32_t *gsrc, const uint32_t *gdst,
                       const uint32_t *out_deg, double damping, int iterations,
                       double init, double *rank);
void ego_spmv_pull(uint32_t n, const uint64_t *in_ptr, const uint32_t *in_idx,
                   const float *in_w, const double *x, double *y);
void ego_spmv_grid(uint32_t n, uint32_t P, const uint64_t *cell_ptr,
                   const uint32_t *gsrc, const uint32_t *gdst, const float *gw,
                   const double *x, double *y);
int ego_bfs(uint32_t n, uint64_t m, const uint64_t *out_ptr,
            const uint32_t *out_col, const uint64_t *in_ptr,
            const uint32_t *in_idx, uint32_t root, int start_pull,
            int switch_mode, uint32_t *dist, uint64_t *level_sizes,
            int max_levels);
int ego_bfs_grid(uint32_t n, uint32_t P, const uint64_t *cell_ptr,
                 const uint32_t *gsrc, const uint32_t *gdst, uint32_t root,
                 uint32_t *dist);
int ego_bfs_edgelist(uint32_t n, uint64_t m, const uint32_t *src,
                     const uint32_t *dst, uint32_t root, uint32_t *dist);
int ego_wcc_push(uint32_t n, const uint64_t *out_ptr, const uint32_t *out_col,
                 int64_t *labels);
int ego_wcc_edges(uint32_t n, uint64_t m, const uint32_t *src,
                  const uint32_t *dst, int64_t *labels);
void ego_ppr(uint32_t n, const uint64_t *in_ptr, const uint32_t *in_idx,
             const uint32_t *out_deg, uint32_t nseeds, const uint32_t *seeds,
             double alpha, int iterations, double *out);
int ego_check_bfs_parents(uint32_t n, const uint64_t *out_ptr,
                          const uint32_t *out_col, uint32_t root,
                          const uint32_t *dist, const uint32_t *parent);

/* Single-source shortest paths (sssp_pushpull.c:39-190): dist[root] = 0,
 * unreached = UINT32_MAX - 1 (sssp_pushpull.c:8).  Edge weights by kind:
 * 0 = (u % 10) + (v % 10) (CREATE_WEIGHT, init_all.c:661-686), 1 = unit,
 * 2 = kind 0 plus one (strictly positive; used only to pin the oracle
 * against library routines that read zero as "no edge").  Returns the number
 * of relaxation rounds. */
int ego_sssp(uint32_t n, const uint64_t *out_ptr, const uint32_t *out_col, uint32_t root,
             int kind, uint32_t *dist);
uint32_t ego_sssp_weight(uint32_t u, uint32_t v, int kind);
/* Legality of an SSSP tree: parent[root] = root, unreached = UINT32_MAX - 1,
 * else the edge parent[v] -> v exists and dist[parent] + w = dist[v].
 * Returns the number of violations. */
int ego_check_sssp_parents(uint32_t n, const uint64_t *out_ptr, const uint32_t *out_col,
                           uint32_t root, int kind, const uint32_t *dist,
                           const uint32_t *parent);

int ego_num_threads(void);

/* Full-size checks from the input records (both = 1: each record also
 * stored reversed, -u).  Row content (count, column sum, column xor) of every
 * vertex from the records and from a CSR; the rows of selected vertices
 * (ascending, distinct) in stored order; single-source BFS level invariants
 * (returns the number of violations). */
void ego_record_row_sums(uint32_t n, uint64_t m, const uint32_t *src,
                         const uint32_t *dst, int both, uint64_t *cnt,
                         uint64_t *sum, uint32_t *xr);
void ego_csr_row_sums(uint32_t n, const uint64_t *ptr, const uint32_t *col,
                      uint64_t *sum, uint32_t *xr);
void ego_rows_of(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                 int both, uint32_t nsel, const uint32_t *sel, uint64_t *out_ptr,
                 uint32_t *out_col);
int64_t ego_check_bfs_levels(uint32_t n, uint64_t m, const uint32_t *src,
                             const uint32_t *dst, int both, uint32_t root,
                             const uint32_t *dist);

#ifdef __cplusplus
}
#endif
#endif
