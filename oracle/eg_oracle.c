/*
 * eg_oracle.c -- CPU oracle (plain C + OpenMP) for the EverythingGraph
 * data-parallel path.
 *
 * TEST INFRASTRUCTURE ONLY: called by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product.  Every routine restates
 * the reference loop it cites; reference defects (R1..R14 in the survey of the
 * reference) are not reproduced, the mathematical contract is.
 *
 * Parity of this oracle is pinned in tests/test_oracle.py against independent
 * routines (numpy stable argsort, scipy.sparse products, scipy.sparse.csgraph
 * shortest paths and connected components, textbook queue BFS) and against
 * closed forms (cycles, stars, complete graphs, paths).
 */
#include "eg_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ull

int ego_num_threads(void) { return omp_get_max_threads(); }

/* SplitMix64 finaliser (Steele, Lea, Flood 2014).  The generator is ours: the
 * reference reads edge files and has no generator. */
uint64_t ego_mix64(uint64_t z) {
  z += GOLDEN;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* R-MAT (Chakrabarti, Zhan, Faloutsos 2004), no noise, no permutation.
 * Edge e draws `scale` quadrants from the counter-based stream
 * h_j = mix64(mix64(mix64(seed) ^ e) + (j+1)*GOLDEN), two 32-bit draws per
 * hash (high half first).  Quadrant thresholds are integer: r < ta -> (0,0),
 * r < tab -> (0,1), r < tabc -> (1,0), else (1,1); bits fill from the MSB. */
void ego_gen_rmat(int scale, uint64_t m_in, uint64_t seed, uint32_t ta,
                  uint32_t tab, uint32_t tabc, uint32_t *src, uint32_t *dst) {
  const uint64_t skey = ego_mix64(seed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m_in; e++) {
    const uint64_t ekey = ego_mix64(skey ^ (uint64_t)e);
    uint32_t s = 0, d = 0;
    uint64_t h = 0;
    for (int l = 0; l < scale; l++) {
      uint32_t r;
      if ((l & 1) == 0) {
        h = ego_mix64(ekey + (uint64_t)(l / 2 + 1) * GOLDEN);
        r = (uint32_t)(h >> 32);
      } else {
        r = (uint32_t)h;
      }
      s <<= 1;
      d <<= 1;
      if (r < ta) {
      } else if (r < tab) {
        d |= 1u;
      } else if (r < tabc) {
        s |= 1u;
      } else {
        s |= 1u;
        d |= 1u;
      }
    }
    src[e] = s;
    dst[e] = d;
  }
}

/* Erdos-Renyi G(n, m): independent uniform (src, dst) pairs, mapped to [0, n)
 * by Lemire's multiply-shift on the two 32-bit halves of one hash. */
void ego_gen_er(uint32_t n, uint64_t m_in, uint64_t seed, uint32_t *src,
                uint32_t *dst) {
  const uint64_t skey = ego_mix64(seed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m_in; e++) {
    const uint64_t h = ego_mix64(ego_mix64(skey ^ (uint64_t)e) + GOLDEN);
    src[e] = (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
    dst[e] = (uint32_t)(((h & 0xFFFFFFFFull) * (uint64_t)n) >> 32);
  }
}

/* Weights w = ((h >> 40) + 1) * 2^-24, exactly representable, in (0, 1]. */
void ego_gen_weights(uint64_t m_in, uint64_t seed, float *w) {
  const uint64_t wkey = ego_mix64(seed ^ 0x5745494748545321ull);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m_in; e++) {
    const uint64_t h = ego_mix64(wkey ^ (uint64_t)e);
    w[e] = (float)((h >> 40) + 1) * 0x1p-24f;
  }
}

/* preload_graph, init_all.c:84-107: record i is copied to slot i and, with
 * -u, its reversal to slot i + m_in.  The reference leaves the weight of the
 * reversed copy unset; here it carries the record's weight. */
void ego_symmetrize(uint64_t m_in, const uint32_t *src, const uint32_t *dst,
                    const float *w, uint32_t *osrc, uint32_t *odst, float *ow) {
  for (uint64_t i = 0; i < m_in; i++) {
    osrc[i] = src[i];
    odst[i] = dst[i];
    osrc[i + m_in] = dst[i];
    odst[i + m_in] = src[i];
    if (w && ow) {
      ow[i] = w[i];
      ow[i + m_in] = w[i];
    }
  }
}

/* get_partition_id, init_all.c:54-60 with the geometry of init_all.c:838-842:
 * the first n % P partitions hold ceil(n/P) vertices, the rest one fewer. */
uint32_t ego_partition_id(uint32_t n, uint32_t P, uint32_t v) {
  const uint64_t mod_val = n % P;
  const uint64_t psize = mod_val ? (uint64_t)n / P + 1 : (uint64_t)n / P;
  const uint64_t split = mod_val * psize;
  if (mod_val == 0) return (uint32_t)(v / psize);
  return (uint32_t)(v < split ? v / psize : (v - split) / (psize - 1) + mod_val);
}

/* random.c:69-71, in 64-bit arithmetic (the reference overflows, R13). */
uint32_t ego_default_P(uint32_t n) {
  uint64_t P = ((uint64_t)n * 8 / 1024) / 20;
  if (P > 2000) P = 256;
  if (P == 0) P = 4;
  return (uint32_t)P;
}

/* Stable counting sort of records by key[i] < nkeys.  ptr gets nkeys+1 offsets;
 * perm[k] = index of the record placed at k.  This is the contract of the
 * reference's stable radix sort intSort::iSort (radixSort_ligra.h:355, stable
 * because radixBlock places back to front, :158-161).  Small inputs take one
 * serial counting pass; large ones (the full-size configurations) an
 * OpenMP LSD radix sort over 11-bit digits with per-thread digit counts over
 * contiguous blocks, which is stable digit by digit and so yields the same
 * permutation. */
static void serial_count_sort(uint64_t nkeys, uint64_t m, const uint32_t *key,
                              uint64_t *ptr, uint64_t *perm) {
  uint64_t *cur = (uint64_t *)calloc(nkeys + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i < m; i++) cur[key[i] + 1]++;
  for (uint64_t k = 0; k < nkeys; k++) cur[k + 1] += cur[k];
  if (ptr) memcpy(ptr, cur, (nkeys + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; i++) perm[cur[key[i]]++] = i;
  free(cur);
}

#define EGO_DIGIT 11
#define EGO_BINS (1u << EGO_DIGIT)

static void parallel_radix_sort(uint64_t nkeys, uint64_t m, const uint32_t *key,
                                uint64_t *ptr, uint64_t *perm) {
  int bits = 1;
  while (bits < 32 && (1ull << bits) < nkeys) bits++;
  const int T = omp_get_max_threads();
  uint32_t *kbuf[2];
  uint64_t *pbuf[2];
  kbuf[0] = (uint32_t *)malloc(m * sizeof(uint32_t));
  kbuf[1] = (uint32_t *)malloc(m * sizeof(uint32_t));
  pbuf[0] = perm;
  pbuf[1] = (uint64_t *)malloc(m * sizeof(uint64_t));
  uint64_t *cnt = (uint64_t *)malloc((size_t)T * EGO_BINS * sizeof(uint64_t));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)m; i++) {
    kbuf[0][i] = key[i];
    pbuf[0][i] = (uint64_t)i;
  }
  int cur = 0;
  for (int shift = 0; shift < bits; shift += EGO_DIGIT) {
    const uint32_t *kin = kbuf[cur];
    const uint64_t *pin = pbuf[cur];
    uint32_t *kout = kbuf[cur ^ 1];
    uint64_t *pout = pbuf[cur ^ 1];
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      const uint64_t lo = m * (uint64_t)t / (uint64_t)T, hi = m * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t *c = cnt + (size_t)t * EGO_BINS;
      memset(c, 0, EGO_BINS * sizeof(uint64_t));
      for (uint64_t i = lo; i < hi; i++) c[(kin[i] >> shift) & (EGO_BINS - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        /* digit-major, then block order: the offsets of a stable pass */
        uint64_t run = 0;
        for (uint32_t b = 0; b < EGO_BINS; b++)
          for (int u = 0; u < T; u++) {
            const uint64_t v = cnt[(size_t)u * EGO_BINS + b];
            cnt[(size_t)u * EGO_BINS + b] = run;
            run += v;
          }
      }
      for (uint64_t i = lo; i < hi; i++) {
        const uint64_t at = c[(kin[i] >> shift) & (EGO_BINS - 1)]++;
        kout[at] = kin[i];
        pout[at] = pin[i];
      }
    }
    cur ^= 1;
  }
  if (cur == 1) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)m; i++) perm[i] = pbuf[1][i];
  }
  if (ptr) {
    /* ptr[k] = first position whose key is >= k (key sorted ascending) */
    const uint32_t *ks = kbuf[cur];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k <= (int64_t)nkeys; k++) {
      uint64_t a = 0, b = m;
      while (a < b) {
        const uint64_t mid = a + (b - a) / 2;
        if (ks[mid] < (uint64_t)k) a = mid + 1; else b = mid;
      }
      ptr[k] = a;
    }
  }
  free(cnt);
  free(pbuf[1]);
  free(kbuf[1]);
  free(kbuf[0]);
}

static void stable_count_sort(uint64_t nkeys, uint64_t m, const uint32_t *key,
                              uint64_t *ptr, uint64_t *perm) {
  if (m < ((uint64_t)1 << 22) || omp_get_max_threads() == 1)
    serial_count_sort(nkeys, m, key, ptr, perm);
  else
    parallel_radix_sort(nkeys, m, key, ptr, perm);
}

/* init_adj_sort, init_all.c:629-697 (modes 2/3; mode 4's count sort,
 * init_all.c:444-496, has the same content): stable sort of the records by
 * src, row_ptr from the boundary detection + scanIBack(min), which equals the
 * exclusive scan of the out-degree histogram. */
void ego_build_csr(uint32_t n, uint64_t m, const uint32_t *src,
                   const uint32_t *dst, const float *w, uint64_t *row_ptr,
                   uint32_t *col, float *wout) {
  uint64_t *perm = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  stable_count_sort(n, m, src, row_ptr, perm);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)m; k++) {
    col[k] = dst[perm[k]];
    if (w && wout) wout[k] = w[perm[k]];
  }
  free(perm);
}

/* init_adj_sort, init_all.c:714-767: the src-sorted records are stably sorted
 * by dst; the in-list entry is the record's src.  Input is the CSR. */
void ego_build_csc(uint32_t n, uint64_t m, const uint64_t *row_ptr,
                   const uint32_t *col, const float *w, uint64_t *col_ptr,
                   uint32_t *row_idx, float *win) {
  uint64_t *perm = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  uint32_t *rowof = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; v++)
    for (uint64_t k = row_ptr[v]; k < row_ptr[v + 1]; k++) rowof[k] = (uint32_t)v;
  stable_count_sort(n, m, col, col_ptr, perm);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)m; k++) {
    row_idx[k] = rowof[perm[k]];
    if (w && win) win[k] = w[perm[k]];
  }
  free(rowof);
  free(perm);
}

/* init_grid_sort, init_all.c:252-338 (modes 0/6): sort by src, split rows by
 * src partition, then a stable sort of every row by dst partition.  The result
 * is the records ordered by cell (part(src), part(dst)) with (src, input order)
 * inside a cell; cell offsets are an exact exclusive scan (R10 not reproduced).
 * Mode 6 re-sorts cells by src (init_all.c:344-358), which changes nothing.
 * presort_by_src = 0 gives mode 1 (init_grid_nosort, init_all.c:363-424), whose
 * in-cell order the reference leaves to lock timing; here it is input order. */
void ego_build_grid(uint32_t n, uint32_t P, uint64_t m, const uint32_t *src,
                    const uint32_t *dst, const float *w, int presort_by_src,
                    uint64_t *cell_ptr, uint32_t *gsrc, uint32_t *gdst,
                    float *gw) {
  uint64_t *perm0 = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  if (presort_by_src) {
    stable_count_sort(n, m, src, NULL, perm0);
  } else {
    for (uint64_t i = 0; i < m; i++) perm0[i] = i;
  }
  uint32_t *cell = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)m; k++) {
    const uint64_t i = perm0[k];
    cell[k] = ego_partition_id(n, P, src[i]) * P + ego_partition_id(n, P, dst[i]);
  }
  uint64_t *perm1 = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  stable_count_sort((uint64_t)P * P, m, cell, cell_ptr, perm1);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)m; k++) {
    const uint64_t i = perm0[perm1[k]];
    gsrc[k] = src[i];
    gdst[k] = dst[i];
    if (w && gw) gw[k] = w[i];
  }
  free(perm1);
  free(cell);
  free(perm0);
}

/* init_edgelist, init_all.c:573-628: mode 7 sorts by src, mode 8 by src then by
 * dst inside each src (both stable), mode 9 keeps input order. */
void ego_build_edgelist(uint32_t n, uint64_t m, const uint32_t *src,
                        const uint32_t *dst, const float *w, int mode,
                        uint32_t *osrc, uint32_t *odst, float *ow) {
  uint64_t *perm = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
  if (mode == 9) {
    for (uint64_t i = 0; i < m; i++) perm[i] = i;
  } else if (mode == 7) {
    stable_count_sort(n, m, src, NULL, perm);
  } else {
    /* LSD: stable by dst, then stable by src. */
    uint64_t *p1 = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
    stable_count_sort(n, m, dst, NULL, p1);
    uint32_t *k2 = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
    for (uint64_t k = 0; k < m; k++) k2[k] = src[p1[k]];
    uint64_t *p2 = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
    stable_count_sort(n, m, k2, NULL, p2);
    for (uint64_t k = 0; k < m; k++) perm[k] = p1[p2[k]];
    free(p2);
    free(k2);
    free(p1);
  }
  for (uint64_t k = 0; k < m; k++) {
    osrc[k] = src[perm[k]];
    odst[k] = dst[perm[k]];
    if (w && ow) ow[k] = w[perm[k]];
  }
  free(perm);
}

/* pr_algo_pull, pagerank_simple.c:72-84, iterated as the recurrence it states:
 * r'[v] = (1-d)/n + d * sum_{u in in(v)} r[u]/outdeg(u), every iteration from
 * the previous one (the reference never swaps in pull mode, R2).  FP64.
 * dangling_uniform adds d * (sum of ranks of out-degree-0 vertices) / n; the
 * reference drops that mass (R3).  dangling_hist / l1_hist receive, per
 * iteration, the dangling mass of the input iterate and sum |r' - r|. */
void ego_pagerank_pull(uint32_t n, const uint64_t *in_ptr,
                       const uint32_t *in_idx, const uint32_t *out_deg,
                       double damping, int iterations, double init,
                       int dangling_uniform, double *rank,
                       double *dangling_hist, double *l1_hist) {
  double *contrib = (double *)malloc((size_t)n * sizeof(double));
  double *next = (double *)malloc((size_t)n * sizeof(double));
  const double base = (1.0 - damping) / (double)n;
  for (uint32_t v = 0; v < n; v++) rank[v] = init;
  for (int it = 0; it < iterations; it++) {
    double dangling = 0.0;
#pragma omp parallel for reduction(+ : dangling) schedule(static)
    for (int64_t u = 0; u < (int64_t)n; u++) {
      if (out_deg[u]) {
        contrib[u] = rank[u] / (double)out_deg[u];
      } else {
        contrib[u] = 0.0;
        dangling += rank[u];
      }
    }
    const double extra = dangling_uniform ? damping * dangling / (double)n : 0.0;
    double l1 = 0.0;
#pragma omp parallel for reduction(+ : l1) schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)n; v++) {
      double s = 0.0;
      for (uint64_t k = in_ptr[v]; k < in_ptr[v + 1]; k++) s += contrib[in_idx[k]];
      next[v] = base + damping * s + extra;
      l1 += fabs(next[v] - rank[v]);
    }
    memcpy(rank, next, (size_t)n * sizeof(double));
    if (dangling_hist) dangling_hist[it] = dangling;
    if (l1_hist) l1_hist[it] = l1;
  }
  free(next);
  free(contrib);
}

/* prgrid_algo + apply, prgrid_cilk.c:62-76,131-135: rows of cells in sequence,
 * cells of a row in parallel (each owns a disjoint dst range), then
 * r = (1-d)/n + d*acc.  The divisor is the true out-degree (R4 not
 * reproduced).  FP64. */
void ego_pagerank_grid(uint32_t n, uint32_t P, const uint64_t *cell_ptr,
                       const uint32_t *gsrc, const uint32_t *gdst,
                       const uint32_t *out_deg, double damping, int iterations,
                       double init, double *rank) {
  double *acc = (double *)calloc((size_t)n, sizeof(double));
  const double base = (1.0 - damping) / (double)n;
  for (uint32_t v = 0; v < n; v++) rank[v] = init;
  for (int it = 0; it < iterations; it++) {
    for (uint32_t i = 0; i < P; i++) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t j = 0; j < (int64_t)P; j++) {
        const uint64_t c = (uint64_t)i * P + (uint64_t)j;
        for (uint64_t k = cell_ptr[c]; k < cell_ptr[c + 1]; k++)
          acc[gdst[k]] += rank[gsrc[k]] / (double)out_deg[gsrc[k]];
      }
    }
    for (uint32_t v = 0; v < n; v++) {
      rank[v] = base + damping * acc[v];
      acc[v] = 0.0;
    }
  }
  free(acc);
}

/* spmv_algo (pull), spmv.c:98-106, with the arithmetic redefined (R7): FP64
 * y[v] = sum_{u in in(v)} w(u->v) * x[u], i.e. y = A^T x.  in_w NULL means
 * unit weights. */
void ego_spmv_pull(uint32_t n, const uint64_t *in_ptr, const uint32_t *in_idx,
                   const float *in_w, const double *x, double *y) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; v++) {
    double s = 0.0;
    for (uint64_t k = in_ptr[v]; k < in_ptr[v + 1]; k++)
      s += (in_w ? (double)in_w[k] : 1.0) * x[in_idx[k]];
    y[v] = s;
  }
}

/* The same product streamed over the grid cells (prgrid_cilk.c:62-76 traversal
 * applied to SpMV). */
void ego_spmv_grid(uint32_t n, uint32_t P, const uint64_t *cell_ptr,
                   const uint32_t *gsrc, const uint32_t *gdst, const float *gw,
                   const double *x, double *y) {
  for (uint32_t v = 0; v < n; v++) y[v] = 0.0;
  for (uint32_t i = 0; i < P; i++) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j = 0; j < (int64_t)P; j++) {
      const uint64_t c = (uint64_t)i * P + (uint64_t)j;
      for (uint64_t k = cell_ptr[c]; k < cell_ptr[c + 1]; k++)
        y[gdst[k]] += (gw ? (double)gw[k] : 1.0) * x[gsrc[k]];
    }
  }
}

/* bfs iterator, bfs_simple.c:136-229: level-synchronous BFS that runs a push
 * step (bfs_push, :121-134) over the queue or a pull step (bfs_pull, :104-120,
 * in-edges with early exit) over unvisited vertices, switching with the rule of
 * :191-197 (frontier size + frontier out-degree sum > m/20).  dist uses the
 * reference's encoding of correct levels (bfsgrid_cilk.c:58-60): root = 1,
 * level L -> L+1, unreached = 0 (R9 not reproduced).  Returns the number of
 * levels; level_sizes[l] = vertices at level l. */
int ego_bfs(uint32_t n, uint64_t m, const uint64_t *out_ptr,
            const uint32_t *out_col, const uint64_t *in_ptr,
            const uint32_t *in_idx, uint32_t root, int start_pull,
            int switch_mode, uint32_t *dist, uint64_t *level_sizes,
            int max_levels) {
  uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  uint32_t *nq = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  char *front = (char *)calloc(n, 1);
  char *nfront = (char *)calloc(n, 1);
  memset(dist, 0, (size_t)n * sizeof(uint32_t));
  dist[root] = 1;
  front[root] = 1;
  q[0] = root;
  uint64_t qn = 1;
  int pull = start_pull;
  int levels = 0;
  while (qn) {
    if (levels < max_levels && level_sizes) level_sizes[levels] = qn;
    levels++;
    const uint32_t lvl = (uint32_t)levels; /* dist of the current frontier */
    uint64_t nqn = 0, degsum = 0;
    if (!pull) {
      for (uint64_t i = 0; i < qn; i++) {
        const uint32_t u = q[i];
        for (uint64_t k = out_ptr[u]; k < out_ptr[u + 1]; k++) {
          const uint32_t v = out_col[k];
          if (dist[v] == 0) {
            dist[v] = lvl + 1;
            nfront[v] = 1;
            nq[nqn++] = v;
            degsum += out_ptr[v + 1] - out_ptr[v];
          }
        }
      }
    } else {
      for (uint32_t v = 0; v < n; v++) {
        if (dist[v] != 0) continue;
        for (uint64_t k = in_ptr[v]; k < in_ptr[v + 1]; k++) {
          if (front[in_idx[k]]) {
            dist[v] = lvl + 1;
            nfront[v] = 1;
            nq[nqn++] = v;
            degsum += out_ptr[v + 1] - out_ptr[v];
            break;
          }
        }
      }
    }
    const double thr = (double)m / 20.0;
    if ((switch_mode && !pull && (double)(nqn + degsum) > thr) ||
        (pull && (double)(nqn + degsum) > thr))
      pull = 1;
    else
      pull = 0;
    for (uint64_t i = 0; i < qn; i++) front[q[i]] = 0;
    char *tf = front;
    front = nfront;
    nfront = tf;
    uint32_t *tq = q;
    q = nq;
    nq = tq;
    qn = nqn;
  }
  free(q);
  free(nq);
  free(front);
  free(nfront);
  return levels;
}

/* bfsgrid iterator, bfsgrid_cilk.c:48-91,121-168: per level, rows of cells
 * whose src partition holds a frontier vertex are streamed in order;
 * dist[dst] = dist[src] + 1 for unvisited dst.  Root = 1. */
int ego_bfs_grid(uint32_t n, uint32_t P, const uint64_t *cell_ptr,
                 const uint32_t *gsrc, const uint32_t *gdst, uint32_t root,
                 uint32_t *dist) {
  char *front = (char *)calloc(n, 1);
  char *nfront = (char *)calloc(n, 1);
  char *active = (char *)calloc(P, 1);
  memset(dist, 0, (size_t)n * sizeof(uint32_t));
  dist[root] = 1;
  front[root] = 1;
  active[ego_partition_id(n, P, root)] = 1;
  uint64_t items = 1;
  int levels = 0;
  while (items) {
    levels++;
    for (uint32_t i = 0; i < P; i++) {
      if (!active[i]) continue;
      for (uint64_t k = cell_ptr[(uint64_t)i * P]; k < cell_ptr[(uint64_t)(i + 1) * P]; k++) {
        const uint32_t s = gsrc[k], d = gdst[k];
        if (front[s] && dist[d] == 0) {
          dist[d] = dist[s] + 1;
          nfront[d] = 1;
        }
      }
    }
    char *t = front;
    front = nfront;
    nfront = t;
    memset(active, 0, P);
    items = 0;
    for (uint32_t v = 0; v < n; v++) {
      if (front[v]) {
        items++;
        active[ego_partition_id(n, P, v)] = 1;
      }
      nfront[v] = 0;
    }
  }
  free(front);
  free(nfront);
  free(active);
  return levels;
}

/* bfs_edgecentric, bfs_simple.c:58-103: full sweeps over the edge array,
 * dist[dst] = dist[src] + 1 for frontier src and unvisited dst. */
int ego_bfs_edgelist(uint32_t n, uint64_t m, const uint32_t *src,
                     const uint32_t *dst, uint32_t root, uint32_t *dist) {
  char *front = (char *)calloc(n, 1);
  char *nfront = (char *)calloc(n, 1);
  memset(dist, 0, (size_t)n * sizeof(uint32_t));
  dist[root] = 1;
  front[root] = 1;
  uint64_t items = 1;
  int levels = 0;
  while (items) {
    levels++;
    for (uint64_t k = 0; k < m; k++) {
      if (front[src[k]] && dist[dst[k]] == 0) {
        dist[dst[k]] = dist[src[k]] + 1;
        nfront[dst[k]] = 1;
      }
    }
    items = 0;
    for (uint32_t v = 0; v < n; v++) {
      items += nfront[v];
      front[v] = nfront[v];
      nfront[v] = 0;
    }
  }
  free(front);
  free(nfront);
  return levels;
}

/* wcc_algo worklist, wcc.c:30-42 with the driver of :262-277 and reset of
 * :128-135: labels start at the vertex id and the minimum is pushed along
 * out-edges to the fixpoint.  label[v] = min id of the vertices that reach v
 * (with -u: the min id of v's component).  Returns the number of rounds. */
int ego_wcc_push(uint32_t n, const uint64_t *out_ptr, const uint32_t *out_col,
                 int64_t *labels) {
  uint32_t *q = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  uint32_t *nq = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  char *inq = (char *)calloc(n, 1);
  uint64_t qn = n;
  for (uint32_t v = 0; v < n; v++) {
    labels[v] = v;
    q[v] = v;
  }
  int rounds = 0;
  while (qn) {
    rounds++;
    uint64_t nqn = 0;
    for (uint64_t i = 0; i < qn; i++) {
      const uint32_t u = q[i];
      for (uint64_t k = out_ptr[u]; k < out_ptr[u + 1]; k++) {
        const uint32_t v = out_col[k];
        if (labels[u] < labels[v]) {
          labels[v] = labels[u];
          if (!inq[v]) {
            inq[v] = 1;
            nq[nqn++] = v;
          }
        }
      }
    }
    for (uint64_t i = 0; i < nqn; i++) inq[nq[i]] = 0;
    uint32_t *t = q;
    q = nq;
    nq = t;
    qn = nqn;
  }
  free(q);
  free(nq);
  free(inq);
  return rounds;
}

/* wcc grid / edge-array sweeps, wcc.c:193-261: every edge relaxes both
 * endpoints with writeMin until a sweep changes nothing.  The fixpoint is the
 * min vertex id of each weakly connected component. */
int ego_wcc_edges(uint32_t n, uint64_t m, const uint32_t *src,
                  const uint32_t *dst, int64_t *labels) {
  for (uint32_t v = 0; v < n; v++) labels[v] = v;
  int rounds = 0;
  int changed = 1;
  while (changed) {
    changed = 0;
    rounds++;
    for (uint64_t k = 0; k < m; k++) {
      const uint32_t s = src[k], d = dst[k];
      if (labels[s] < labels[d]) {
        labels[d] = labels[s];
        changed = 1;
      } else if (labels[d] < labels[s]) {
        labels[s] = labels[d];
        changed = 1;
      }
    }
  }
  return rounds;
}

/* Personalized PageRank by power iteration (no reference counterpart; the
 * batched form of the pull recurrence above): p = uniform over the distinct
 * seeds, x0 = p, x' = alpha*p + (1-alpha) * sum_{u in in(v)} x[u]/outdeg(u).
 * Dangling mass is dropped, as in the reference PageRank.  FP64. */
void ego_ppr(uint32_t n, const uint64_t *in_ptr, const uint32_t *in_idx,
             const uint32_t *out_deg, uint32_t nseeds, const uint32_t *seeds,
             double alpha, int iterations, double *out) {
  double *p = (double *)calloc(n, sizeof(double));
  double *contrib = (double *)malloc((size_t)n * sizeof(double));
  uint32_t distinct = 0;
  for (uint32_t i = 0; i < nseeds; i++)
    if (p[seeds[i]] == 0.0) {
      p[seeds[i]] = 1.0;
      distinct++;
    }
  for (uint32_t v = 0; v < n; v++) {
    p[v] = p[v] / (double)distinct;
    out[v] = p[v];
  }
  for (int it = 0; it < iterations; it++) {
#pragma omp parallel for schedule(static)
    for (int64_t u = 0; u < (int64_t)n; u++)
      contrib[u] = out_deg[u] ? out[u] / (double)out_deg[u] : 0.0;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)n; v++) {
      double s = 0.0;
      for (uint64_t k = in_ptr[v]; k < in_ptr[v + 1]; k++) s += contrib[in_idx[k]];
      out[v] = alpha * p[v] + (1.0 - alpha) * s;
    }
  }
  free(contrib);
  free(p);
}

/* sssp_pushpull.c relaxes frontier vertices until no distance changes
 * (push: :45-62, pull: :75-104, loop: :139-188).  The fixpoint is the
 * shortest-path distance, whatever order the relaxations take, so the oracle
 * runs a FIFO worklist Bellman-Ford (each vertex queued at most once at a
 * time).  The reference's push adds 1 per edge while its pull adds the edge
 * weight (:54 vs :88); both directions here use the edge weight. */
uint32_t ego_sssp_weight(uint32_t u, uint32_t v, int kind) {
  if (kind == 1) return 1u;
  return (u % 10u) + (v % 10u) + (kind == 2 ? 1u : 0u);
}

int ego_sssp(uint32_t n, const uint64_t *out_ptr, const uint32_t *out_col, uint32_t root,
             int kind, uint32_t *dist) {
  const uint32_t MAXD = UINT32_MAX - 1u;
  for (uint32_t v = 0; v < n; v++) dist[v] = MAXD;
  if (root >= n) return 0;
  uint32_t *queue = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
  unsigned char *inq = (unsigned char *)calloc(n, 1);
  uint64_t head = 0, tail = 0, pops = 0;
  dist[root] = 0;
  queue[tail++ % n] = root;
  inq[root] = 1;
  while (head < tail) {
    const uint32_t u = queue[head++ % n];
    inq[u] = 0;
    pops++;
    for (uint64_t k = out_ptr[u]; k < out_ptr[u + 1]; k++) {
      const uint32_t v = out_col[k];
      if (v == u) continue; /* sssp_pushpull.c:56,84 skip self loops */
      const uint64_t nd = (uint64_t)dist[u] + ego_sssp_weight(u, v, kind);
      if (nd < dist[v]) {
        dist[v] = (uint32_t)nd;
        if (!inq[v]) {
          inq[v] = 1;
          queue[tail++ % n] = v;
        }
      }
    }
  }
  free(queue);
  free(inq);
  return (int)(pops > 0x7fffffff ? 0x7fffffff : pops);
}

int ego_check_sssp_parents(uint32_t n, const uint64_t *out_ptr, const uint32_t *out_col,
                           uint32_t root, int kind, const uint32_t *dist,
                           const uint32_t *parent) {
  const uint32_t MAXD = UINT32_MAX - 1u;
  int bad = 0;
  for (uint32_t v = 0; v < n; v++) {
    if (v == root) {
      if (parent[v] != root || dist[v] != 0) bad++;
      continue;
    }
    if (dist[v] == MAXD) {
      if (parent[v] != MAXD) bad++;
      continue;
    }
    const uint32_t p = parent[v];
    if (p >= n || dist[p] == MAXD ||
        (uint64_t)dist[p] + ego_sssp_weight(p, v, kind) != dist[v]) {
      bad++;
      continue;
    }
    int found = 0;
    for (uint64_t k = out_ptr[p]; k < out_ptr[p + 1] && !found; k++) found = out_col[k] == v;
    if (!found) bad++;
  }
  return bad;
}

/* Legality of a BFS tree with the parent encoding of bfs_numa.c:79-83,257:
 * parent[root] = root, unreached = n, and for every other reached v the edge
 * parent[v] -> v exists and dist[parent[v]] = dist[v] - 1.  Returns the number
 * of violations. */
int ego_check_bfs_parents(uint32_t n, const uint64_t *out_ptr,
                          const uint32_t *out_col, uint32_t root,
                          const uint32_t *dist, const uint32_t *parent) {
  /* tree edges found in one pass over the rows: entry u -> v is v's tree
   * edge when parent[v] == u (a scan of the parent's row per vertex would
   * cost the hub's degree for every child of a hub) */
  unsigned char *edge = (unsigned char *)calloc(n ? n : 1, 1);
  if (!edge) return -1;
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < (int64_t)n; u++)
    for (uint64_t k = out_ptr[u]; k < out_ptr[u + 1]; k++) {
      const uint32_t v = out_col[k];
      if (v < n && parent[v] == (uint32_t)u) {
#pragma omp atomic write
        edge[v] = 1;
      }
    }
  int bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const uint32_t v = (uint32_t)i;
    if (v == root) {
      if (parent[v] != root || dist[v] != 1) bad++;
      continue;
    }
    if (dist[v] == 0) {
      if (parent[v] != n) bad++;
      continue;
    }
    const uint32_t p = parent[v];
    if (p >= n || dist[p] + 1 != dist[v] || !edge[v]) bad++;
  }
  free(edge);
  return bad;
}

/* ---- full-size checks (graphs past 2^32 stored entries) ----
 * These read the input records, not a built layout, so a layout is checked
 * against what it must hold without a second full sort on the host.  With
 * both = 1 the stored entries are the records followed by their reversed
 * copies (-u, init_all.c:102-106). */

/* Row content of every vertex from the records: cnt[v] entries, sum[v] =
 * sum of their columns, xr[v] = xor of their columns. */
void ego_record_row_sums(uint32_t n, uint64_t m, const uint32_t *src,
                         const uint32_t *dst, int both, uint64_t *cnt,
                         uint64_t *sum, uint32_t *xr) {
  memset(cnt, 0, (size_t)n * sizeof(uint64_t));
  memset(sum, 0, (size_t)n * sizeof(uint64_t));
  memset(xr, 0, (size_t)n * sizeof(uint32_t));
  for (int pass = 0; pass < (both ? 2 : 1); pass++) {
    const uint32_t *key = pass ? dst : src, *val = pass ? src : dst;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)m; i++) {
      const uint32_t v = key[i], c = val[i];
#pragma omp atomic
      cnt[v] += 1;
#pragma omp atomic
      sum[v] += c;
#pragma omp atomic
      xr[v] ^= c;
    }
  }
}

/* The same row content read from a CSR (ptr[n+1], col). */
void ego_csr_row_sums(uint32_t n, const uint64_t *ptr, const uint32_t *col,
                      uint64_t *sum, uint32_t *xr) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; v++) {
    uint64_t s = 0;
    uint32_t x = 0;
    for (uint64_t k = ptr[v]; k < ptr[v + 1]; k++) {
      s += col[k];
      x ^= col[k];
    }
    sum[v] = s;
    xr[v] = x;
  }
}

/* Rows of the vertices sel[0..nsel) (ascending, distinct) in stored order:
 * row v lists dst[i] for the records with src[i] == v in input order, then
 * (both) src[i] for the records with dst[i] == v: the rows of a stable sort
 * of the stored entries by source (init_all.c:629-697).  out_ptr[nsel+1];
 * out_col (may be NULL to count only) holds out_ptr[nsel] entries. */
void ego_rows_of(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                 int both, uint32_t nsel, const uint32_t *sel, uint64_t *out_ptr,
                 uint32_t *out_col) {
  int32_t *slot = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  for (uint32_t v = 0; v < n; v++) slot[v] = -1;
  for (uint32_t j = 0; j < nsel; j++) slot[sel[j]] = (int32_t)j;
  const int T = omp_get_max_threads();
  const int passes = both ? 2 : 1;
  /* counts per (pass, thread block, selected vertex) */
  uint64_t *c = (uint64_t *)calloc((size_t)passes * T * (nsel ? nsel : 1), sizeof(uint64_t));
  for (int pass = 0; pass < passes; pass++) {
    const uint32_t *key = pass ? dst : src;
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      const uint64_t lo = m * (uint64_t)t / (uint64_t)T, hi = m * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t *ct = c + ((size_t)pass * T + t) * nsel;
      for (uint64_t i = lo; i < hi; i++) {
        const int32_t j = slot[key[i]];
        if (j >= 0) ct[j]++;
      }
    }
  }
  /* row j: pass 0 blocks in order, then pass 1 blocks: exclusive offsets */
  uint64_t run = 0;
  for (uint32_t j = 0; j < nsel; j++) {
    out_ptr[j] = run;
    for (int pass = 0; pass < passes; pass++)
      for (int t = 0; t < T; t++) {
        uint64_t *x = c + ((size_t)pass * T + t) * nsel + j;
        const uint64_t v = *x;
        *x = run;
        run += v;
      }
  }
  out_ptr[nsel] = run;
  if (out_col) {
    for (int pass = 0; pass < passes; pass++) {
      const uint32_t *key = pass ? dst : src, *val = pass ? src : dst;
#pragma omp parallel num_threads(T)
      {
        const int t = omp_get_thread_num();
        const uint64_t lo = m * (uint64_t)t / (uint64_t)T, hi = m * (uint64_t)(t + 1) / (uint64_t)T;
        uint64_t *ct = c + ((size_t)pass * T + t) * nsel;
        for (uint64_t i = lo; i < hi; i++) {
          const int32_t j = slot[key[i]];
          if (j >= 0) out_col[ct[j]++] = val[i];
        }
      }
    }
  }
  free(c);
  free(slot);
}

/* Single-source BFS invariants of levels dist (root 1, unreached 0) over the
 * edges of the records (both: each record in both directions): dist[root] =
 * 1; for every edge u -> v with u reached, v is reached and dist[v] <=
 * dist[u] + 1; every reached v != root has an in-neighbour u with dist[u] =
 * dist[v] - 1.  Together these make dist - 1 the hop distance from root.
 * Returns the number of violations. */
int64_t ego_check_bfs_levels(uint32_t n, uint64_t m, const uint32_t *src,
                             const uint32_t *dst, int both, uint32_t root,
                             const uint32_t *dist) {
  unsigned char *pred = (unsigned char *)calloc(n ? n : 1, 1);
  int64_t bad = dist[root] != 1;
  for (int pass = 0; pass < (both ? 2 : 1); pass++) {
    const uint32_t *from = pass ? dst : src, *to = pass ? src : dst;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = 0; i < (int64_t)m; i++) {
      const uint32_t u = from[i], v = to[i];
      if (!dist[u]) continue;
      if (!dist[v] || dist[v] > dist[u] + 1) bad++;
      if (dist[v] == dist[u] + 1) {
#pragma omp atomic write
        pred[v] = 1;
      }
    }
  }
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t v = 0; v < (int64_t)n; v++)
    if ((uint32_t)v != root && dist[v] && !pred[v]) bad++;
  free(pred);
  return bad;
}
