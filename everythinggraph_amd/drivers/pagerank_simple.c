/*
 * pagerank_simple -- drop-in for pagerank_simple.c over libeg.
 * Pull PageRank over in-edges for the adjacency modes (pr_algo_pull,
 * pagerank_simple.c:72-84), the edge-array form for modes > 6
 * (pr_algo_edgecentric, :40-61).  The push (-a 0) and pull (-a 1) variants
 * compute the same recurrence, so both run the pull kernel.
 */
#include <stdlib.h>

#include "eg_driver.h"

static float *rank_out;

static void pr_construct(void) { rank_out = (float *)malloc((size_t)NB_NODES * sizeof(float)); }
static void pr_destruct(void) { free(rank_out); }
static void pr_reset(struct node *nodes) { (void)nodes; }

/* pagerank_simple.c:143-162 (its main hook is named prgrid). */
static void prgrid(struct node *nodes) {
  (void)nodes;
  eg_pr_opts o = {0.85, opt_iterations,
                  opt_dangling ? EG_PR_DANGLING_UNIFORM : EG_PR_DANGLING_DROP, opt_init,
                  EG_LAYOUT_AUTO};
  double *l1 = (double *)calloc(opt_iterations > 0 ? opt_iterations : 1, sizeof(double));
  const double t0 = eg_now();
  eg_check(eg_pagerank(G, &o, rank_out, NULL, l1), "pagerank");
  const double t1 = eg_now();
  for (int it = 0; it < opt_iterations; it++) printf(" Iter %d delta L1 %.3e\n", it + 1, l1[it]);
  printf("#Algo time ( %.3f sec) \n", t1 - t0);
  for (uint32_t i = 0; i < 10 && i < NB_NODES; i++) printf("Rank[%u] = %.5f\n", i, rank_out[i]);
  free(l1);
  eg_dump(rank_out, (size_t)NB_NODES * sizeof(float));
}

struct algo_func current_algo = {
    .reset = pr_reset, .main = prgrid, .construct = pr_construct, .destruct = pr_destruct};
