/*
 * prgrid_cilk -- drop-in for prgrid_cilk.c over libeg.
 * Grid modes 0/1/6 run the cell-streaming kernel (prgrid_algo,
 * prgrid_cilk.c:62-76), edge-array modes the edge form (prgrid_algo_col,
 * :79-88); the adjacency modes, where the reference reads freed memory (R5),
 * run the pull kernel.
 */
#include <stdlib.h>

#include "eg_driver.h"

static float *rank_out;

static void prgrid_construct(void) {
  rank_out = (float *)malloc((size_t)NB_NODES * sizeof(float));
}
static void prgrid_destruct(void) { free(rank_out); }
static void prgrid_reset(struct node *nodes) { (void)nodes; }

static void prgrid(struct node *nodes) {
  (void)nodes;
  eg_pr_opts o = {0.85, opt_iterations,
                  opt_dangling ? EG_PR_DANGLING_UNIFORM : EG_PR_DANGLING_DROP, opt_init,
                  EG_LAYOUT_AUTO};
  const double t0 = eg_now();
  eg_check(eg_pagerank(G, &o, rank_out, NULL, NULL), "pagerank");
  const double t1 = eg_now();
  printf("Iter %d time %f \n", opt_iterations, t1 - t0);
  printf("#Algo time ( %.3f sec) \n", t1 - t0);
  for (uint32_t i = 0; i < 10 && i < NB_NODES; i++) printf("Rank[%u] = %.5f\n", i, rank_out[i]);
  eg_dump(rank_out, (size_t)NB_NODES * sizeof(float));
}

struct algo_func current_algo = {.reset = prgrid_reset,
                                 .main = prgrid,
                                 .construct = prgrid_construct,
                                 .destruct = prgrid_destruct};
