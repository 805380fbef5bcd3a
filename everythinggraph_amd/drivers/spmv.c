/*
 * spmv -- drop-in for spmv.c over libeg.
 * One product y[v] = sum_{u->v} w(u,v) x[u] over in-edges (spmv_algo,
 * spmv.c:98-106), with FP32 vectors and the input weights (w = 1 when the
 * graph is unweighted) instead of u32 vectors and the constant 0.001 (R7).
 * x[i] = i as in spmv.c:126-127.
 */
#include <stdlib.h>

#include "eg_driver.h"

static float *x, *y;

static void spmv_construct(void) {
  x = (float *)malloc((size_t)NB_NODES * sizeof(float));
  y = (float *)malloc((size_t)NB_NODES * sizeof(float));
}
static void spmv_destruct(void) {
  free(x);
  free(y);
}
static void spmv_reset(struct node *nodes) {
  (void)nodes;
  for (uint32_t i = 0; i < NB_NODES; i++) x[i] = (float)i;
}

static void spmv(struct node *nodes) {
  (void)nodes;
  const double t0 = eg_now();
  eg_check(eg_spmv(G, EG_LAYOUT_AUTO, x, y), "spmv");
  const double t1 = eg_now();
  double sum = 0.0;
  for (uint32_t i = 0; i < NB_NODES; i++) sum += y[i];
  printf("#Algo time ( %.3f sec) \n", t1 - t0);
  printf("sum(y) = %.6e\n", sum);
  for (uint32_t i = 0; i < 10 && i < NB_NODES; i++) printf("y[%u] = %.5f\n", i, y[i]);
  eg_dump(y, (size_t)NB_NODES * sizeof(float));
}

struct algo_func current_algo = {
    .reset = spmv_reset, .main = spmv, .construct = spmv_construct, .destruct = spmv_destruct};
