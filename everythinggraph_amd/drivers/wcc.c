/*
 * wcc -- drop-in for wcc.c over libeg: min-label propagation to the fixpoint
 * (wcc.c:186-286) and the component summary of validate() (:138-182).  The
 * layouts the reference accepts are kept (grid mode 0, edge array mode 8,
 * adjacency modes 2/3/4; wcc.c:278-280 otherwise).
 */
#include <stdlib.h>

#include "eg_driver.h"

static int64_t *components;

static void wcc_construct(void) {
  components = (int64_t *)malloc((size_t)NB_NODES * sizeof(int64_t));
}
static void wcc_destruct(void) { free(components); }
static void wcc_reset(struct node *nodes) {
  (void)nodes;
  for (uint32_t i = 0; i < NB_NODES; i++) components[i] = i;
}

static void validate(void) {
  uint32_t *diff_comp = (uint32_t *)calloc(NB_NODES, sizeof(uint32_t));
  for (uint32_t i = 0; i < NB_NODES; i++) diff_comp[components[i]]++;
  uint32_t max_comp = 0, total_comp = 0;
  for (uint32_t i = 0; i < NB_NODES; i++) {
    if (diff_comp[i] > max_comp) max_comp = diff_comp[i];
    if (diff_comp[i] != 0) total_comp++;
  }
  printf("Total comp %u , max comp %u \n", total_comp, max_comp);
  free(diff_comp);
}

static void wcc(struct node *nodes) {
  (void)nodes;
  if (!(load_mode == 0 || load_mode == 8 || load_mode == 2 || load_mode == 3 || load_mode == 4)) {
    printf("Don't know how to run with this layout\n");
    exit(1);
  }
  int32_t rounds = 0;
  const double t0 = eg_now();
  eg_check(eg_wcc(G, components, &rounds), "wcc");
  const double t1 = eg_now();
  printf("Iter %d\n", rounds);
  printf("Algo time %f \n", t1 - t0);
  validate();
  eg_dump(components, (size_t)NB_NODES * sizeof(int64_t));
}

struct algo_func current_algo = {
    .reset = wcc_reset, .main = wcc, .construct = wcc_construct, .destruct = wcc_destruct};
