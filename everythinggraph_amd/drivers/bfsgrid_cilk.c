/*
 * bfsgrid_cilk -- drop-in for bfsgrid_cilk.c over libeg: level-synchronous
 * BFS over the grid cells from -p, root level 1 (bfsgrid_cilk.c:175-204).
 * Only the grid modes are accepted (bfsgrid_cilk.c:94-97).
 */
#include <stdlib.h>

#include "eg_driver.h"

static uint32_t *dist;

static void bfsgrid_construct(void) {
  if (load_mode != 0 && load_mode != 1 && load_mode != 6) {
    printf("To run this, you need to give grid as the data layout, otherwise work with pagerank_simple\n");
    exit(1);
  }
  dist = (uint32_t *)calloc(NB_NODES, sizeof(uint32_t));
}

static void bfsgrid_destruct(void) {
  uint64_t nodes_discovered = 0;
  for (uint64_t i = 0; i < NB_NODES; i++)
    if (dist[i] != 0) nodes_discovered++;
  printf("Total nodes discovered:: %lu\n", (unsigned long)nodes_discovered);
  eg_dump(dist, (size_t)NB_NODES * sizeof(uint32_t));
  free(dist);
}

static void bfsgrid_reset(struct node *nodes) { (void)nodes; }

static void bfsgrid(struct node *nodes) {
  (void)nodes;
  eg_bfs_stats st;
  const double t0 = eg_now();
  eg_check(eg_bfs(G, BFS_ROOT, 0, 0, dist, NULL, &st), "bfs");
  const double t1 = eg_now();
  printf("Done in %d iterations\n", st.levels);
  printf("#Algo time ( %.3f sec) device %.3f ms\n", t1 - t0, st.device_ms);
}

struct algo_func current_algo = {.reset = bfsgrid_reset,
                                 .main = bfsgrid,
                                 .construct = bfsgrid_construct,
                                 .destruct = bfsgrid_destruct};
