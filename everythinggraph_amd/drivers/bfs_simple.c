/*
 * bfs_simple -- drop-in for bfs_simple.c over libeg.
 * Direction-optimising BFS from -p (the reference hard-codes root 0, R8) with
 * true levels in the reference encoding (root 1, unreached 0; R9 not
 * reproduced).  -a 1 starts in pull mode, -r disables switching
 * (bfs_simple.c:191-197).  Grid modes are refused as in bfs_simple.c:256-259.
 */
#include <stdlib.h>

#include "eg_driver.h"

static uint32_t *dist;

static void bfs_construct(void) {
  dist = (uint32_t *)calloc(NB_NODES, sizeof(uint32_t));
  if (load_mode < 2 || load_mode == 6) {
    printf("Use the BFS grid file, don't know how to run BFS here with this layout\n");
    exit(1);
  }
}

static void bfs_destruct(void) {
  uint64_t nodes_discovered = 0;
  for (uint64_t i = 0; i < NB_NODES; i++)
    if (dist[i] != 0) nodes_discovered++;
  printf("Total nodes discovered:: %lu\n", (unsigned long)nodes_discovered);
  eg_dump(dist, (size_t)NB_NODES * sizeof(uint32_t));
  free(dist);
}

static void bfs_reset(struct node *nodes) { (void)nodes; }

static void bfs(struct node *nodes) {
  (void)nodes;
  eg_bfs_stats st;
  const double t0 = eg_now();
  eg_check(eg_bfs(G, BFS_ROOT, mode == PULL, switch_mode, dist, NULL, &st), "bfs");
  const double t1 = eg_now();
  printf("Done in %d iterations and found %lu nodes (%d pull levels)\n", st.levels,
         (unsigned long)st.reached, st.pull_levels);
  printf("#Algo time ( %.3f sec) device %.3f ms, %.3f GTEPS\n", t1 - t0, st.device_ms,
         st.device_ms > 0 ? (double)st.edges_traversed / (st.device_ms * 1e6) : 0.0);
}

struct algo_func current_algo = {
    .reset = bfs_reset, .main = bfs, .construct = bfs_construct, .destruct = bfs_destruct};
