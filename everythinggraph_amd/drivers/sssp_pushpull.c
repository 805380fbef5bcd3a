/*
 * sssp_pushpull -- drop-in for sssp_pushpull.c over libeg.
 * Single-source shortest paths from -p with the reference's generated edge
 * weights (u % 10) + (v % 10) (CREATE_WEIGHT, init_all.c:661-686).  -a 1
 * starts in pull mode and -r disables the push/pull switch
 * (sssp_pushpull.c:174-179).  Distances: root 0, unreached UINT32_MAX - 1
 * (:8); the push step adds the edge weight like the pull step (the reference
 * adds 1, :54).  Prints the lines of sssp_pushpull.c:226-235 and :113-123.
 */
#include <stdlib.h>

#include "eg_driver.h"

static uint32_t *dist;

static void sssp_construct(void) {
  dist = (uint32_t *)malloc((size_t)NB_NODES * sizeof(uint32_t));
  if (load_mode < 2 || load_mode > 4) {
    printf("SSSP runs over adjacency lists (load modes 2, 3, 4)\n");
    exit(1);
  }
}

static void sssp_destruct(void) {
  uint64_t nodes_discovered = 0;
  for (uint64_t i = 0; i < NB_NODES; i++)
    if (dist[i] != EG_SSSP_UNREACHED) nodes_discovered++;
  printf("Total nodes discovered:: %lu\n", (unsigned long)nodes_discovered);
  eg_dump(dist, (size_t)NB_NODES * sizeof(uint32_t));
  free(dist);
}

static void sssp_reset(struct node *nodes) { (void)nodes; }

static void sssp(struct node *nodes) {
  (void)nodes;
  eg_sssp_stats st;
  const double t0 = eg_now();
  eg_check(eg_sssp(G, BFS_ROOT, mode == PULL, switch_mode, EG_SSSP_WEIGHT_REF, dist, NULL, &st),
           "sssp");
  const double t1 = eg_now();
  printf("#Algo time ( %.3f sec) device %.3f ms, %d rounds (%d pull)\n", t1 - t0, st.device_ms,
         st.iterations, st.pull_iterations);
  uint32_t t = 0;
  for (uint32_t i = 0; i < NB_NODES; i++)
    if (dist[i] != EG_SSSP_UNREACHED && dist[i] > t) t = dist[i];
  printf("Max dist = %d and dist[1] is %d\n", (int)t, NB_NODES > 1 ? (int)dist[1] : -1);
}

struct algo_func current_algo = {
    .reset = sssp_reset, .main = sssp, .construct = sssp_construct, .destruct = sssp_destruct};
