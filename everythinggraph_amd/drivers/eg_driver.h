/*
 * eg_driver.h -- the reference's plugin host in C over libeg.
 *
 * Each drop-in binary (pagerank_simple, prgrid_cilk, spmv, bfs_simple,
 * bfsgrid_cilk, wcc) links eg_driver.c, which holds main(): the flag parsing
 * of random.c:50-116, the graph load that replaces init() (init_all.c:791),
 * and the lifecycle construct -> reset -> main -> destruct of random.c:128-143.
 * The algorithm file defines `struct algo_func current_algo` (random.h:64-69).
 */
#ifndef EG_DRIVER_H
#define EG_DRIVER_H

#include <stdint.h>
#include <stdio.h>

#include "eg.h"

struct node; /* the reference's vertex record; libeg keeps it on the device */

struct algo_func {
  void (*reset)(struct node *);
  void (*main)(struct node *);
  void (*construct)(void);
  void (*destruct)(void);
};
extern struct algo_func current_algo;

typedef enum { PUSH, PULL } algo_mode; /* random.h:115-120 */

/* Globals of random.c:4-24 and init_all.c:6-23 that the algorithms read. */
extern eg_graph *G;
extern char *filename;
extern uint32_t NB_NODES;
extern uint64_t nb_edges;
extern uint32_t P;
extern uint32_t BFS_ROOT;
extern int load_mode, isSymmetric, createUndir, weighted_graph, switch_mode;
extern algo_mode mode;
/* Options added here (defaults reproduce the reference). */
extern int opt_iterations; /* -i: PageRank iterations (10, pagerank_simple.c:115) */
extern double opt_init;    /* -I: initial rank (0.15, pagerank_simple.c:95); 0 = 1/n */
extern int opt_dangling;   /* -D: 1 = redistribute dangling mass uniformly */

/* Exits with the library's message when rc != EG_OK. */
void eg_check(int rc, const char *what);
/* Writes the result array of the algorithm (the global the reference leaves
 * behind: dist, components, rank, y) as raw bytes to the file named by the
 * EG_DUMP environment variable, when it is set; a no-op otherwise. */
void eg_dump(const void *data, size_t bytes);
double eg_now(void);

#endif
