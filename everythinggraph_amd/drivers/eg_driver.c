/*
 * eg_driver.c -- main() of the drop-in binaries (random.c:50-145 over libeg).
 */
#define _GNU_SOURCE
#include "eg_driver.h"

#include <getopt.h>
#include <locale.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

eg_graph *G = NULL;
char *filename = NULL;
uint32_t NB_NODES = 0;
uint64_t nb_edges = 0;
uint32_t P = 0;
uint32_t BFS_ROOT = 0;
int load_mode = 2, isSymmetric = 0, createUndir = 0, weighted_graph = 0, switch_mode = 1;
algo_mode mode = PUSH;
int opt_iterations = 10;
double opt_init = 0.15;
int opt_dangling = 0;

double eg_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void eg_check(int rc, const char *what) {
  if (rc != EG_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, eg_last_error());
    exit(1);
  }
}

void eg_dump(const void *data, size_t bytes) {
  const char *path = getenv("EG_DUMP");
  if (!path || !*path) return;
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(data, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot write %s\n", path);
    exit(1);
  }
  if (fclose(f) != 0) {
    fprintf(stderr, "cannot write %s\n", path);
    exit(1);
  }
}

/* EG_TUNE="param=value,param=value": kernel shapes for the graph handle
 * (eg_tune_set), applied before the file is loaded; no reference
 * counterpart.  A malformed entry or an unknown parameter is an error. */
static void apply_tune_env(eg_graph *g) {
  const char *env = getenv("EG_TUNE");
  if (!env || !*env) return;
  char buf[512];
  if (strlen(env) >= sizeof(buf)) {
    fprintf(stderr, "EG_TUNE is too long\n");
    exit(1);
  }
  strcpy(buf, env);
  for (char *item = strtok(buf, ","); item; item = strtok(NULL, ",")) {
    char *eq = strchr(item, '=');
    char *end = NULL;
    if (!eq || eq == item) {
      fprintf(stderr, "EG_TUNE: expected param=value, got '%s'\n", item);
      exit(1);
    }
    *eq = '\0';
    const long v = strtol(eq + 1, &end, 10);
    if (end == eq + 1 || *end != '\0') {
      fprintf(stderr, "EG_TUNE: bad value for %s\n", item);
      exit(1);
    }
    eg_check(eg_tune_set(g, item, (int32_t)v), "EG_TUNE");
  }
}

static void usage(void) {
  /* random.c:32-48, plus the options added by this implementation. */
  printf("Usage: ./random -f <graph file> -n <nb_nodes> [-m -u -r -L -w -b -a [mode] -p [bfs_root] -s -N ]\n");
  printf("\t-u: create undirected on load (For example for WCC running with load mode 2 or 3)\n");
  printf("\t-r [switch]: switch between PUSH and PULL where applicable (default 1)\n");
  printf("\t-n: NB_NODES, e.g.,\n");
  printf("\t\trmat20 1048576\n\t\trmat23 8388608\n\t\trmat25 33554432\n\t\trmat27 134217728\n");
  printf("\t-a [mode] : Algo mode, 0 PUSH, 1 PULL, (default 0)\n");
  printf("\t-w: weighted input graph\n");
  printf("\t-p [bfs_root]: BFS & SSSP root\n");
  printf("\t-s:  symmetric graph, if not given set of incoming edges will be created \n");
  printf("\t-N: accepted and ignored (NUMA variants are not part of this build)\n");
  printf("\t-m [0..9] : load mode (default 2): 0 grid sorted; 1 grid nosort; 2 adj sorted; 3 adj sorted;"
         " 4 adj count sort; 6 grid cells by src; 7 edge array by src; 8 by src,dst; 9 raw\n");
  printf("\t-P [p]: grid partitions per side (default: (n*8/1024)/20, 256 if > 2000, 4 if 0)\n");
  printf("\t-i [it]: PageRank iterations (default 10)\n");
  printf("\t-I [r0]: initial rank (default 0.15; 0 means 1/n)\n");
  printf("\t-D: add the dangling mass back uniformly (default: dropped, as the reference)\n");
  fflush(stdout);
  _exit(-1);
}

int main(int argc, char **argv) {
  setlocale(LC_NUMERIC, "");
  int c;
  while ((c = getopt(argc, argv, "f:n:m:uwLa:p:srNhP:i:I:D")) != -1) {
    switch (c) {
      case 'f': filename = optarg; break;
      case 'm':
        load_mode = atoi(optarg);
        printf("Load mode = %d\n", load_mode);
        break;
      case 'n':
        NB_NODES = (uint32_t)strtoul(optarg, NULL, 10);
        break;
      case 'u':
        createUndir = 1;
        printf("#Creating Undir\n");
        break;
      case 'r': switch_mode = 0; break;
      case 's':
        isSymmetric = 1;
        printf(" no in-edges created\n");
        break;
      case 'a': mode = atoi(optarg) == 0 ? PUSH : PULL; break;
      case 'w': weighted_graph = 1; break;
      case 'p': BFS_ROOT = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 'N': break;
      case 'P': P = (uint32_t)strtoul(optarg, NULL, 10); break;
      case 'i': opt_iterations = atoi(optarg); break;
      case 'I': opt_init = atof(optarg); break;
      case 'D': opt_dangling = 1; break;
      case 'h': usage(); break;
      case '?':
        if (optopt == 'f' || optopt == 'n' || optopt == 'a')
          fprintf(stderr, "Option -%c requires an argument.\n", optopt);
        else
          fprintf(stderr, "Unknown option `-%c'.\n", optopt);
        usage();
        break;
      default: usage();
    }
  }
  if (!NB_NODES || !filename) usage();

  const uint32_t flags = (createUndir ? EG_UNDIRECTED : 0u) | (isSymmetric ? EG_SYMMETRIC : 0u) |
                         (weighted_graph ? EG_WEIGHTED : 0u);
  double t0 = eg_now();
  eg_check(eg_graph_create(NB_NODES, flags, load_mode, P, &G), "graph create");
  eg_graph_info info;
  eg_check(eg_graph_get_info(G, &info), "graph info");
  P = info.P;
  printf("P = %u\n", P);
  eg_device_info dev;
  if (eg_device_query(&dev) == EG_OK)
    printf("#Running on %s (%d CUs, %.1f GiB HBM)\n", dev.name, dev.compute_units,
           (double)dev.hbm_bytes / (1u << 30));
  apply_tune_env(G);
  eg_check(eg_graph_load_file(G, filename), "graph load");
  eg_check(eg_graph_get_info(G, &info), "graph info");
  nb_edges = info.m;
  printf("#NB edges = %lu\n", (unsigned long)info.m_in);
  printf("#Layout build (device) %.3f ms\n", info.build_ms);
  double t1 = eg_now();
  printf("#Total loading time ( %fs )\n", t1 - t0);

  current_algo.construct();
  t0 = eg_now();
  current_algo.reset(NULL);
  double t2 = eg_now();
  printf("#Total reset  time ( %fs )\n", t2 - t0);
  current_algo.main(NULL);
  t1 = eg_now();
  printf("#Total algo time ( %fs )\n", t1 - t0);
  current_algo.destruct();
  eg_graph_free(G);
  return 0;
}
