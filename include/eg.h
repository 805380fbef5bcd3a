/*
 * eg.h -- C ABI of libeg, the MI355X-native implementation of EverythingGraph's
 * data-parallel path (layout builders, pull PageRank / SpMV, direction-
 * optimising BFS, min-label WCC) plus its batched and serving forms.
 *
 * The reference is a link-time plugin system: each algorithm binary defines
 * `struct algo_func current_algo = {reset, main, construct, destruct}`
 * (random.h:64-69) and the driver (random.c:124-143) calls it after `init()`
 * (init_all.c:791) has built the graph selected by `-m load_mode` into process
 * globals.  libeg replaces `init()` and the algorithm bodies behind plain C
 * calls: a graph handle owns the device-resident layout and a HIP stream, and
 * results come back in caller-owned host buffers in the reference encodings.
 *
 * Conventions: every call returns EG_OK (0) or a negative EG_E* code and
 * eg_last_error() describes the last failure of the calling thread.  A handle
 * is not thread-safe.  Vertex ids are uint32 and must be < n (EG_ERANGE
 * otherwise).  Layouts with fewer than 2^32 stored entries use 32-bit
 * offsets.  From 2^32 stored entries on, the adjacency modes (2/3/4) keep
 * 64-bit row offsets (the reference's u64 edge offsets, random.h:87-93):
 * eg_bfs, eg_graph_get_degrees / _csr / _csc and the info calls run on such a
 * graph, every other call returns EG_ENOTSUP; the other modes refuse the load.
 */
#ifndef EG_H
#define EG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference exits or asserts: utils.h:98-102) ---- */
#define EG_OK 0
#define EG_EINVAL (-1)   /* bad argument */
#define EG_ENOMEM (-2)   /* device or host allocation failed */
#define EG_EIO (-3)      /* edge file unreadable or misaligned */
#define EG_EDEVICE (-4)  /* HIP runtime error or no GPU */
#define EG_ERANGE (-5)   /* a vertex id >= n */
#define EG_ENOTSUP (-6)  /* layout / algorithm combination not supported */
#define EG_ESTATE (-7)   /* graph not loaded yet */
#define EG_EAGAIN (-8)   /* serving: no capacity now */
#define EG_ETIMEDOUT (-9) /* serving: wait timed out */

/* ---- graph flags: the -u, -s and -w switches of random.c:57-98 ---- */
#define EG_UNDIRECTED 1u /* -u: append reversed records (init_all.c:102-106) */
#define EG_SYMMETRIC 2u  /* -s: do not build in-edges (init_all.c:714) */
#define EG_WEIGHTED 4u   /* -w: 12-byte [src][dst][f32 w] records (init_all.h:8-14) */

/* ---- load modes: the -m numbering of init_all.c:855-888 ---- */
#define EG_MODE_GRID_SORT 0     /* init_grid_sort(0,0) */
#define EG_MODE_GRID_NOSORT 1   /* init_grid_nosort */
#define EG_MODE_ADJ_SORT 2      /* init_adj_sort: CSR (+CSC) */
#define EG_MODE_ADJ_SORT_FULL 3 /* same as 2 (init_all.c:864-869) */
#define EG_MODE_ADJ_COUNT 4     /* init_adj_count: count-sort CSR (+CSC) */
#define EG_MODE_ADJ_DYNAMIC 5   /* init_adj_dynamic: frees what it builds; EG_ENOTSUP */
#define EG_MODE_GRID_SORT_SRC 6 /* init_grid_sort(full,1) */
#define EG_MODE_EDGES_SRC 7     /* init_edgelist(.,0,1): sorted by src */
#define EG_MODE_EDGES_SRC_DST 8 /* init_edgelist(.,1,1): by src then dst */
#define EG_MODE_EDGES_RAW 9     /* init_edgelist(.,0,0): input order */

typedef struct eg_graph eg_graph;

typedef struct {
  char name[64];
  int32_t compute_units;
  int32_t gcn_arch; /* e.g. 950 */
  uint64_t hbm_bytes;
  uint64_t free_bytes;
  double peak_hbm_gbs; /* published peak used for roofline fractions */
} eg_device_info;

typedef struct {
  uint32_t n;        /* vertices (NB_NODES, -n) */
  uint32_t P;        /* grid partitions per side */
  uint32_t flags;    /* EG_UNDIRECTED | EG_SYMMETRIC | EG_WEIGHTED */
  int32_t load_mode; /* -m */
  uint64_t m_in;     /* input records (nb_edges after init_all.c:821) */
  uint64_t m;        /* stored entries: 2*m_in with EG_UNDIRECTED */
  int32_t has_csr, has_csc, has_grid, has_edgelist;
  double build_ms;   /* device time of the last layout build */
  uint64_t device_bytes;
} eg_graph_info;

/* Synthetic input generated on the device (no reference counterpart; the
 * reference only reads files).  kind 0 = R-MAT(scale, a, b, c), 1 = uniform
 * Erdos-Renyi pairs over [0, n). Identical to the oracle's generator. */
typedef struct {
  int32_t kind;
  int32_t scale;      /* R-MAT: n = 2^scale */
  uint64_t m_in;      /* records to generate */
  uint64_t seed;      /* graph seed */
  double a, b, c;     /* R-MAT quadrant probabilities */
  uint64_t weight_seed; /* used when the graph is EG_WEIGHTED */
} eg_gen_params;

/* PageRank options.  The reference hard-codes d = 0.85 (pagerank_simple.c:4),
 * 10 iterations (:115) and prev = 0.15 (:95). */
#define EG_PR_DANGLING_DROP 0    /* reference recurrence: mass of sinks dropped */
#define EG_PR_DANGLING_UNIFORM 1 /* add d * dangling / n to every vertex */
#define EG_LAYOUT_AUTO 0
#define EG_LAYOUT_ADJ 1   /* pull over in-edges (CSC, or CSR when symmetric) */
#define EG_LAYOUT_GRID 2  /* 2-D grid cells */
#define EG_LAYOUT_EDGES 3 /* edge array */
typedef struct {
  double damping;     /* 0.85 */
  int32_t iterations; /* 10 in the reference */
  int32_t dangling;   /* EG_PR_DANGLING_* */
  double init;        /* initial rank of every vertex (0 -> 1/n) */
  int32_t layout;     /* EG_LAYOUT_* */
} eg_pr_opts;

typedef struct {
  int32_t levels;          /* number of BFS levels (frontiers) */
  int32_t pull_levels;     /* levels run bottom-up */
  uint64_t reached;        /* vertices with dist != 0 */
  uint64_t edges_traversed; /* input records inside the traversed component */
  double device_ms;        /* device time of the traversal */
} eg_bfs_stats;

/* ---- library and device ---- */
const char *eg_last_error(void);
int eg_device_query(eg_device_info *out);
const char *eg_version(void);

/* ---- graph lifecycle (replaces init(), init_all.c:791-893) ---- */
/* P = 0 selects the reference formula (random.c:69-71). */
int eg_graph_create(uint32_t n, uint32_t flags, int32_t load_mode, uint32_t P,
                    eg_graph **out);
/* init_all.c:812-850: read the binary edge file (8- or 12-byte records; a
 * trailing partial record is dropped as st_size/item_size does) and build the
 * layout of the handle's load mode on the device. */
int eg_graph_load_file(eg_graph *g, const char *path);
/* Same from raw record bytes already in host memory. */
int eg_graph_load_records(eg_graph *g, const void *records, uint64_t nbytes);
/* Same from separate host arrays (w may be NULL). */
int eg_graph_load_edges(eg_graph *g, uint64_t m_in, const uint32_t *src,
                        const uint32_t *dst, const float *w);
/* Same from synthetic records generated in device memory. */
int eg_graph_generate(eg_graph *g, const eg_gen_params *p);
/* Generates p->m_in records on the device and copies them to host arrays
 * (w may be NULL): the input of an edge file, for tools that write one. */
int eg_generate_edges(const eg_gen_params *p, uint32_t n, uint32_t *src, uint32_t *dst,
                      float *w);
int eg_graph_free(eg_graph *g);
int eg_graph_get_info(const eg_graph *g, eg_graph_info *out);

/* One-time preparations of a loaded graph and their device memory (no
 * reference counterpart): the layout build of the load mode, and what the
 * kernels build on first use -- the degree-ordered copy of the pull
 * orientation, the binned PageRank / SpMV plan and the hubs-first BFS rows --
 * with the wall milliseconds each took (0 until built). */
typedef struct {
  double layout_ms;       /* device time of the layout build (eg_graph_info.build_ms) */
  double order_ms, binned_plan_ms, bfs_rows_ms;
  uint64_t layout_bytes;  /* eg_graph_info.device_bytes */
  uint64_t order_bytes, binned_plan_bytes, bfs_rows_bytes;
} eg_prep_info;
int eg_graph_prep_info(const eg_graph *g, eg_prep_info *out);

/* ---- layout downloads (parity checks, drop-in globals) ---- */
/* Out-degree (nb_out_edges) and in-degree (nb_in_edges) of random.h:87-93. */
int eg_graph_get_degrees(eg_graph *g, uint32_t *out_deg, uint32_t *in_deg);
/* CSR: row_ptr[n+1] (outgoing_edges offsets), col[m] (edge_array_out[].dst),
 * w[m] or NULL.  init_all.c:629-697. */
int eg_graph_get_csr(eg_graph *g, uint64_t *row_ptr, uint32_t *col, float *w);
/* CSC: col_ptr[n+1], row[m] (edge_array_in[].dst holds the src), w or NULL.
 * init_all.c:714-767. */
int eg_graph_get_csc(eg_graph *g, uint64_t *col_ptr, uint32_t *row, float *w);
/* Grid: cell_ptr[P*P+1] (row-major cells (part(src), part(dst))), and the
 * cell-ordered records.  init_all.c:252-424. */
int eg_graph_get_grid(eg_graph *g, uint64_t *cell_ptr, uint32_t *src,
                      uint32_t *dst, float *w);
/* Edge array of modes 7/8/9, init_all.c:573-628. */
int eg_graph_get_edgelist(eg_graph *g, uint32_t *src, uint32_t *dst, float *w);

/* ---- CSR build that allocates no device memory ----
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  It is the CSR of eg_graph_load_edges in load mode 2 -- the
 * records stably sorted by src (init_adj_sort, init_all.c:629-697) -- over m
 * records (src[i], dst[i]) already in device memory (append the reversed
 * copies first for -u): row_ptr[n + 1] (u32 offsets) and col[m], both device
 * memory owned by the caller.  Ids must be < n (not checked: that would need a
 * host round trip).  Every temporary lives in one fixed device array of 48 MiB
 * that each device holds until the process ends, so m is at most
 * EG_NOALLOC_MAX_ENTRIES (EG_EINVAL beyond).  The work is enqueued on stream
 * (a hipStream_t; NULL = the null stream) and the call returns without
 * waiting.  Calls on one device share its array: calls whose work overlaps on
 * one device race, so serialise them (one stream, or wait in between). */
#define EG_NOALLOC_MAX_ENTRIES 2097152u
int eg_csr_build_noalloc(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                         uint32_t *row_ptr, uint32_t *col, void *stream);

/* ---- algorithms (replace current_algo.main of each plugin) ---- */
/* pagerank_simple.c:72-133 / prgrid_cilk.c:62-141.  FP32 storage, compensated
 * sums, FP64 dangling-mass and L1-delta reductions on the device.  rank may be
 * NULL (result stays on the device).  dangling_hist / l1_hist (length
 * iterations, may be NULL) receive per iteration the dangling mass of the
 * input iterate and sum |r' - r|. */
int eg_pagerank(eg_graph *g, const eg_pr_opts *opts, float *rank,
                double *dangling_hist, double *l1_hist);
/* spmv.c:98-136 with FP32 weights: y[v] = sum_{u->v} w(u,v) x[u] (pull).
 * Unweighted graphs use w = 1.  layout as eg_pr_opts.layout. */
int eg_spmv(eg_graph *g, int32_t layout, const float *x, float *y);
/* bfs_simple.c:136-281 (CSR modes), bfsgrid_cilk.c:121-204 (grid modes),
 * bfs_edgecentric (edge modes).  start_pull = -a 1, switch_mode = !-r.
 * dist: root = 1, level L -> L+1, unreached 0.  parent (may be NULL):
 * parent[root] = root, unreached = n (bfs_numa.c:79-83).  dist may be NULL. */
int eg_bfs(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
           uint32_t *dist, uint32_t *parent, eg_bfs_stats *stats);
/* wcc.c:186-286: labels[v] = fixpoint of min-label propagation (int64, as
 * `long components[]`).  rounds may be NULL. */
int eg_wcc(eg_graph *g, int64_t *labels, int32_t *rounds);

/* sssp_pushpull.c:39-242: single-source shortest paths by frontier
 * relaxation, push over out-edges or pull over in-edges with the bfs_simple
 * switch rule (switch_mode = !-r, start_pull = -a 1).  Edge weights: the
 * reference's CREATE_WEIGHT weights (u % 10) + (v % 10) (init_all.c:661-686)
 * or unit weights.  dist: root 0, unreached EG_SSSP_UNREACHED
 * (sssp_pushpull.c:8); both directions add the edge weight (the reference's
 * push adds 1, :54).  parent (may be NULL): parent[root] = root, unreached
 * EG_SSSP_UNREACHED, else the first in-neighbour on a shortest path.
 * Adjacency modes (2/3/4) only. */
#define EG_SSSP_WEIGHT_REF 0
#define EG_SSSP_WEIGHT_UNIT 1
#define EG_SSSP_UNREACHED 0xFFFFFFFEu
typedef struct {
  int32_t iterations;      /* relaxation rounds */
  int32_t pull_iterations; /* rounds run bottom-up */
  uint64_t reached;        /* vertices with a finite distance */
  double device_ms;
} eg_sssp_stats;
int eg_sssp(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
            int32_t weights, uint32_t *dist, uint32_t *parent, eg_sssp_stats *stats);

/* ---- batched form (no reference counterpart) ---- */
/* k BFS traversals from roots[k] as multi-source sweeps of up to "ms_lanes"
 * lanes (64, 128 or 256 sources per traversal).  Levels run device-driven:
 * the kernels apply the bfs_simple.c:191-197 direction rule to the frontier
 * of all lanes and end the traversal themselves, levels are enqueued in
 * batches (replayed from a hipGraph) and the host waits once per batch, not
 * per level.  dist (may be NULL) is k x n, source-major, in the eg_bfs
 * encoding.  A traversal deeper than 65534 levels (the u16 level store) is
 * rerun source by source on the single-source path, so no depth is refused. */
int eg_bfs_batch(eg_graph *g, uint32_t k, const uint32_t *roots,
                 uint32_t *dist, double *device_ms);
typedef struct {
  uint32_t traversals;  /* multi-source traversals (groups of sources) */
  uint32_t lanes;       /* sources per traversal */
  uint32_t levels;      /* levels run, summed over the traversals */
  uint32_t push_levels; /* of which pushed */
  uint32_t host_syncs;  /* host waits on the device while levels ran */
  uint32_t fallbacks;   /* traversals rerun source by source (depth past 65534) */
  double device_ms;
} eg_bfs_batch_info;
/* Figures of the last eg_bfs_batch call on g. */
int eg_bfs_batch_info_get(eg_graph *g, eg_bfs_batch_info *out);
/* k personalized PageRank vectors as one FP32 SpMM over the in-edges.
 * Query q's seeds are seeds[seed_ptr[q] .. seed_ptr[q+1]); p = uniform over
 * its distinct seeds, x' = alpha p + (1-alpha) A^T D^-1 x.  out (may be NULL)
 * is k x n, query-major. */
int eg_ppr_batch(eg_graph *g, uint32_t k, const uint32_t *seed_ptr,
                 const uint32_t *seeds, double alpha, int32_t iterations,
                 float *out, double *device_ms);
/* One personalized PageRank vector on the single-vector PageRank kernels
 * (the plain per-query path; binned or pull as "pr_binned" selects): x0 = p, iterated as eg_ppr_batch until the L1 change
 * of an iteration is <= tol or max_iter iterations ran.  out (may be NULL):
 * n floats.  iterations / delta (may be NULL): iterations run, last L1 change. */
int eg_ppr(eg_graph *g, const uint32_t *seeds, uint32_t nseeds, double alpha, double tol,
           int32_t max_iter, float *out, int32_t *iterations, double *delta);

/* ---- kernel-shape autotuning (no reference counterpart) ----
 * Tunable shapes: "pull_batch" (pull PageRank / SpMV / PPR gather depth),
 * "spmm_depth" (batched PPR row gathers in flight), "ms_batch" (multi-source
 * BFS index pass depth), each 8 or 16; "ms_lanes" (sources per multi-source
 * traversal: 64, 128 or 256, default 256); "ms_device" (1: device-driven
 * batched levels, the default; 0: the host decides and waits every level);
 * "ms_push_div" (1..1000: a multi-source level on a symmetric graph pushes
 * while its union frontier's vertices plus out-degrees are at most
 * m / ms_push_div; 20, the default, is bfs_simple.c's rule; levels are the
 * same for every value);
 * "pull_hot" (sources cached normally in the degree-ordered pull, a vertex
 * count, 0 = all); "pull_persist" (0, or 1-2
 * persistent pull workgroups per CU keeping the highest-degree sources' values
 * in LDS); "pr_binned" (PageRank and single-vector PPR on adjacency lists:
 * 1 = binned two-phase
 * kernels whose row sums are exact in 64-bit fixed point, 0 = the pull kernel
 * with compensated FP32 sums, -1 = binned from 2^22 stored entries up; the two
 * forms round differently, each is deterministic); "pr_tile" (source offsets
 * per binned scatter workgroup, 0 = 32768); "pr_scatter" (binned scatter
 * source offsets in flight per lane, 8, 16 or 32); "pr_hot" (binned PageRank:
 * sources below this id, at most 389120, are read from an LDS copy in phase 2
 * instead of being scattered, one window of 24320 sources at a time; the plan
 * keeps the leading windows that hold at least 2% of the entries each, and
 * none when they hold under 15% together; 0 = none, default 389120);
 * "pr_persist" (1: phase-2 workgroups loop over work items, one per CU, the
 * default; 0: one workgroup per work item); "pr_xcd" (binned
 * scatter: 0 = each of the eight XCDs walks a contiguous eighth of the
 * chunk-ordered tiles, the default; 1 = each XCD takes short runs from every
 * eighth by turns); "bfs_order"
 * (symmetric graphs: 2 = BFS runs on the degree-ordered vertex ids with every
 * row's neighbours by descending degree, results written in the caller's ids
 * (the default; 1 when the ordered copy is disabled); 1 = the layout's ids,
 * rows by descending neighbour degree; 0 = rows as built; the copies are built
 * on first use; levels are the same, parents may differ); "wcc_order" (undirected
 * / symmetric graphs: 1 = WCC rounds pull over the degree-ordered ids, rows
 * hubs first, the default; 0 = edge relaxation in the layout's ids; labels are
 * the same); "grid_band" (grid source
 * partitions per launch, 0 = all); "graphs" (1: batched PPR iterations
 * replayed from a hipGraph); "ingest_kib" (record ingest chunk of the pinned
 * two-slot load pipeline, 0 = 64 MiB; set before loading); "wide_part"
 * (adjacency modes, set before loading: 10..31 builds the layout with 64-bit
 * row offsets from stably sorted parts of 2^wide_part records, as every load
 * past 2^32 stored entries is built in parts of 2^31; 0 = the default).
 * Every variant
 * except pr_binned computes bit-identical results.  eg_autotune times the
 * variants of the selected kernels on the graph itself, keeps the fastest, and
 * reports each timing. */
#define EG_TUNE_PULL 1u
#define EG_TUNE_SPMM 2u
#define EG_TUNE_MSBFS 4u
#define EG_TUNE_GRID 8u /* grid PageRank / SpMV: "grid_band" source partitions per launch */
typedef struct {
  char param[24];
  int32_t value;
  int32_t chosen; /* 1 for the variant kept */
  double ms;      /* device ms per iteration (pull, SpMM) or per multi-source call */
} eg_tune_entry;
int eg_autotune(eg_graph *g, uint32_t kernels, eg_tune_entry *out, int32_t cap,
                int32_t *count);
/* Geometry of the binned plan of g (built by the first eg_pagerank, eg_ppr
 * or eg_spmv call that uses the binned kernels; no reference counterpart):
 * out[0..6] =
 * destination bins, source chunks, scatter tiles, accumulation work items,
 * rows summed in parts, non-empty (chunk, bin) segments, hot sources (read
 * from LDS in phase 2).  EG_ESTATE if no plan exists. */
int eg_pr_binned_geometry(eg_graph *g, uint32_t *out);
int eg_tune_set(eg_graph *g, const char *param, int32_t value);
int eg_tune_get(eg_graph *g, const char *param, int32_t *value);

/* ---- serving form (no reference counterpart) ----
 * A server owns a loaded graph (which must not be used by other calls until
 * eg_serve_stop) and runs one engine thread that batches continuously: BFS
 * queries join a multi-source traversal at the next level boundary (64 lanes
 * per lane group, each lane with its own level count) and PPR queries join
 * the 64-column SpMM at the next iteration boundary, each column retiring when
 * its own L1 change falls to ppr_tol.  The two kinds run on separate streams
 * and overlap on the device.  Lane groups and the PPR block are allocated
 * only while free device memory stays above hbm_reserve; a query waits for a
 * free slot, and fails with EG_ENOMEM if no slot of its kind can ever be
 * allocated.  A query's result does not depend on what else is in flight:
 * BFS levels are exact, and a PPR vector is bit-identical whichever queries
 * share its block (with idle_plain = -1; by default a query that arrives at
 * light load runs on the single-vector kernel, which sums in another order and
 * agrees to FP32 rounding). */
typedef struct eg_server eg_server;

#define EG_REQ_BFS 0
#define EG_REQ_PPR 1

typedef struct {
  uint32_t bfs_lanes;   /* BFS queries in flight, rounded up to 64; 0 = 64 */
  uint32_t ppr_columns; /* PPR queries in flight, 1..64; 0 = 64 */
  uint64_t hbm_reserve; /* device bytes left free; 0 = 1 GiB */
  double ppr_alpha;     /* teleport probability; 0 = 0.15 */
  double ppr_tol;       /* L1 change that ends a PPR query; 0 = 1e-6 */
  int32_t ppr_max_iter; /* iteration cap; 0 = 200 */
  int32_t plain;        /* 1: one query at a time, BFS through eg_bfs */
  uint32_t max_pending; /* queued queries before submit says EG_EAGAIN; 0 = 1<<20 */
  uint32_t ppr_grid;    /* workgroups of an SpMM iteration, leaving the rest of the
                           device to BFS levels; 0 = 8 per compute unit */
  int32_t idle_plain;   /* light load: with no BFS lane busy and at most this many
                           BFS queries waiting, the oldest runs on the single-source
                           path (eg_bfs); with no PPR column busy and at most
                           L PPR queries waiting, the oldest runs on the
                           single-vector kernel (eg_ppr), stepped beside BFS work.
                           L is the larger of twice this value and the measured
                           cost of an SpMM step over that many columns in
                           single-vector steps (4 + columns/8 until both are
                           timed), capped at 63.  0 = 2, -1 = always batch
                           (every PPR result then bit-identical at any load) */
  uint32_t pad;
} eg_serve_opts;

typedef struct {
  int32_t kind;      /* EG_REQ_BFS or EG_REQ_PPR */
  int32_t status;    /* EG_OK or this query's error */
  uint32_t levels;   /* BFS: largest dist value; PPR: iterations run */
  uint32_t pad;
  uint64_t reached;  /* BFS: vertices reached; PPR: 0 */
  double delta;      /* PPR: L1 change of the last iteration */
  double queue_ms;   /* submit -> first device step */
  double latency_ms; /* submit -> result written */
} eg_serve_result;

typedef struct {
  uint64_t submitted, completed, failed;
  uint32_t pending, bfs_inflight, ppr_inflight;
  uint32_t bfs_lanes, ppr_columns; /* slots allocated so far */
  uint32_t pad;
  uint64_t bfs_levels, ppr_steps;  /* device steps run */
  uint64_t device_bytes;           /* device memory held by the server */
  uint64_t served_plain;           /* queries served through the plain path */
} eg_serve_stats;

/* Builds every layout the engine's kernels read (pull plan or ordered copy,
 * the binned plan of the single-vector path, the hubs-first BFS columns of a
 * symmetric graph) before its thread starts, so no query waits for a build. */
int eg_serve_start(eg_graph *g, const eg_serve_opts *opts, eg_server **out);
/* out (may be NULL): host buffer of n entries, written before the query
 * completes: uint32 levels in the eg_bfs encoding. */
int eg_serve_submit_bfs(eg_server *s, uint32_t root, uint32_t *out, uint64_t *ticket);
/* Personalized PageRank over the distinct seeds (as eg_ppr_batch); out (may
 * be NULL): n floats. */
int eg_serve_submit_ppr(eg_server *s, const uint32_t *seeds, uint32_t nseeds,
                        float *out, uint64_t *ticket);
/* Waits for a ticket (timeout_ms < 0: no limit) and releases it.  Returns
 * EG_ETIMEDOUT while it is still running, else EG_OK with the query's own
 * status in res->status.  One thread at a time may wait on a ticket: a second
 * concurrent wait on it returns EG_EINVAL; after EG_OK the ticket is unknown. */
int eg_serve_wait(eg_server *s, uint64_t ticket, double timeout_ms, eg_serve_result *res);
int eg_serve_stats_get(eg_server *s, eg_serve_stats *st);
/* Completes every submitted query, stops the engine and frees the server. */
int eg_serve_stop(eg_server *s);

/* Poisson request stream: exponential inter-arrival times at rate qps,
 * EG_REQ_BFS with probability bfs_fraction (root uniform over vertices with
 * out-edges), else PPR with 1..ppr_max_seeds seeds drawn the same way; all
 * draws from `seed`.  Results are summaries only. */
typedef struct {
  double qps;
  uint32_t requests;
  uint32_t ppr_max_seeds; /* 0 = 8 */
  double bfs_fraction;
  uint64_t seed;
} eg_replay_params;

typedef struct {
  uint32_t completed, failed, bfs, ppr;
  double wall_s, qps; /* first submit -> last completion; completed / wall_s */
  double p50_ms, p99_ms, mean_ms, max_ms;
  double bfs_p50_ms, bfs_p99_ms, ppr_p50_ms, ppr_p99_ms;
  eg_serve_stats stats;
} eg_replay_report;

int eg_serve_replay(eg_graph *g, const eg_serve_opts *opts, const eg_replay_params *p,
                    eg_replay_report *r);

/* ---- device timing of named kernels on the graph's stream ---- */
typedef struct {
  char name[32];
  uint64_t launches;
  double total_ms;
} eg_kernel_time;
int eg_prof_enable(eg_graph *g, int32_t on);
int eg_prof_reset(eg_graph *g);
int eg_prof_read(eg_graph *g, eg_kernel_time *out, int32_t max, int32_t *count);
int eg_sync(eg_graph *g);

#ifdef __cplusplus
}
#endif
#endif
