// eg_batch.h -- multi-source state and step functions shared by the batched
// API (eg_bfs_batch, eg_ppr_batch) and the serving engine.
#pragma once

#include "eg_internal.h"

// Pull orientation of the batched kernels: the degree-ordered copy when
// enabled (perm = new id of every vertex, deg = out-degree in new ids), else
// the in-edges as built.  sym: the orientation also lists the out-edges
// (symmetric graph), which allows push levels.
struct BatchView {
  const uint32_t *ptr, *idx, *deg, *perm;
  // MS-BFS rows: idx with every row's entries ascending (hubs first) on
  // symmetric graphs, where a pull row stops once it holds every needed lane;
  // otherwise idx.  OR is order-free, so levels do not depend on it.
  const uint32_t *ms_idx;
  PullPlan *pl;
  bool sym;
  // Chunks [static_chunk0, nchunks) hold only rows without entries and
  // without out-edges: their PPR values are the seed term (or 0) and their
  // contributions 0 from the first step on, so once two full steps have run
  // since the last column change both iterate buffers hold them and the steps
  // skip those chunks.
  uint32_t static_chunk0;
};
int eg_batch_view(eg_graph *g, BatchView *v);

// Up to 64 nw BFS lanes over one traversal (nw = 1, 2 or 4 lane words per
// vertex): lane i of vertex v is bit i % 64 of word v * nw + i / 64 of the
// masks.  Levels are u16, lane-major (lvl[lane * n + v]), root = 1; a level
// holds only where the lane's seen bit is set (elsewhere it may be left from
// an earlier traversal: nothing clears the store).
struct MsState {
  uint32_t nw = 1;
  hipStream_t stream = nullptr;  // every step is enqueued here
  void *scan_tmp = nullptr;      // frontier degree scan workspace
  uint64_t *seen = nullptr, *visit = nullptr, *nvisit = nullptr, *pend = nullptr;
  uint16_t *lvl = nullptr;
  uint64_t *seg_or = nullptr;  // [nseg] long-row partial ORs
  uint32_t seg_cap = 0;
  uint32_t *queue = nullptr, *qoff = nullptr;  // frontier vertices with edges, degree scan
  uint64_t *fbits = nullptr;  // device-driven levels: bit v = v is in the frontier and has edges
  unsigned long long *ctr = nullptr;  // [0] progress, [1] count, [2] degsum, [3] qlen, [8..) partials
  uint16_t *lvl_tab = nullptr;                 // [64 nw] level written per lane this step
  uint32_t *roots = nullptr;                   // [512]: lanes, roots (staging)
  uint32_t *ctl = nullptr;                     // device-driven levels: control words
  hipGraphExec_t exec = nullptr;               // two device-driven levels, replayed
  uint64_t exec_key[6] = {};                   // what the recorded levels were launched with
  uint32_t depth_hint = 0;                     // levels of the last device-driven traversal
  bool pushed = false;                         // the last host-driven level pushed
  unsigned long long *h = nullptr;             // pinned: [0..5) ctr mirror, [8..8+64nw) lane sums
  uint32_t qlen = 0;                           // frontier queue length (host copy)
  uint64_t qdeg = 0;                           // its out-degree sum
};
int eg_ms_alloc(eg_graph *g, MsState *st, uint32_t nw);
void eg_ms_free(MsState *st);
// Clears every lane.
int eg_ms_reset(eg_graph *g, MsState *st);
// Starts lane[i] at roots[i] (vertex ids of the layout), i < k.
int eg_ms_seed(eg_graph *g, const BatchView &bv, MsState *st, uint32_t k, const uint32_t *lanes,
               const uint32_t *roots);
// One level for the lanes in `active` (nw words); tab[i] is the level value
// written for lane i (64 nw entries).  progress (nw words) = lanes that
// reached a new vertex.
int eg_ms_level(eg_graph *g, const BatchView &bv, MsState *st, const uint64_t *active,
                const uint16_t *tab, uint64_t *progress);
// out[v] (device, n words) = level of lane for vertex v of the layout.
int eg_ms_extract_lane(eg_graph *g, const BatchView &bv, MsState *st, uint32_t lane,
                       uint32_t *out);
// Frees lanes (a mask) for reuse; each must have finished (no frontier).
int eg_ms_clear_lanes(eg_graph *g, MsState *st, const uint64_t *lanes);
// reached[i] = vertices with a level in lane lanes[i], i < k <= 64.
int eg_ms_reached(eg_graph *g, MsState *st, uint32_t k, const uint32_t *lanes,
                  uint64_t *reached);

// Up to 64 personalized PageRank columns (FP32, row-major, 64 per vertex).
struct PprState {
  hipStream_t stream = nullptr;
  float *X = nullptr, *Y = nullptr, *C = nullptr, *Cn = nullptr;
  uint64_t *mask = nullptr;  // seed columns of every vertex
  float *w = nullptr;        // [128]: 1/|S_q| and alpha/|S_q| per column
  float2 *seg_part = nullptr;
  uint64_t seg_cap = 0;
  uint32_t *seeds = nullptr;  // staging for column initialisation
  uint32_t seeds_cap = 0;
  float hw[128];              // host copy of w
  double *dpart = nullptr;    // per-workgroup L1 delta partials [blocks][64]
  uint64_t dpart_cap = 0;
  double *delta = nullptr;    // [64] L1 change of every column in the last step
  double *h_delta = nullptr;  // pinned mirror
  uint64_t active = ~0ull;    // columns in use: steps gather only up to the highest one
  uint32_t max_grid = 0;      // workgroups of the SpMM launch (0: one per work item)
  bool capturing = false;     // a step is being recorded into a hipGraph
  uint32_t full_steps = 0;    // steps over every chunk since the last column change
  bool skip_y = false;        // non-delta steps may leave Y unwritten (a later step writes it)
};
int eg_ppr_alloc(eg_graph *g, PprState *st);
void eg_ppr_free(PprState *st);
int eg_ppr_reset(eg_graph *g, PprState *st);
// Column q starts personalized at seeds[0..ns) (vertex ids of the layout).
int eg_ppr_col_init(eg_graph *g, const BatchView &bv, PprState *st, uint32_t q,
                    const uint32_t *seeds, uint32_t ns, double alpha);
// One iteration x' = alpha p + (1-alpha) A^T D^-1 x for all 64 columns,
// enqueued on st->stream.  With want_delta, the L1 change of every column is
// reduced in a fixed order and copied to st->h_delta (valid once the stream
// reaches the end of the step).
int eg_ppr_step(eg_graph *g, const BatchView &bv, PprState *st, double alpha,
                bool want_delta = false);
int eg_ppr_extract_col(eg_graph *g, const BatchView &bv, PprState *st, uint32_t q, float *out);
// Zeroes columns (a mask) for reuse.
int eg_ppr_clear_cols(eg_graph *g, PprState *st, uint64_t cols);
