// eg_internal.h -- shared declarations of libeg (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/eg.h"

// ---------------------------------------------------------------------------
// Errors: a thread-local message plus a negative status code.
// ---------------------------------------------------------------------------
int eg_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define EG_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return eg_fail(EG_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call,    \
                     hipGetErrorString(e_));                                   \
  } while (0)

#define EG_TRY(call)                                                           \
  do {                                                                         \
    int rc_ = (call);                                                          \
    if (rc_ != EG_OK) return rc_;                                              \
  } while (0)

// Launch check: reports configuration errors of the launch just issued.
#define EG_LAUNCHED() EG_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Device memory helpers.
// ---------------------------------------------------------------------------
int eg_dalloc(void **p, size_t bytes, const char *what);
void eg_dfree(void **p);
template <typename T>
static inline int eg_dalloc_t(T **p, size_t count, const char *what) {
  return eg_dalloc(reinterpret_cast<void **>(p), count * sizeof(T), what);
}
template <typename T>
static inline void eg_dfree_t(T **p) {
  eg_dfree(reinterpret_cast<void **>(p));
}

// ---------------------------------------------------------------------------
// Kernel timing: hipEvents recorded around named launches on the graph's
// stream, resolved when read.
// ---------------------------------------------------------------------------
struct EgProfSlot {
  std::string name;
  std::vector<hipEvent_t> pending;  // start, stop pairs
  uint64_t launches = 0;
  double total_ms = 0.0;
};

// Work decomposition for a pull kernel over one CSR-like orientation
// (ptr[n+1], idx[m]).  Short rows are grouped into chunks of the merged
// (row, entry) sequence; rows with more than kLongRow entries are split into
// segments of kSegLen entries whose compensated partial sums are combined in
// segment order.
constexpr uint32_t kPlanChunk = 2048;    // merged (row + entry) units per chunk
constexpr uint32_t kPlanLongRow = 2048;  // rows longer than this are segmented
constexpr uint32_t kPlanSegLen = 4096;   // entries per segment of a long row

struct PullPlan {
  bool built = false;
  uint32_t nchunks = 0;     // stream chunks
  uint32_t nlong = 0;       // long rows
  uint32_t nseg = 0;        // segments over all long rows
  uint4 *chunk_info = nullptr;    // [nchunks] short rows [x, y), entries [z, w)
  uint4 *seg_info = nullptr;      // [nseg] entries [x, y) of long row index z
  uint32_t *long_rows = nullptr;  // [nlong]
  uint32_t *seg_base = nullptr;   // [nlong + 1]
  float2 *seg_part = nullptr;     // [nseg] (sum, compensation)
  double2 *chunk_part = nullptr;  // [nchunks] FP64 (dangling, l1)
  double2 *fin_part = nullptr;    // [kFinBlocks]
  uint32_t *fin_counter = nullptr;
  // First chunk whose rows all have no entries and no out-edges (such rows'
  // batched PPR values stop changing); UINT32_MAX until computed.
  uint32_t static_chunk0 = 0xFFFFFFFFu;
};

struct GridCompute;
struct BatchState;
struct PbPlan;

// Degree-ordered copy of the pull orientation (eg_order.hip): vertex i of the
// copy is old vertex old_of_new[i]; rows keep their entry order.
struct EgOrdered {
  bool built = false;
  uint32_t *old_of_new = nullptr, *new_of_old = nullptr;
  uint32_t *ptr = nullptr, *idx = nullptr;
  float *w = nullptr;
  uint32_t *deg = nullptr;  // out-degree in new ids
  PullPlan plan;
};

// Binned PageRank: sources per hot window (phase 2 keeps one window's
// contributions in LDS beside its accumulators) and windows at most.
constexpr int kPbHotWindow = 24320;
constexpr int kPbHotWindows = 16;

// Kernel shapes of a graph, set by eg_autotune or eg_tune_set.  Every
// variant computes bit-identical results; only the speed differs.
struct EgTune {
  int pull_batch = 16;   // k_pull: entries per lane per gather pass (8 or 16)
  int spmm_depth = 8;    // k_ppr_spmm: row gathers in flight per wave (8 or 16)
  int ms_batch = 8;      // k_ms_pull: entries per lane per gather pass (8 or 16)
  int graphs = 1;        // batched PPR iterations replayed from a hipGraph (0: eager)
  int grid_band = 1;     // grid pull: source partitions per launch (0: all)
  int ingest_kib = 0;    // record ingest chunk in KiB (0: 64 MiB)
  int pull_persist = 0;  // k_pull on the degree-ordered copy: 0 = a workgroup per work
                         // item; 1 or 2 = persistent workgroups per CU holding the
                         // highest-degree sources' values in LDS
  int pull_hot = 0;      // k_pull on the degree-ordered copy: sources below this id
                         // are cached normally, the rest loaded non-temporally (0: all normal)
  int pr_binned = -1;    // PageRank on adjacency lists: 1 = binned two-phase kernels
                         // (exact fixed-point row sums), 0 = pull kernel (compensated
                         // FP32 sums), -1 = binned when the graph has >= 2^22 entries
  int pr_tile = 0;       // binned PageRank: source offsets per scatter tile (0: 32768)
  int pr_scatter = 16;   // binned scatter: source offsets in flight per lane (8, 16 or 32)
  int pr_hot = kPbHotWindow * kPbHotWindows;  // binned PageRank: sources below this id are
                         // read from an LDS copy in phase 2 instead of being scattered (0:
                         // none; in windows of kPbHotWindow sources; the plan keeps the
                         // windows that pay)
  int wcc_order = 1;     // undirected / symmetric graphs: WCC pulls on the ordered ids (0: edges)
  int bfs_order = 2;     // symmetric graphs: BFS on CSR rows ordered hubs first (1), on the
                         // degree-ordered ids with rows hubs first (2), or as built (0)
  int pr_persist = 1;    // binned PageRank with hot sources: phase-2 workgroups loop over
                         // items (one per CU) instead of one workgroup per item
  int pr_xcd = 0;        // binned scatter: every XCD walks a contiguous eighth of the
                         // chunk-ordered tiles (0) or runs from every eighth by turns (1)
  int ms_lanes = 256;    // eg_bfs_batch: sources per multi-source traversal (64, 128, 256)
  int ms_device = 1;     // eg_bfs_batch: device-driven levels (1) or host-driven (0)
  int ms_push_div = 20;  // multi-source BFS on symmetric graphs: a level pushes while the
                         // union frontier's vertices plus out-degrees are at most
                         // m / ms_push_div (20: the single-source rule of bfs_simple.c:191)
  int wide_part = 0;     // adjacency loads: build with 64-bit row offsets from sorted parts
                         // of 2^wide_part entries (10..31); 0: only past 2^32 entries,
                         // in parts of 2^31
};

struct eg_graph {
  uint32_t n = 0, P = 0, flags = 0;
  int32_t load_mode = 2;
  uint64_t m_in = 0, m = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  bool loaded = false;
  double build_ms = 0.0;

  // Records in input order (memblock), reversed copies appended with -u.
  uint32_t *e_src = nullptr, *e_dst = nullptr;
  float *e_w = nullptr;
  // CSR (out-edges) and the sorted src of every entry.
  uint32_t *out_ptr = nullptr, *out_col = nullptr, *out_src = nullptr;
  float *out_w = nullptr;
  // CSC (in-edges).
  uint32_t *in_ptr = nullptr, *in_idx = nullptr;
  // Adjacency with 64-bit row offsets (wide): out_ptr64 / in_ptr64 replace
  // out_ptr / in_ptr (both null), out_src is not kept.
  bool wide = false;
  uint64_t *out_ptr64 = nullptr, *in_ptr64 = nullptr;
  float *in_w = nullptr;
  uint32_t *out_deg = nullptr, *in_deg = nullptr;
  // 2-D grid.
  uint32_t *cell_ptr = nullptr, *g_src = nullptr, *g_dst = nullptr;
  float *g_w = nullptr;
  // Edge array of modes 7/8/9.
  uint32_t *el_src = nullptr, *el_dst = nullptr;
  float *el_w = nullptr;

  PullPlan plan_in;   // pull over in-edges
  PullPlan plan_out;  // pull over out-edges (symmetric graphs, BFS)

  // Algorithm state (allocated on first use, kept for reuse).
  float *rank_a = nullptr, *rank_b = nullptr, *contrib_a = nullptr,
        *contrib_b = nullptr;
  double *pr_stats = nullptr;  // [2 * max_iters]
  int pr_stats_cap = 0;
  uint32_t *bfs_dist = nullptr, *bfs_parent = nullptr;
  uint64_t *bfs_visited = nullptr, *bfs_front = nullptr, *bfs_next = nullptr;
  uint32_t *bfs_q = nullptr, *bfs_nq = nullptr;
  uint64_t *bfs_f3 = nullptr;                    // third frontier bitmap (device-driven BFS)
  uint32_t *bfs_col = nullptr;  // symmetric graphs: CSR columns, each row by descending degree
  uint32_t *bfs_ocol = nullptr; // symmetric graphs: degree-ordered rows, entries ascending (new ids)
  uint32_t bfs_onz = 0;         // vertices with edges (a prefix of the ordered ids)
  uint32_t *bfs_qoff = nullptr, *bfs_nqoff = nullptr;  // queue edge offsets
  uint32_t bfs_depth_hint = 0;  // levels of the last traversal (first level batch)
  unsigned long long *bfs_part = nullptr;  // per-workgroup partial counts
  uint32_t *bfs_bmcnt = nullptr;           // bitmap -> queue block counts
  uint32_t ndangling = 0xFFFFFFFFu;        // vertices without out-edges (cached)
  uint32_t *bfs_ctr = nullptr;      // device counters
  uint32_t *h_ctr = nullptr;        // pinned mirror
  int64_t *wcc_label = nullptr;
  uint32_t *wcc_lab32 = nullptr;
  uint32_t *wcc_tmp = nullptr;  // ordered-id labels and component minima (2n)
  uint32_t *wide_seg = nullptr;            // wide BFS: row segments per frontier vertex
  unsigned long long *wide_ctr = nullptr;  // wide BFS: next-queue length, degree sum

  // Scratch buffers shared by the builders (grown on demand).
  void *scratch = nullptr;
  size_t scratch_bytes = 0;

  GridCompute *gridc = nullptr;  // compute copy of the grid (eg_grid.hip)
  BatchState *batch = nullptr;   // multi-source state (eg_batch.hip)
  EgOrdered *ord = nullptr;      // degree-ordered pull copy (eg_order.hip)
  PbPlan *pb = nullptr;          // binned PageRank plan (eg_binned.hip)

  int prof_on = 0;
  std::vector<EgProfSlot> prof;
  eg_server *server = nullptr;  // set while a server owns the graph
  EgTune tune;
  uint32_t *sssp_segmin = nullptr;  // SSSP pull: minimum per long-row segment
  uint32_t sssp_segmin_cap = 0;
  // One-time preparations built on first use (eg_graph_prep_info): wall ms
  // and device bytes of the degree-ordered copy, the binned plan and the
  // hubs-first BFS rows.
  double prep_order_ms = 0.0, prep_plan_ms = 0.0, prep_bfs_ms = 0.0;
  uint64_t prep_order_bytes = 0, prep_plan_bytes = 0, prep_bfs_bytes = 0;
  float *ppr_base = nullptr;     // single-vector PPR: alpha p
  uint32_t *ppr_seeds = nullptr;
  uint32_t ppr_seeds_cap = 0;
};

void eg_grid_free(eg_graph *g);
void eg_batch_free(eg_graph *g);
void eg_order_free(eg_graph *g);
void eg_pb_free(eg_graph *g);
// Binned PageRank / SpMV (eg_binned.hip).  The entries of row v are the
// sources u of its in-edges, from an adjacency orientation (ptr[n+1], idx[m])
// or an edge list (row[k], idx[k]; ptr null); w (may be null): entry weights,
// kept for SpMV.  The plan is built on first use for that source; the
// fixed-point scale for row sums <= bound; an SpMV pass y = A^T x over the
// plan (x, y on the device, xmax = max |x|).
struct PbSource {
  const uint32_t *ptr = nullptr, *row = nullptr, *idx = nullptr;
  const float *w = nullptr;
};
int eg_pb_build(eg_graph *g, const PbSource &src);
int eg_pb_scale(double bound);
int eg_pb_spmv(eg_graph *g, const float *x, float *y, double xmax, hipStream_t stream = nullptr);
// True when eg_pb_spmv can sum exactly over the built plan: finite weights
// and |w| |x| below FLT_MAX (else the pull kernels take the product).
bool eg_pb_spmv_exact(const eg_graph *g, double xmax);
// Automatic choices weigh the one-time preparations against the work asked
// for (measured at R-MAT 24 -u: the degree-ordered copy costs about 8
// pull iterations and saves about a quarter of each; the binned plan costs
// about 80 and saves about 60% of each): an orientation's degree-ordered copy
// is built for at least kAutoOrderIters iterations, a binned plan (pr_binned
// = -1) for at least kAutoBinnedIters; once built, either is used.
constexpr int kAutoOrderIters = 32;
constexpr int kAutoBinnedIters = 128;
// True when PageRank / SpMV / PPR over the entry source `key` (an
// orientation's row pointers, or an edge list's rows) uses the binned kernels
// for a run of `iterations` iterations.
bool eg_pb_selected(const eg_graph *g, int iterations, const void *key);
bool eg_reorder_enabled();
// True when a run of `iterations` pull / binned iterations uses the
// degree-ordered copy (built already, or worth building).
bool eg_use_order(const eg_graph *g, int iterations);
// Builds (on first use) and returns the degree-ordered pull copy.
int eg_ordered(eg_graph *g, EgOrdered **out);
// dst[i] = src[perm[i]] for 32-bit vectors of length n (on the graph stream).
int eg_gather_vec(eg_graph *g, const void *src, const uint32_t *perm, void *dst,
                  hipStream_t stream = nullptr);

bool eg_has_in_edges(const eg_graph *g);
int eg_scratch(eg_graph *g, size_t bytes, void **out);
// EG_OK if g has a layout with 32-bit offsets and no server owns it.
int eg_require_loaded(const eg_graph *g);
// The same for the entry points that also run on 64-bit row offsets (BFS,
// degrees, the CSR / CSC downloads).
int eg_require_loaded_wide_ok(const eg_graph *g);
// Direction-optimizing BFS over 64-bit row offsets (eg_wide.hip): dist /
// parent of eg_bfs in g->bfs_dist / g->bfs_parent, levels in *st.
int eg_wide_bfs(eg_graph *g, uint32_t root, int start_pull, int switch_mode, bool want_parent,
                eg_bfs_stats *st);
// Single-vector personalized PageRank on the PageRank pull kernel (the plain
// per-query path): stops once the L1 change is <= tol or after max_iter
// iterations.  out (host, may be null) receives the n ranks.
int eg_ppr_single(eg_graph *g, const uint32_t *seeds, uint32_t ns, double alpha, double tol,
                  int max_iter, float *out, int *iters, double *delta);

// The same query stepped on an explicit stream (the serving engine overlaps
// it with other work).  It uses the graph's PageRank vectors, so one runs at a
// time.  eg_ppr1_step enqueues one iteration and the copy of its L1 change to
// *h_delta (valid once the stream has passed it); eg_ppr1_output enqueues the
// result's download to host memory and waits for it.
struct Ppr1 {
  const uint32_t *ptr = nullptr, *idx = nullptr, *deg = nullptr, *unperm = nullptr;
  const float *w = nullptr;  // the orientation's weights (the binned plan keeps them)
  PullPlan *pl = nullptr;
  float *rank = nullptr, *rank_next = nullptr, *contrib = nullptr, *contrib_next = nullptr;
  double alpha = 0.15;
  int it = 0, max_iter = 0;
  hipStream_t stream = nullptr;
  double *h_delta = nullptr;  // pinned, owned by the caller
  bool binned = false;        // iterations on the binned kernels
  int scale = 0;              // their fixed-point scale
};
int eg_ppr1_begin(eg_graph *g, Ppr1 *st, const uint32_t *seeds, uint32_t ns, double alpha,
                  int max_iter, hipStream_t stream, double *h_delta);
int eg_ppr1_step(eg_graph *g, Ppr1 *st);
// Builds what eg_ppr1_begin would build on first use (orientation, vectors,
// binned plan) without starting a query.
int eg_ppr1_prepare(eg_graph *g, int max_iter);
// Builds the adjacency copies a BFS would build on first use.
int eg_bfs_prepare(eg_graph *g);
// Undirected / symmetric graphs: the rows of the degree-ordered copy with
// every row's entries ascending (hubs first), built on first use; *rows =
// null where the ordered copy is disabled.
int eg_hubs_first_rows(eg_graph *g, const uint32_t **rows);
// Takes back the last eg_ppr1_step (queued or finished): a step reads the
// iterate in one buffer and writes the other, so the iterate before it is
// intact; work queued after this call sees the earlier iterate.
int eg_ppr1_undo(eg_graph *g, Ppr1 *st);
int eg_ppr1_output(eg_graph *g, Ppr1 *st, float *out);
// eg_bfs without the ownership check (the serving engine's plain path).
int eg_bfs_host(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
                uint32_t *dist, uint32_t *parent, eg_bfs_stats *stats);

// Wall-clock milliseconds since an earlier eg_wall_ms() (a steady clock).
double eg_wall_ms();

// Profiling scope around one launch.
struct EgProf {
  eg_graph *g;
  EgProfSlot *slot = nullptr;
  hipEvent_t ev[2];
  hipStream_t stream;
  // Times the work enqueued on stream s (the graph's stream when null)
  // between construction and destruction.
  EgProf(eg_graph *g_, const char *name, hipStream_t s = nullptr);
  ~EgProf();
};

// ---------------------------------------------------------------------------
// Primitives (eg_prims.hip).
// ---------------------------------------------------------------------------
// out[0..count) = exclusive scan of in, out[count] = total.  in and out may
// alias.  tmp must hold eg_scan_tmp_bytes(count).
size_t eg_scan_tmp_bytes(uint64_t count);
int eg_scan_excl(hipStream_t s, const uint32_t *in, uint32_t *out,
                 uint64_t count, void *tmp);
// a[0..count) = exclusive scan of a (in place), a[count] = total; tmp holds
// eg_scan64_tmp_bytes() bytes.
size_t eg_scan64_tmp_bytes();
int eg_scan_excl_u64(hipStream_t s, uint64_t *a, uint64_t count, void *tmp);
// deg[k] += number of keys equal to k (deg zeroed by the caller).
int eg_histogram(hipStream_t s, const uint32_t *keys, uint64_t count,
                 uint32_t *deg, uint32_t nbins);
// ptr[0..n] = offsets of the runs of ascending keys; deg (may be NULL) their
// lengths.
int eg_ptr_from_sorted(hipStream_t s, const uint32_t *keys, uint64_t m,
                       uint32_t n, uint32_t *ptr, uint32_t *deg);
// *result (host) = number of zero entries of a[0..n); d_tmp holds 1025 words.
int eg_count_zero(hipStream_t s, const uint32_t *a, uint32_t n,
                  uint32_t *d_tmp, uint32_t *result);
// Stable LSD radix sort of count keys < 2^key_bits with up to three 32-bit
// payload columns.  kout may be NULL when the sorted keys are not needed.
// Inputs are not modified.
struct EgSortCols {
  const uint32_t *in[3] = {nullptr, nullptr, nullptr};
  uint32_t *out[3] = {nullptr, nullptr, nullptr};
  int np = 0;
};
// Optional id check folded into the sort's first pass: flag (device, zeroed
// by the caller) is set when a key or vals[i] is >= limit (vals = payload
// column 0 is checked as the first scatter reads it).
struct EgIdCheck {
  uint32_t limit = 0;
  const uint32_t *vals = nullptr;
  uint32_t *flag = nullptr;
};
size_t eg_radix_tmp_bytes(uint64_t count, int np);
int eg_radix_sort(hipStream_t s, uint64_t count, int key_bits,
                  const uint32_t *kin, uint32_t *kout, const EgSortCols &cols,
                  void *tmp, const EgIdCheck *chk = nullptr);
int eg_bits_for(uint64_t nkeys);

// Device generator (eg_prims.hip).
int eg_gen_device(hipStream_t s, const eg_gen_params *p, uint32_t n,
                  uint32_t *src, uint32_t *dst, float *w);
// Unpack raw AoS records into SoA (+ reversed copies when undirected).
int eg_unpack_records(hipStream_t s, const void *rec, uint64_t m_in,
                      int weighted, uint32_t *src, uint32_t *dst, float *w);
int eg_append_reverse(hipStream_t s, uint64_t m_in, uint32_t *src,
                      uint32_t *dst, float *w);
// Returns 1 in *bad if any id >= n.
int eg_check_ids(hipStream_t s, const uint32_t *a, const uint32_t *b,
                 uint64_t count, uint32_t n, uint32_t *d_flag, int *bad);
int eg_cell_keys(hipStream_t s, const uint32_t *src, const uint32_t *dst,
                 uint64_t count, uint32_t n, uint32_t P, uint32_t *keys);

// Pull planning (eg_pull.hip).
int eg_plan_build(eg_graph *g, PullPlan *pl, const uint32_t *ptr);
void eg_plan_free(PullPlan *pl);
// In-edge orientation of the pull kernels (CSC, or the CSR of a symmetric
// graph) and its plan, built on first use.
int eg_pull_orientation(eg_graph *g, const uint32_t **ptr, const uint32_t **idx,
                        const float **w, PullPlan **pl);

// ---------------------------------------------------------------------------
// Device helpers.
// ---------------------------------------------------------------------------
#define EG_WAVE 64

__device__ __forceinline__ uint32_t eg_partition_id(uint32_t v, uint32_t n,
                                                     uint32_t P) {
  // get_partition_id, init_all.c:54-60.
  const uint32_t mod_val = n % P;
  const uint32_t psize = mod_val ? n / P + 1 : n / P;
  if (mod_val == 0) return v / psize;
  const uint64_t split = (uint64_t)mod_val * psize;
  return v < split ? v / psize : (uint32_t)((v - split) / (psize - 1)) + mod_val;
}

// Error-free transformation a + b = s + e (Knuth TwoSum).
__device__ __forceinline__ void eg_two_sum(float a, float b, float &s,
                                           float &e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated accumulator (Neumaier's variant of Kahan summation).
struct EgKahan {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    float t, e;
    eg_two_sum(s, x, t, e);
    s = t;
    c += e;
  }
  __device__ __forceinline__ void merge(float s2, float c2) {
    float t, e;
    eg_two_sum(s, s2, t, e);
    s = t;
    c += e + c2;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// Butterfly merge of compensated pairs across the 64 lanes of a wave.  Every
// lane ends with the same (s, c); the combination order is fixed.
__device__ __forceinline__ void eg_wave_kahan_allreduce(EgKahan &k) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const float s2 = __shfl_xor(k.s, o);
    const float c2 = __shfl_xor(k.c, o);
    // Combine in lane order so that both partners compute the same pair.
    const int lane = __lane_id();
    EgKahan lo, hi;
    if (lane & o) {
      lo.s = s2; lo.c = c2; hi.s = k.s; hi.c = k.c;
    } else {
      lo.s = k.s; lo.c = k.c; hi.s = s2; hi.c = c2;
    }
    lo.merge(hi.s, hi.c);
    k = lo;
  }
}

__device__ __forceinline__ double eg_wave_sum_f64(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double w = __shfl_xor(v, o);
    const int lane = __lane_id();
    v = (lane & o) ? (w + v) : (v + w);
  }
  return v;
}

__device__ __forceinline__ uint32_t eg_wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ uint64_t eg_lanemask_lt() {
  return (1ull << __lane_id()) - 1ull;
}
