// eg_bfs.hip -- direction-optimising BFS over bitmap frontiers.
//
// Reference: the bfs_simple iterator (bfs_simple.c:136-229) alternates a push
// step over a vertex queue (bfs_push, :121-134) and a pull step over all
// unvisited vertices that stops at the first in-neighbour in the frontier
// (bfs_pull, :104-120), switching with the rule of :191-197.  Grid layouts
// run bfsgrid (bfsgrid_cilk.c:48-91, rows of cells whose src partition holds a
// frontier vertex) and edge arrays run bfs_edgecentric (bfs_simple.c:58-103).
// Levels use the encoding of bfsgrid_cilk.c:58-60 (root 1, unreached 0);
// parents that of bfs_numa.c:79-83,257 (parent[root] = root, unreached n).
//
// GPU form:
//  * visited / frontier sets are 1-bit bitmaps (64 vertices per word, the
//    Bitmap of bitmap.h:16-50 with the word as a wave's ballot).
//  * push: the frontier queue's out-degrees are scanned and every workgroup
//    expands an equal slice of the frontier's edges, finding each edge's
//    owner by binary search in LDS (load balanced regardless of hubs).  A
//    vertex is claimed with one atomicOr on its visited word; the winner
//    writes dist / parent and appends to the next queue with one atomic per
//    wave (ballot + mbcnt).
//  * pull: one wave per bitmap word; each unvisited lane scans its in-edges
//    until it meets a frontier vertex.  The wave writes the next-frontier and
//    visited words whole (no atomics).
//  * The host reads two counters per level (frontier size, its out-degree
//    sum) and applies the reference's switch rule.
#include "eg_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kEdgesPerBlock = 2048;

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

// Counters (u32 words of g->bfs_ctr): [0] next frontier size, [1] next queue
// length, [2..3] next frontier out-degree sum (u64).  The queue holds only
// vertices with out-edges: a vertex without out-edges is never an
// in-neighbour, so it cannot discover anything in either direction.
struct Ctr {
  uint32_t *count;
  uint32_t *qlen;
  unsigned long long *degsum;
};

// Counts the wave's discoveries and appends those with out-edges to nq (one
// atomic per wave).  Must be called by all 64 lanes.
__device__ __forceinline__ void wave_emit(Ctr c, bool found, uint32_t v, uint32_t deg,
                                          uint32_t *__restrict__ nq) {
  const uint64_t b = __ballot(found);
  if (!b) return;
  const bool enq = found && deg > 0;
  const uint64_t qb = __ballot(enq);
  unsigned long long ds = found ? deg : 0u;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) ds += __shfl_xor(ds, o);
  uint32_t base = 0;
  if (__lane_id() == 0) {
    atomicAdd(c.count, (uint32_t)__popcll(b));
    atomicAdd(c.degsum, ds);
    if (qb) base = atomicAdd(c.qlen, (uint32_t)__popcll(qb));
  }
  base = __shfl(base, 0);
  if (enq && nq) nq[base + __popcll(qb & eg_lanemask_lt())] = v;
}

__global__ void k_bfs_init(uint32_t n, uint32_t root, uint32_t *__restrict__ dist,
                           uint32_t *__restrict__ parent, uint64_t *__restrict__ visited,
                           uint64_t *__restrict__ front, uint32_t nwords) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dist[i] = i == root ? 1u : 0u;
    if (parent) parent[i] = i == root ? root : n;
  }
  if (i < nwords) {
    const uint64_t w = (root >> 6) == i ? (1ull << (root & 63)) : 0ull;
    visited[i] = w;
    front[i] = w;
  }
}

__global__ void k_frontier_degrees(const uint32_t *__restrict__ q, uint32_t qn,
                                   const uint32_t *__restrict__ ptr, uint32_t *__restrict__ deg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= qn) return;
  const uint32_t u = q[i];
  deg[i] = ptr[u + 1] - ptr[u];
}

// Push step over the queue: block b expands frontier edges
// [b*kEdgesPerBlock, (b+1)*kEdgesPerBlock) of the scanned degree sequence.
// Queue entries all have out-edges, so at most kEdgesPerBlock + 1 of them
// overlap a block's slice.
__global__ __launch_bounds__(kBlock) void k_bfs_push(
    const uint32_t *__restrict__ q, uint32_t qn, const uint32_t *__restrict__ qoff,
    const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ col,
    uint64_t *__restrict__ visited, uint32_t *__restrict__ dist, uint32_t *__restrict__ parent,
    uint32_t newdist, uint32_t *__restrict__ nq, Ctr ctr) {
  __shared__ uint32_t s_off[kEdgesPerBlock + 1];
  __shared__ uint32_t s_u[kEdgesPerBlock + 1];
  __shared__ uint32_t s_first, s_cnt;
  const uint32_t e0 = blockIdx.x * kEdgesPerBlock;
  const uint32_t total = qoff[qn];
  const uint32_t e1 = e0 + kEdgesPerBlock < total ? e0 + kEdgesPerBlock : total;
  if (threadIdx.x == 0) {
    uint32_t lo = 0, hi = qn;  // last i with qoff[i] <= e0
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (qoff[mid] <= e0) lo = mid; else hi = mid;
    }
    uint32_t a = lo, b = qn;  // first i > lo with qoff[i] >= e1
    while (a < b) {
      const uint32_t mid = a + (b - a) / 2;
      if (qoff[mid] < e1) a = mid + 1; else b = mid;
    }
    s_first = lo;
    s_cnt = a - lo;
  }
  __syncthreads();
  const uint32_t first = s_first;
  const uint32_t cnt = s_cnt <= kEdgesPerBlock + 1 ? s_cnt : kEdgesPerBlock + 1;
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
    s_off[i] = qoff[first + i];
    s_u[i] = q[first + i];
  }
  __syncthreads();
  for (uint32_t base = e0; base < e1; base += kBlock) {
    const uint32_t e = base + threadIdx.x;
    bool won = false;
    uint32_t v = 0, deg = 0;
    if (e < e1) {
      uint32_t lo = 0, hi = cnt;  // last j with s_off[j] <= e
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_off[mid] <= e) lo = mid; else hi = mid;
      }
      const uint32_t u = s_u[lo];
      v = col[ptr[u] + (e - s_off[lo])];
      const uint64_t bit = 1ull << (v & 63);
      if (!(__atomic_load_n(&visited[v >> 6], __ATOMIC_RELAXED) & bit)) {
        const uint64_t old = atomicOr(reinterpret_cast<unsigned long long *>(&visited[v >> 6]),
                                      (unsigned long long)bit);
        if (!(old & bit)) {
          won = true;
          dist[v] = newdist;
          if (parent) parent[v] = u;
          deg = ptr[v + 1] - ptr[v];
        }
      }
    }
    wave_emit(ctr, won, v, deg, nq);
  }
}

// Frontier queue -> frontier bitmap (before a pull level).
__global__ void k_queue_to_bitmap(const uint32_t *__restrict__ q, uint32_t qn,
                                  uint64_t *__restrict__ bm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= qn) return;
  const uint32_t v = q[i];
  atomicOr(reinterpret_cast<unsigned long long *>(&bm[v >> 6]), 1ull << (v & 63));
}

// Pull step: wave per visited word.
__global__ __launch_bounds__(kBlock) void k_bfs_pull(
    uint32_t n, uint32_t nwords, const uint32_t *__restrict__ ptr,
    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ out_ptr,
    const uint64_t *__restrict__ front, uint64_t *__restrict__ next,
    uint64_t *__restrict__ visited, uint32_t *__restrict__ dist, uint32_t *__restrict__ parent,
    uint32_t newdist, uint32_t *__restrict__ nq, Ctr ctr) {
  const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = __lane_id();
  if (w >= nwords) return;
  const uint64_t vis = visited[w];
  const uint32_t v = w * 64 + lane;
  bool found = false;
  uint32_t deg = 0;
  if (v < n && !((vis >> lane) & 1ull)) {
    const uint32_t k1 = ptr[v + 1];
    for (uint32_t k = ptr[v]; k < k1; k++) {
      const uint32_t u = idx[k];
      if ((front[u >> 6] >> (u & 63)) & 1ull) {
        found = true;
        dist[v] = newdist;
        if (parent) parent[v] = u;
        deg = out_ptr[v + 1] - out_ptr[v];
        break;
      }
    }
  }
  const uint64_t nw = __ballot(found);
  if (lane == 0) {
    next[w] = nw;
    if (nw) visited[w] = vis | nw;
  }
  wave_emit(ctr, found, v, deg, nq);
}

// Edge sweep for grid (row-active filter) and edge-array layouts.
__global__ void k_bfs_edges(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            uint64_t m, const uint64_t *__restrict__ front,
                            uint64_t *__restrict__ visited, uint64_t *__restrict__ next,
                            uint32_t *__restrict__ dist, uint32_t *__restrict__ parent,
                            uint32_t newdist, const uint32_t *__restrict__ deg, Ctr ctr) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool won = false;
  uint32_t d = 0;
  if (k < m) {
    const uint32_t s = src[k];
    if ((front[s >> 6] >> (s & 63)) & 1ull) {
      const uint32_t v = dst[k];
      const uint64_t bit = 1ull << (v & 63);
      if (!(__atomic_load_n(&visited[v >> 6], __ATOMIC_RELAXED) & bit)) {
        const uint64_t old = atomicOr(reinterpret_cast<unsigned long long *>(&visited[v >> 6]),
                                      (unsigned long long)bit);
        if (!(old & bit)) {
          won = true;
          dist[v] = newdist;
          if (parent) parent[v] = s;
          atomicOr(reinterpret_cast<unsigned long long *>(&next[v >> 6]),
                   (unsigned long long)bit);
          d = deg[v];
        }
      }
    }
  }
  wave_emit(ctr, won, 0, d, nullptr);
}

// Sum of out-degrees over reached vertices (GTEPS numerator).
__global__ void k_reached_edges(uint32_t n, const uint32_t *__restrict__ dist,
                                const uint32_t *__restrict__ deg,
                                unsigned long long *__restrict__ acc,
                                uint32_t *__restrict__ reached) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  const bool r = v < n && dist[v] != 0;
  unsigned long long e = r ? deg[v] : 0u;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) e += __shfl_xor(e, o);
  const uint64_t b = __ballot(r);
  if (__lane_id() == 0 && b) {
    atomicAdd(acc, e);
    atomicAdd(reached, (uint32_t)__popcll(b));
  }
}

}  // namespace

static int bfs_state(eg_graph *g) {
  const uint32_t n = g->n;
  const uint32_t nwords = (n + 63) / 64;
  if (!g->bfs_dist) {
    EG_TRY(eg_dalloc_t(&g->bfs_dist, n, "dist"));
    EG_TRY(eg_dalloc_t(&g->bfs_parent, n, "parent"));
    EG_TRY(eg_dalloc_t(&g->bfs_visited, nwords, "visited"));
    EG_TRY(eg_dalloc_t(&g->bfs_front, nwords, "frontier"));
    EG_TRY(eg_dalloc_t(&g->bfs_next, nwords, "next frontier"));
    EG_TRY(eg_dalloc_t(&g->bfs_q, (uint64_t)n + 1, "queue"));
    EG_TRY(eg_dalloc_t(&g->bfs_nq, (uint64_t)n + 1, "queue"));
  }
  if (!g->bfs_ctr) EG_TRY(eg_dalloc_t(&g->bfs_ctr, 64, "counters"));
  return EG_OK;
}

struct BfsCounts {
  uint32_t count, qlen;
  uint64_t degsum;
};

static int read_counts(eg_graph *g, BfsCounts *c) {
  EG_HIP(hipMemcpyAsync(g->h_ctr, g->bfs_ctr, 16, hipMemcpyDeviceToHost, g->stream));
  EG_HIP(hipStreamSynchronize(g->stream));
  c->count = g->h_ctr[0];
  c->qlen = g->h_ctr[1];
  c->degsum = (uint64_t)g->h_ctr[2] | ((uint64_t)g->h_ctr[3] << 32);
  return EG_OK;
}

// Adjacency BFS.  The traversal stays on the device; the host reads the
// per-level counters to apply the switch rule and size the next launch.
static int bfs_adjacency(eg_graph *g, uint32_t root, int start_pull, int switch_mode,
                         bool want_parent, eg_bfs_stats *st) {
  const uint32_t n = g->n;
  const uint32_t nwords = (n + 63) / 64;
  hipStream_t s = g->stream;
  const uint32_t *optr = g->out_ptr, *ocol = g->out_col;
  const uint32_t *iptr = g->in_ptr ? g->in_ptr : g->out_ptr;
  const uint32_t *iidx = g->in_ptr ? g->in_idx : g->out_col;
  uint32_t *parent = want_parent ? g->bfs_parent : nullptr;
  Ctr ctr{g->bfs_ctr, g->bfs_ctr + 1, reinterpret_cast<unsigned long long *>(g->bfs_ctr + 2)};
  void *tmp = nullptr;
  const size_t qoff_bytes = (((uint64_t)n + 1) * 4 + 255) / 256 * 256;
  EG_TRY(eg_scratch(g, eg_scan_tmp_bytes(n) + qoff_bytes, &tmp));
  uint32_t *qoff = reinterpret_cast<uint32_t *>(tmp);
  void *scan_tmp = reinterpret_cast<uint8_t *>(tmp) + qoff_bytes;

  hipLaunchKernelGGL(k_bfs_init, dim3(blocks_for(n > nwords ? n : nwords, 256)), dim3(256), 0,
                     s, n, root, g->bfs_dist, parent, g->bfs_visited, g->bfs_front, nwords);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(g->bfs_q, &root, 4, hipMemcpyHostToDevice, s));
  uint64_t count = 1, qdeg = 0;
  uint32_t qn = 1;
  {
    uint32_t h[2];
    EG_HIP(hipMemcpyAsync(h, optr + root, 8, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    qdeg = h[1] - h[0];
  }
  if (qdeg == 0) qn = 0;
  bool pull = start_pull != 0;
  bool bm_valid = true;  // bfs_front holds the current frontier
  const double thr = (double)g->m / 20.0;
  int levels = 0, pull_levels = 0;
  while (count) {
    levels++;
    if ((uint64_t)levels > (uint64_t)n + 1)
      return eg_fail(EG_EDEVICE, "BFS did not terminate within n levels");
    if (qn == 0) break;  // no frontier vertex has out-edges: this is the last level
    const uint32_t newdist = (uint32_t)levels + 1;
    EG_HIP(hipMemsetAsync(g->bfs_ctr, 0, 16, s));
    if (!pull) {
      hipLaunchKernelGGL(k_frontier_degrees, dim3(blocks_for(qn, 256)), dim3(256), 0, s,
                         g->bfs_q, qn, optr, qoff);
      EG_LAUNCHED();
      EG_TRY(eg_scan_excl(s, qoff, qoff, qn, scan_tmp));
      {
        EgProf prof(g, "bfs_push");
        hipLaunchKernelGGL(k_bfs_push, dim3(blocks_for(qdeg, kEdgesPerBlock)), dim3(kBlock), 0, s,
                           g->bfs_q, qn, qoff, optr, ocol, g->bfs_visited, g->bfs_dist, parent,
                           newdist, g->bfs_nq, ctr);
        EG_LAUNCHED();
      }
      bm_valid = false;
    } else {
      pull_levels++;
      if (!bm_valid) {
        EG_HIP(hipMemsetAsync(g->bfs_front, 0, (size_t)nwords * 8, s));
        hipLaunchKernelGGL(k_queue_to_bitmap, dim3(blocks_for(qn, 256)), dim3(256), 0, s,
                           g->bfs_q, qn, g->bfs_front);
        EG_LAUNCHED();
      }
      {
        EgProf prof(g, "bfs_pull");
        hipLaunchKernelGGL(k_bfs_pull, dim3(blocks_for((uint64_t)nwords * 64, kBlock)),
                           dim3(kBlock), 0, s, n, nwords, iptr, iidx, optr, g->bfs_front,
                           g->bfs_next, g->bfs_visited, g->bfs_dist, parent, newdist, g->bfs_nq,
                           ctr);
        EG_LAUNCHED();
      }
      std::swap(g->bfs_front, g->bfs_next);
      bm_valid = true;
    }
    std::swap(g->bfs_q, g->bfs_nq);
    BfsCounts c;
    EG_TRY(read_counts(g, &c));
    count = c.count;
    qn = c.qlen;
    qdeg = c.degsum;
    // bfs_simple.c:191-197.
    const double load = (double)count + (double)qdeg;
    pull = (switch_mode && !pull && load > thr) || (pull && load > thr);
  }
  st->levels = levels;
  st->pull_levels = pull_levels;
  return EG_OK;
}

// Grid (bfsgrid) and edge-array (bfs_edgecentric) traversals: one sweep of the
// records per level.
static int bfs_edges(eg_graph *g, uint32_t root, bool want_parent, eg_bfs_stats *st) {
  const uint32_t n = g->n;
  const uint32_t nwords = (n + 63) / 64;
  hipStream_t s = g->stream;
  const uint32_t *src = g->cell_ptr ? g->g_src : g->el_src;
  const uint32_t *dst = g->cell_ptr ? g->g_dst : g->el_dst;
  uint32_t *parent = want_parent ? g->bfs_parent : nullptr;
  Ctr ctr{g->bfs_ctr, g->bfs_ctr + 1, reinterpret_cast<unsigned long long *>(g->bfs_ctr + 2)};
  hipLaunchKernelGGL(k_bfs_init, dim3(blocks_for(n > nwords ? n : nwords, 256)), dim3(256), 0,
                     s, n, root, g->bfs_dist, parent, g->bfs_visited, g->bfs_front, nwords);
  EG_LAUNCHED();
  uint32_t qn = 1;
  int levels = 0;
  while (qn) {
    levels++;
    if ((uint64_t)levels > (uint64_t)n + 1)
      return eg_fail(EG_EDEVICE, "BFS did not terminate within n levels");
    EG_HIP(hipMemsetAsync(g->bfs_ctr, 0, 16, s));
    EG_HIP(hipMemsetAsync(g->bfs_next, 0, (size_t)nwords * 8, s));
    {
      EgProf prof(g, g->cell_ptr ? "bfs_grid" : "bfs_edges");
      hipLaunchKernelGGL(k_bfs_edges, dim3(blocks_for(g->m, 256)), dim3(256), 0, s, src, dst,
                         g->m, g->bfs_front, g->bfs_visited, g->bfs_next, g->bfs_dist, parent,
                         (uint32_t)levels + 1, g->out_deg, ctr);
      EG_LAUNCHED();
    }
    std::swap(g->bfs_front, g->bfs_next);
    BfsCounts c;
    EG_TRY(read_counts(g, &c));
    qn = c.count;
  }
  st->levels = levels;
  st->pull_levels = 0;
  return EG_OK;
}

// Device-resident BFS used by the batched baseline and the benchmark.
int eg_bfs_device(eg_graph *g, uint32_t root, int start_pull, int switch_mode, bool want_parent,
                  eg_bfs_stats *st) {
  EG_TRY(bfs_state(g));
  if (g->out_ptr) return bfs_adjacency(g, root, start_pull, switch_mode, want_parent, st);
  return bfs_edges(g, root, want_parent, st);
}

extern "C" int eg_bfs(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
                      uint32_t *dist, uint32_t *parent, eg_bfs_stats *stats) {
  EG_TRY(eg_require_loaded(g));
  if (root >= g->n) return eg_fail(EG_ERANGE, "root %u >= n = %u", root, g->n);
  EG_TRY(bfs_state(g));
  eg_bfs_stats st = {};
  hipStream_t s = g->stream;
  hipEvent_t e0, e1;
  EG_HIP(hipEventCreate(&e0));
  EG_HIP(hipEventCreate(&e1));
  EG_HIP(hipEventRecord(e0, s));
  int rc = eg_bfs_device(g, root, start_pull, switch_mode, true, &st);
  if (rc == EG_OK) {
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    st.device_ms = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  // Reached vertices and traversed input records (Graph500 convention).
  EG_HIP(hipMemsetAsync(g->bfs_ctr + 8, 0, 16, s));
  hipLaunchKernelGGL(k_reached_edges, dim3(blocks_for(g->n, 256)), dim3(256), 0, s, g->n,
                     g->bfs_dist, g->out_deg,
                     reinterpret_cast<unsigned long long *>(g->bfs_ctr + 8), g->bfs_ctr + 10);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(g->h_ctr + 8, g->bfs_ctr + 8, 12, hipMemcpyDeviceToHost, s));
  if (dist) EG_HIP(hipMemcpyAsync(dist, g->bfs_dist, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
  if (parent)
    EG_HIP(hipMemcpyAsync(parent, g->bfs_parent, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  const uint64_t esum = (uint64_t)g->h_ctr[8] | ((uint64_t)g->h_ctr[9] << 32);
  st.reached = g->h_ctr[10];
  st.edges_traversed = (g->flags & EG_UNDIRECTED) ? esum / 2 : esum;
  if (stats) *stats = st;
  return EG_OK;
}
