// eg_wide.hip -- direction-optimising BFS over adjacency with 64-bit row
// offsets (graphs with 2^32 stored entries or more).
//
// Reference: the bfs_simple iterator (bfs_simple.c:136-229): a push step over
// the frontier queue (bfs_push, :121-134), a pull step over every unvisited
// vertex that stops at its first in-neighbour in the frontier (bfs_pull,
// :104-120), and the switch rule of :191-197 on the new frontier's size plus
// its out-degree sum against m / 20.  Levels: root 1, level L -> L + 1,
// unreached 0 (bfsgrid_cilk.c:58-60); parents: root itself, unreached n.
//
// GPU form: dist is the visited set (a vertex of the frontier of level L has
// dist L, one found in it gets L + 1, so a pull test needs no frontier
// bitmap).  Push splits every frontier row into segments of kSeg entries; the
// workgroups stride over the scanned segment counts and claim a vertex with
// one compare-and-swap on its dist.  Pull gives each unvisited vertex a lane
// that scans its in-edges four at a time.  Both append the vertices they find
// to the next queue with one atomic per wave and add up their out-degrees;
// the host reads the two counters once per level and applies the rule.
#include <string.h>

#include "eg_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kSeg = 4096;         // row entries per push segment
constexpr uint32_t kPushGrid = 4096;    // workgroups of a push level
constexpr uint32_t kStrideGrid = 8192;  // cap of the vertex-strided kernels

inline uint32_t grid_for(uint64_t count) {
  const uint64_t b = (count + kBlock - 1) / kBlock;
  return b < kStrideGrid ? (b ? (uint32_t)b : 1u) : kStrideGrid;
}

__global__ void k_wide_init(uint32_t n, uint32_t root, uint32_t *__restrict__ dist,
                            uint32_t *__restrict__ parent, uint32_t *__restrict__ q) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    dist[i] = i == root ? 1u : 0u;
    if (parent) parent[i] = i == root ? root : n;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) q[0] = root;
}

__global__ void k_wide_reset(unsigned long long *ctr) {
  if (threadIdx.x < 2) ctr[threadIdx.x] = 0;
}

// segs[i] = push segments of the row of frontier vertex q[i].
__global__ void k_wide_segs(const uint32_t *__restrict__ q, uint32_t qn,
                            const uint64_t *__restrict__ ptr, uint32_t *__restrict__ segs) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < qn; i += stride) {
    const uint32_t u = q[i];
    segs[i] = (uint32_t)((ptr[u + 1] - ptr[u] + kSeg - 1) / kSeg);
  }
}

// Appends the lanes' finds to the next queue (one atomic per wave) and adds
// their out-degrees to the degree sum.
__device__ __forceinline__ void wave_append(bool found, uint32_t v,
                                            const uint32_t *__restrict__ out_deg,
                                            uint32_t *__restrict__ nq,
                                            unsigned long long *__restrict__ ctr) {
  const uint64_t mask = __ballot(found);
  if (!mask) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  unsigned long long deg = found ? out_deg[v] : 0ull;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) deg += __shfl_xor(deg, o);
  unsigned long long base = 0;
  if (lane == leader) {
    base = atomicAdd(&ctr[0], (unsigned long long)__popcll(mask));
    atomicAdd(&ctr[1], deg);
  }
  base = __shfl(base, leader);
  if (found) nq[base + __popcll(mask & eg_lanemask_lt())] = v;
}

// Push: dist[v] = L + 1 for every unvisited out-neighbour v of the frontier
// (queue q, qn vertices), parent[v] = the claiming row's vertex.
__global__ __launch_bounds__(kBlock) void k_wide_push(
    const uint32_t *__restrict__ q, uint32_t qn, const uint32_t *__restrict__ segs,
    const uint64_t *__restrict__ ptr, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ out_deg, uint32_t *dist, uint32_t *__restrict__ parent,
    uint32_t *__restrict__ nq, unsigned long long *__restrict__ ctr, uint32_t L) {
  const uint32_t total = segs[qn];
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    // the frontier vertex whose segments cover t: last i with segs[i] <= t
    uint32_t a = 0, b = qn;
    while (b - a > 1) {
      const uint32_t mid = a + (b - a) / 2;
      if (segs[mid] <= t) a = mid; else b = mid;
    }
    const uint32_t u = q[a];
    const uint64_t lo = ptr[u] + (uint64_t)(t - segs[a]) * kSeg;
    const uint64_t end = ptr[u + 1];
    const uint64_t hi = lo + kSeg < end ? lo + kSeg : end;
    for (uint64_t k0 = lo; k0 < hi; k0 += kBlock) {
      const uint64_t k = k0 + threadIdx.x;
      bool found = false;
      uint32_t v = 0;
      if (k < hi) {
        v = col[k];
        if (dist[v] == 0 && atomicCAS(&dist[v], 0u, L + 1) == 0u) {
          found = true;
          if (parent) parent[v] = u;
        }
      }
      wave_append(found, v, out_deg, nq, ctr);
    }
  }
}

// Pull: every unvisited v scans its in-edges (iptr, iidx) in order and stops
// at the first u with dist[u] == L: dist[v] = L + 1, parent[v] = u.
__global__ __launch_bounds__(kBlock) void k_wide_pull(
    uint32_t n, const uint64_t *__restrict__ iptr, const uint32_t *__restrict__ iidx,
    const uint32_t *__restrict__ out_deg, uint32_t *dist, uint32_t *__restrict__ parent,
    uint32_t *__restrict__ nq, unsigned long long *__restrict__ ctr, uint32_t L) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t b = (uint64_t)blockIdx.x * kBlock; b < n; b += stride) {
    const uint64_t v = b + threadIdx.x;
    bool found = false;
    uint32_t par = 0;
    if (v < n && dist[v] == 0) {
      uint64_t k = iptr[v];
      const uint64_t end = iptr[v + 1];
      for (; !found && k + 4 <= end; k += 4) {
        uint32_t u[4], d[4];
#pragma unroll
        for (int j = 0; j < 4; j++) u[j] = iidx[k + j];
#pragma unroll
        for (int j = 0; j < 4; j++) d[j] = dist[u[j]];
#pragma unroll
        for (int j = 3; j >= 0; j--)
          if (d[j] == L) {
            found = true;
            par = u[j];  // the lowest j wins: the first in-neighbour in order
          }
      }
      for (; !found && k < end; k++) {
        const uint32_t u = iidx[k];
        if (dist[u] == L) {
          found = true;
          par = u;
        }
      }
      if (found) {
        dist[v] = L + 1;
        if (parent) parent[v] = par;
      }
    }
    wave_append(found, (uint32_t)v, out_deg, nq, ctr);
  }
}

}  // namespace

int eg_wide_bfs(eg_graph *g, uint32_t root, int start_pull, int switch_mode, bool want_parent,
                eg_bfs_stats *st) {
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  if (!g->out_ptr64) return eg_fail(EG_ESTATE, "no 64-bit row offsets");
  if (!g->wide_seg) EG_TRY(eg_dalloc_t(&g->wide_seg, (uint64_t)n + 1, "push segments"));
  if (!g->wide_ctr) EG_TRY(eg_dalloc_t(&g->wide_ctr, 2, "wide BFS counters"));
  void *scan_tmp = nullptr;
  EG_TRY(eg_scratch(g, eg_scan_tmp_bytes((uint64_t)n + 1), &scan_tmp));
  const uint64_t *optr = g->out_ptr64;
  const uint32_t *ocol = g->out_col;
  // in-edges: the CSC, or the CSR of a graph declared symmetric (-s)
  const uint64_t *iptr = g->in_ptr64 ? g->in_ptr64 : g->out_ptr64;
  const uint32_t *iidx = g->in_ptr64 ? g->in_idx : g->out_col;
  uint32_t *dist = g->bfs_dist;
  uint32_t *parent = want_parent ? g->bfs_parent : nullptr;
  uint32_t *q = g->bfs_q, *nq = g->bfs_nq;
  unsigned long long *ctr = g->wide_ctr;
  hipLaunchKernelGGL(k_wide_init, dim3(grid_for(n)), dim3(kBlock), 0, s, n, root, dist, parent, q);
  EG_LAUNCHED();
  const double thr = (double)g->m / 20.0;
  uint64_t qn = 1;
  int pull = start_pull ? 1 : 0;
  int levels = 0, pulls = 0;
  while (qn) {
    levels++;
    if ((uint64_t)levels > (uint64_t)n + 1)
      return eg_fail(EG_EDEVICE, "BFS did not terminate within n levels");
    const uint32_t L = (uint32_t)levels;  // dist of the current frontier
    hipLaunchKernelGGL(k_wide_reset, dim3(1), dim3(64), 0, s, ctr);
    EG_LAUNCHED();
    if (!pull) {
      EgProf prof(g, "wide_push");
      hipLaunchKernelGGL(k_wide_segs, dim3(grid_for(qn)), dim3(kBlock), 0, s,
                         static_cast<const uint32_t *>(q), (uint32_t)qn, optr, g->wide_seg);
      EG_LAUNCHED();
      EG_TRY(eg_scan_excl(s, g->wide_seg, g->wide_seg, qn, scan_tmp));
      hipLaunchKernelGGL(k_wide_push, dim3(kPushGrid), dim3(kBlock), 0, s,
                         static_cast<const uint32_t *>(q), (uint32_t)qn,
                         static_cast<const uint32_t *>(g->wide_seg), optr, ocol,
                         static_cast<const uint32_t *>(g->out_deg), dist, parent, nq, ctr, L);
      EG_LAUNCHED();
    } else {
      EgProf prof(g, "wide_pull");
      pulls++;
      hipLaunchKernelGGL(k_wide_pull, dim3(grid_for(n)), dim3(kBlock), 0, s, n, iptr, iidx,
                         static_cast<const uint32_t *>(g->out_deg), dist, parent, nq, ctr, L);
      EG_LAUNCHED();
    }
    unsigned long long h[2];
    EG_HIP(hipMemcpyAsync(g->h_ctr, ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    memcpy(h, g->h_ctr, sizeof(h));
    const uint64_t nqn = h[0], degsum = h[1];
    // bfs_simple.c:191-197
    pull = ((switch_mode && !pull && (double)(nqn + degsum) > thr) ||
            (pull && (double)(nqn + degsum) > thr))
               ? 1
               : 0;
    uint32_t *t = q;
    q = nq;
    nq = t;
    qn = nqn;
  }
  st->levels = levels;
  st->pull_levels = pulls;
  return EG_OK;
}
