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
//    writes dist / parent into an LDS queue that the workgroup appends to the
//    next queue with one atomic.
//  * pull: one wave per bitmap word; each unvisited lane scans its in-edges,
//    four at a time, until it meets a frontier vertex.  The wave writes the
//    next-frontier and visited words whole (no atomics).  Frontier counts are
//    per-workgroup partials reduced by one small kernel.
//  * The queue holds only vertices with out-edges: a vertex without out-edges
//    is never an in-neighbour, so it cannot discover anything either way.
//  * The host reads the level's counters (frontier size, its out-degree sum,
//    queue length) and applies the reference's switch rule.
#include <string.h>

#include "eg_internal.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kEdgesPerBlock = 2048;
constexpr uint32_t kMaxParts = 4096;  // workgroups of the grid-stride kernels

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}
inline uint32_t stride_blocks(uint64_t threads) {
  const uint32_t b = blocks_for(threads, kBlock);
  return b < kMaxParts ? (b ? b : 1) : kMaxParts;
}

// Sum of a per-lane 64-bit value over the workgroup (result valid in thread 0).
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v,
                                                            unsigned long long *s_tmp) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  if (__lane_id() == 0) s_tmp[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWaves; w++) t += s_tmp[w];
  __syncthreads();
  return t;
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

// Counters (u64 words of ctr): [0] next frontier size, [1] its out-degree
// sum, [2] next queue length.
__global__ __launch_bounds__(kBlock) void k_bfs_push(
    const uint32_t *__restrict__ q, uint32_t qn, const uint32_t *__restrict__ qoff,
    const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ col,
    uint64_t *__restrict__ visited, uint32_t *__restrict__ dist, uint32_t *__restrict__ parent,
    uint32_t newdist, uint32_t *__restrict__ nq, unsigned long long *__restrict__ ctr) {
  __shared__ uint32_t s_off[kEdgesPerBlock + 1];
  __shared__ uint32_t s_u[kEdgesPerBlock + 1];
  __shared__ uint32_t s_wq[kEdgesPerBlock];
  __shared__ uint32_t s_first, s_cnt, s_nq, s_base;
  __shared__ unsigned long long s_tmp[kWaves];
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
    s_nq = 0;
  }
  __syncthreads();
  const uint32_t first = s_first;
  const uint32_t cnt = s_cnt <= kEdgesPerBlock + 1 ? s_cnt : kEdgesPerBlock + 1;
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
    s_off[i] = qoff[first + i];
    s_u[i] = q[first + i];
  }
  __syncthreads();
  unsigned long long found = 0, degsum = 0;
  for (uint32_t base = e0; base < e1; base += kBlock) {
    const uint32_t e = base + threadIdx.x;
    if (e < e1) {
      uint32_t lo = 0, hi = cnt;  // last j with s_off[j] <= e
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s_off[mid] <= e) lo = mid; else hi = mid;
      }
      const uint32_t u = s_u[lo];
      const uint32_t v = col[ptr[u] + (e - s_off[lo])];
      const uint64_t bit = 1ull << (v & 63);
      if (!(__atomic_load_n(&visited[v >> 6], __ATOMIC_RELAXED) & bit)) {
        const uint64_t old = atomicOr(reinterpret_cast<unsigned long long *>(&visited[v >> 6]),
                                      (unsigned long long)bit);
        if (!(old & bit)) {
          dist[v] = newdist;
          if (parent) parent[v] = u;
          const uint32_t d = ptr[v + 1] - ptr[v];
          found++;
          degsum += d;
          if (d) s_wq[atomicAdd(&s_nq, 1u)] = v;
        }
      }
    }
  }
  found = block_sum_u64(found, s_tmp);
  degsum = block_sum_u64(degsum, s_tmp);
  if (threadIdx.x == 0) {
    if (s_nq) s_base = (uint32_t)atomicAdd(&ctr[2], (unsigned long long)s_nq);
    if (found) {
      atomicAdd(&ctr[0], found);
      atomicAdd(&ctr[1], degsum);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < s_nq; i += kBlock) nq[s_base + i] = s_wq[i];
}

// Frontier queue -> frontier bitmap (before a pull level).
__global__ void k_queue_to_bitmap(const uint32_t *__restrict__ q, uint32_t qn,
                                  uint64_t *__restrict__ bm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= qn) return;
  const uint32_t v = q[i];
  atomicOr(reinterpret_cast<unsigned long long *>(&bm[v >> 6]), 1ull << (v & 63));
}

// Pull step: a wave per visited word, grid-stride; per-workgroup partials of
// (found, out-degree sum) in part[2*block].
__global__ __launch_bounds__(kBlock) void k_bfs_pull(
    uint32_t n, uint32_t nwords, const uint32_t *__restrict__ ptr,
    const uint32_t *__restrict__ idx, const uint32_t *__restrict__ out_ptr,
    const uint64_t *__restrict__ front, uint64_t *__restrict__ next,
    uint64_t *__restrict__ visited, uint32_t *__restrict__ dist, uint32_t *__restrict__ parent,
    uint32_t newdist, unsigned long long *__restrict__ part) {
  __shared__ unsigned long long s_tmp[kWaves];
  const int lane = __lane_id();
  const uint32_t wave0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  unsigned long long found = 0, degsum = 0;
  for (uint32_t w = wave0; w < nwords; w += nwaves) {
    const uint64_t vis = visited[w];
    const uint32_t v = w * 64 + lane;
    bool hit = false;
    if (v < n && !((vis >> lane) & 1ull)) {
      const uint32_t k1 = ptr[v + 1];
      for (uint32_t k = ptr[v]; k < k1 && !hit; k += 4) {
        uint32_t u[4];
#pragma unroll
        for (int t = 0; t < 4; t++) u[t] = k + t < k1 ? idx[k + t] : idx[k];
        uint64_t fw[4];
#pragma unroll
        for (int t = 0; t < 4; t++) fw[t] = front[u[t] >> 6];
#pragma unroll
        for (int t = 0; t < 4; t++) {
          if (!hit && ((fw[t] >> (u[t] & 63)) & 1ull)) {
            hit = true;
            dist[v] = newdist;
            if (parent) parent[v] = u[t];
          }
        }
      }
      if (hit) {
        found++;
        degsum += out_ptr[v + 1] - out_ptr[v];
      }
    }
    const uint64_t nw = __ballot(hit);
    if (lane == 0) {
      next[w] = nw;
      if (nw) visited[w] = vis | nw;
    }
  }
  found = block_sum_u64(found, s_tmp);
  degsum = block_sum_u64(degsum, s_tmp);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = found;
    part[2 * blockIdx.x + 1] = degsum;
  }
}

// out[c] = sum over p of part[p * ncols + c] (one workgroup, fixed order).
__global__ __launch_bounds__(kBlock) void k_sum_parts(const unsigned long long *__restrict__ part,
                                                      uint32_t nparts, uint32_t ncols,
                                                      unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_tmp[kWaves];
  for (uint32_t c = 0; c < ncols; c++) {
    unsigned long long acc = 0;
    for (uint32_t p = threadIdx.x; p < nparts; p += kBlock) acc += part[(uint64_t)p * ncols + c];
    acc = block_sum_u64(acc, s_tmp);
    if (threadIdx.x == 0) out[c] = acc;
  }
}

// Bitmap -> queue of set vertices with out-edges: per-workgroup counts over a
// fixed word range, then an emit pass at the scanned offsets.
constexpr uint32_t kWordsPerBlock = kBlock;

__device__ __forceinline__ uint64_t kept_bits(const uint64_t *__restrict__ bm, uint32_t w,
                                              uint32_t nwords, uint32_t n,
                                              const uint32_t *__restrict__ out_ptr) {
  uint64_t word = w < nwords ? bm[w] : 0ull;
  uint64_t keep = 0;
  while (word) {
    const int b = __ffsll((unsigned long long)word) - 1;
    const uint32_t v = w * 64 + b;
    if (v < n && out_ptr[v + 1] > out_ptr[v]) keep |= 1ull << b;
    word &= word - 1;
  }
  return keep;
}

__global__ __launch_bounds__(kBlock) void k_bm_count(const uint64_t *__restrict__ bm,
                                                     uint32_t nwords, uint32_t n,
                                                     const uint32_t *__restrict__ out_ptr,
                                                     uint32_t *__restrict__ cnt) {
  __shared__ unsigned long long s_tmp[kWaves];
  const uint32_t w = blockIdx.x * kWordsPerBlock + threadIdx.x;
  unsigned long long c = (unsigned long long)__popcll(kept_bits(bm, w, nwords, n, out_ptr));
  c = block_sum_u64(c, s_tmp);
  if (threadIdx.x == 0) cnt[blockIdx.x] = (uint32_t)c;
}

__global__ __launch_bounds__(kBlock) void k_bm_emit(const uint64_t *__restrict__ bm,
                                                    uint32_t nwords, uint32_t n,
                                                    const uint32_t *__restrict__ out_ptr,
                                                    const uint32_t *__restrict__ off,
                                                    uint32_t *__restrict__ q) {
  __shared__ uint32_t s_c[kBlock];
  const uint32_t w = blockIdx.x * kWordsPerBlock + threadIdx.x;
  uint64_t keep = kept_bits(bm, w, nwords, n, out_ptr);
  const uint32_t mine = (uint32_t)__popcll(keep);
  // inclusive scan of the per-word counts over the workgroup (Hillis-Steele)
  s_c[threadIdx.x] = mine;
  __syncthreads();
  for (uint32_t o = 1; o < kBlock; o <<= 1) {
    const uint32_t add = threadIdx.x >= o ? s_c[threadIdx.x - o] : 0u;
    __syncthreads();
    s_c[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t pos = off[blockIdx.x] + s_c[threadIdx.x] - mine;
  while (keep) {
    const int b = __ffsll((unsigned long long)keep) - 1;
    q[pos++] = w * 64 + b;
    keep &= keep - 1;
  }
}

// Edge sweep for grid and edge-array layouts (grid-stride, partial counts).
__global__ __launch_bounds__(kBlock) void k_bfs_edges(
    const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint64_t m,
    const uint64_t *__restrict__ front, uint64_t *__restrict__ visited,
    uint64_t *__restrict__ next, uint32_t *__restrict__ dist, uint32_t *__restrict__ parent,
    uint32_t newdist, unsigned long long *__restrict__ part) {
  __shared__ unsigned long long s_tmp[kWaves];
  unsigned long long found = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride) {
    const uint32_t s = __builtin_nontemporal_load(src + k);
    if ((front[s >> 6] >> (s & 63)) & 1ull) {
      const uint32_t v = __builtin_nontemporal_load(dst + k);
      const uint64_t bit = 1ull << (v & 63);
      if (!(__atomic_load_n(&visited[v >> 6], __ATOMIC_RELAXED) & bit)) {
        const uint64_t old = atomicOr(reinterpret_cast<unsigned long long *>(&visited[v >> 6]),
                                      (unsigned long long)bit);
        if (!(old & bit)) {
          found++;
          dist[v] = newdist;
          if (parent) parent[v] = s;
          atomicOr(reinterpret_cast<unsigned long long *>(&next[v >> 6]),
                   (unsigned long long)bit);
        }
      }
    }
  }
  found = block_sum_u64(found, s_tmp);
  if (threadIdx.x == 0) part[blockIdx.x] = found;
}

// Reached vertices and the sum of their out-degrees (GTEPS numerator).
__global__ __launch_bounds__(kBlock) void k_reached_edges(uint32_t n,
                                                          const uint32_t *__restrict__ dist,
                                                          const uint32_t *__restrict__ out_deg,
                                                          unsigned long long *__restrict__ part) {
  __shared__ unsigned long long s_tmp[kWaves];
  unsigned long long e = 0, r = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    if (dist[v]) {
      r++;
      e += out_deg[v];
    }
  }
  e = block_sum_u64(e, s_tmp);
  r = block_sum_u64(r, s_tmp);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = e;
    part[2 * blockIdx.x + 1] = r;
  }
}


// ---------------------------------------------------------------------------
// SSSP (sssp_pushpull.c): frontier Bellman-Ford on the same bitmaps.
// ---------------------------------------------------------------------------
constexpr uint32_t kSsspInf = 0xFFFFFFFEu;  // UINT32_MAX - 1, sssp_pushpull.c:8

// Edge weight: (u % 10) + (v % 10) as CREATE_WEIGHT builds it
// (init_all.c:661-686), or 1.
__device__ __forceinline__ uint32_t sssp_w(uint32_t u, uint32_t v, int unit) {
  return unit ? 1u : (u % 10u) + (v % 10u);
}

__global__ void k_sssp_init(uint32_t n, uint32_t nwords, uint32_t root, uint32_t *__restrict__ dist,
                            uint64_t *__restrict__ front, uint64_t *__restrict__ next) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dist[i] = i == root ? 0u : kSsspInf;
  if (i < nwords) {
    front[i] = (root >> 6) == i ? 1ull << (root & 63) : 0ull;
    next[i] = 0ull;
  }
}

// Push: the frontier queue's edges, kEdgesPerBlock per workgroup (owners by
// binary search in LDS).  A shorter distance is published with atomicMin and
// the vertex joins the next frontier.
__global__ __launch_bounds__(kBlock) void k_sssp_push(
    const uint32_t *__restrict__ q, uint32_t qn, const uint32_t *__restrict__ qoff,
    const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ col,
    uint32_t *__restrict__ dist, uint64_t *__restrict__ next, int unit) {
  __shared__ uint32_t s_off[kEdgesPerBlock + 1];
  __shared__ uint32_t s_u[kEdgesPerBlock + 1];
  __shared__ uint32_t s_first, s_cnt;
  const uint32_t e0 = blockIdx.x * kEdgesPerBlock;
  const uint32_t total = qoff[qn];
  const uint32_t e1 = e0 + kEdgesPerBlock < total ? e0 + kEdgesPerBlock : total;
  if (threadIdx.x == 0) {
    uint32_t lo = 0, hi = qn;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (qoff[mid] <= e0) lo = mid; else hi = mid;
    }
    uint32_t a = lo, b = qn;
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
    if (e >= e1) continue;
    uint32_t lo = 0, hi = cnt;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (s_off[mid] <= e) lo = mid; else hi = mid;
    }
    const uint32_t u = s_u[lo];
    const uint32_t v = col[ptr[u] + (e - s_off[lo])];
    if (v == u) continue;  // sssp_pushpull.c:56
    const uint64_t nd = (uint64_t)__atomic_load_n(&dist[u], __ATOMIC_RELAXED) + sssp_w(u, v, unit);
    if (nd >= __atomic_load_n(&dist[v], __ATOMIC_RELAXED)) continue;
    const uint32_t old = atomicMin(&dist[v], (uint32_t)nd);
    if (nd < old) {
      const uint64_t bit = 1ull << (v & 63);
      if (!(__atomic_load_n(&next[v >> 6], __ATOMIC_RELAXED) & bit))
        atomicOr(reinterpret_cast<unsigned long long *>(&next[v >> 6]), (unsigned long long)bit);
    }
  }
}

// Pull (sssp_pushpull.c:75-104): every vertex takes the minimum over its
// in-neighbours in the frontier.  SSSP has no early exit, so the work follows
// the pull plan: chunks of short rows stage their in-lists in LDS (a lane per
// row of up to kSerialRow entries, a wave per longer row), and hub rows are
// cut into segments, one workgroup each, whose minima k_sssp_pull_finish
// merges.  A vertex whose distance drops joins the next frontier.
constexpr uint32_t kSerialRow = 32;
constexpr uint64_t kNoCand = ~0ull;

__device__ __forceinline__ uint64_t sssp_cand(uint32_t u, uint32_t v,
                                              const uint64_t *__restrict__ front,
                                              const uint32_t *__restrict__ dist, int unit) {
  if (u == v || !((front[u >> 6] >> (u & 63)) & 1ull)) return kNoCand;
  return (uint64_t)__atomic_load_n(&dist[u], __ATOMIC_RELAXED) + sssp_w(u, v, unit);
}

// Lowers dist[v] to best (the row's only writer in a pull pass).
__device__ __forceinline__ void sssp_lower(uint32_t v, uint64_t best, uint32_t *__restrict__ dist,
                                           uint64_t *__restrict__ next) {
  if (best < __atomic_load_n(&dist[v], __ATOMIC_RELAXED)) {
    __atomic_store_n(&dist[v], (uint32_t)best, __ATOMIC_RELAXED);
    atomicOr(reinterpret_cast<unsigned long long *>(&next[v >> 6]), 1ull << (v & 63));
  }
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const uint64_t y = __shfl_xor(x, o);
    x = y < x ? y : x;
  }
  return x;
}

__global__ __launch_bounds__(kBlock) void k_sssp_pull(
    const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
    const uint4 *__restrict__ chunk_info, const uint4 *__restrict__ seg_info,
    const uint32_t *__restrict__ long_rows, uint32_t nseg, uint32_t *__restrict__ seg_min,
    const uint64_t *__restrict__ front, uint64_t *__restrict__ next,
    uint32_t *__restrict__ dist, int unit) {
  constexpr uint32_t kChunk = kPlanChunk, kLds = kPlanChunk + kPlanLongRow;
  static_assert(kLds <= 65535 && kChunk <= 65535, "u16 LDS offsets");
  __shared__ uint32_t s_idx[kLds];
  __shared__ uint16_t s_ptr[kChunk + 1];
  __shared__ uint16_t s_list[kChunk];
  __shared__ uint32_t s_nlist;
  __shared__ uint64_t s_min[kWaves];
  const uint32_t tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  if (blockIdx.x < nseg) {
    const uint4 si = seg_info[blockIdx.x];
    const uint32_t v = long_rows[si.z];
    uint64_t best = kNoCand;
    for (uint32_t base = si.x; base < si.y; base += kBlock * 8) {
      uint32_t u[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const uint32_t k = base + tid + q * kBlock;
        u[q] = k < si.y ? __builtin_nontemporal_load(idx + k) : v;
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const uint64_t c = sssp_cand(u[q], v, front, dist, unit);
        best = c < best ? c : best;
      }
    }
    best = wave_min_u64(best);
    if (lane == 0) s_min[wave] = best;
    __syncthreads();
    if (tid == 0) {
      uint64_t b = s_min[0];
      for (int q = 1; q < kWaves; q++) b = s_min[q] < b ? s_min[q] : b;
      seg_min[blockIdx.x] = b < kSsspInf ? (uint32_t)b : kSsspInf;
    }
    return;
  }
  const uint4 ci = chunk_info[blockIdx.x - nseg];
  const uint32_t r0 = ci.x, r1 = ci.y, k0 = ci.z, k1 = ci.w;
  const uint32_t nrows = r1 - r0;
  if (tid == 0) s_nlist = 0;
  for (uint32_t i = tid; i <= nrows; i += kBlock) s_ptr[i] = (uint16_t)(ptr[r0 + i] - k0);
  for (uint32_t base = k0; base < k1; base += kBlock * 8) {
    uint32_t u[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint32_t k = base + tid + q * kBlock;
      u[q] = k < k1 ? __builtin_nontemporal_load(idx + k) : 0u;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint32_t k = base + tid + q * kBlock;
      if (k < k1) s_idx[k - k0] = u[q];
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < nrows; i += kBlock) {
    const uint32_t ra = s_ptr[i], rb = s_ptr[i + 1];
    if (rb - ra <= kSerialRow) {
      const uint32_t v = r0 + i;
      uint64_t best = kNoCand;
      for (uint32_t q = ra; q < rb; q++) {
        const uint64_t c = sssp_cand(s_idx[q], v, front, dist, unit);
        best = c < best ? c : best;
      }
      sssp_lower(v, best, dist, next);
    } else {
      s_list[atomicAdd(&s_nlist, 1u)] = (uint16_t)i;
    }
  }
  __syncthreads();
  const uint32_t nlist = s_nlist;
  for (uint32_t li = wave; li < nlist; li += kWaves) {
    const uint32_t i = s_list[li];
    const uint32_t v = r0 + i;
    const uint32_t rb = s_ptr[i + 1];
    uint64_t best = kNoCand;
    for (uint32_t q = s_ptr[i] + lane; q < rb; q += 64) {
      const uint64_t c = sssp_cand(s_idx[q], v, front, dist, unit);
      best = c < best ? c : best;
    }
    best = wave_min_u64(best);
    if (lane == 0) sssp_lower(v, best, dist, next);
  }
}

// Merges the segment minima of each long row (in any order: min is exact).
__global__ void k_sssp_pull_finish(const uint32_t *__restrict__ long_rows,
                                   const uint32_t *__restrict__ seg_base, uint32_t nlong,
                                   const uint32_t *__restrict__ seg_min,
                                   uint32_t *__restrict__ dist, uint64_t *__restrict__ next) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlong) return;
  uint32_t best = kSsspInf;
  for (uint32_t s = seg_base[j]; s < seg_base[j + 1]; s++)
    best = seg_min[s] < best ? seg_min[s] : best;
  sssp_lower(long_rows[j], best, dist, next);
}

// Per-workgroup (set bits, their out-degree sum) of a frontier bitmap.
__global__ __launch_bounds__(kBlock) void k_front_stats(const uint64_t *__restrict__ bm,
                                                        uint32_t nwords,
                                                        const uint32_t *__restrict__ out_ptr,
                                                        unsigned long long *__restrict__ part) {
  __shared__ unsigned long long s_tmp[kWaves];
  unsigned long long c = 0, d = 0;
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += gridDim.x * blockDim.x) {
    uint64_t b = bm[w];
    c += __popcll(b);
    while (b) {
      const uint32_t v = w * 64 + (uint32_t)(__ffsll((unsigned long long)b) - 1);
      d += out_ptr[v + 1] - out_ptr[v];
      b &= b - 1;
    }
  }
  c = block_sum_u64(c, s_tmp);
  d = block_sum_u64(d, s_tmp);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = c;
    part[2 * blockIdx.x + 1] = d;
  }
}

// Parents: the first in-neighbour (in list order) on a shortest path; root =
// root, unreached = UINT32_MAX - 1 (the reference's initial value,
// sssp_pushpull.c:203).  Also per-workgroup counts of reached vertices.
__global__ __launch_bounds__(kBlock) void k_sssp_finish(uint32_t n, uint32_t root,
                                                        const uint32_t *__restrict__ ptr,
                                                        const uint32_t *__restrict__ idx,
                                                        const uint32_t *__restrict__ dist,
                                                        uint32_t *__restrict__ parent, int unit,
                                                        unsigned long long *__restrict__ part) {
  __shared__ unsigned long long s_tmp[kWaves];
  unsigned long long reached = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t dv = dist[v];
    reached += dv != kSsspInf;
    if (!parent) continue;
    uint32_t p = kSsspInf;
    if (v == root) {
      p = root;
    } else if (dv != kSsspInf) {
      for (uint32_t k = ptr[v]; k < ptr[v + 1]; k++) {
        const uint32_t u = idx[k];
        if (u != v && dist[u] != kSsspInf && (uint64_t)dist[u] + sssp_w(u, v, unit) == dv) {
          p = u;
          break;
        }
      }
    }
    parent[v] = p;
  }
  reached = block_sum_u64(reached, s_tmp);
  if (threadIdx.x == 0) part[blockIdx.x] = reached;
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
    EG_TRY(eg_dalloc_t(&g->bfs_part, 2 * (uint64_t)kMaxParts, "bfs partials"));
    EG_TRY(eg_dalloc_t(&g->bfs_bmcnt, (uint64_t)blocks_for(nwords, kWordsPerBlock) + 1,
                       "bitmap counts"));
  }
  if (!g->bfs_ctr) EG_TRY(eg_dalloc_t(&g->bfs_ctr, 64, "counters"));
  return EG_OK;
}

static int read_u64(eg_graph *g, const void *d, int count, uint64_t *h) {
  EG_HIP(hipMemcpyAsync(g->h_ctr, d, (size_t)count * 8, hipMemcpyDeviceToHost, g->stream));
  EG_HIP(hipStreamSynchronize(g->stream));
  memcpy(h, g->h_ctr, (size_t)count * 8);
  return EG_OK;
}

// Queue of the vertices with out-edges in bitmap bm; returns its length.
static int bitmap_to_queue(eg_graph *g, const uint64_t *bm, uint32_t *q, uint32_t *qlen,
                           void *scan_tmp) {
  const uint32_t n = g->n, nwords = (n + 63) / 64;
  hipStream_t s = g->stream;
  const uint32_t nb = blocks_for(nwords, kWordsPerBlock);
  uint32_t *cnt = g->bfs_bmcnt;
  hipLaunchKernelGGL(k_bm_count, dim3(nb), dim3(kBlock), 0, s, bm, nwords, n, g->out_ptr, cnt);
  EG_LAUNCHED();
  EG_TRY(eg_scan_excl(s, cnt, cnt, nb, scan_tmp));
  hipLaunchKernelGGL(k_bm_emit, dim3(nb), dim3(kBlock), 0, s, bm, nwords, n, g->out_ptr, cnt, q);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(g->h_ctr, cnt + nb, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  *qlen = g->h_ctr[0];
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
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(g->bfs_ctr);
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
  bool q_valid = true;   // bfs_q holds the current frontier's vertices with out-edges
  const double thr = (double)g->m / 20.0;
  const uint32_t pull_blocks = stride_blocks((uint64_t)nwords * 64);
  int levels = 0, pull_levels = 0;
  while (count) {
    levels++;
    if ((uint64_t)levels > (uint64_t)n + 1)
      return eg_fail(EG_EDEVICE, "BFS did not terminate within n levels");
    const uint32_t newdist = (uint32_t)levels + 1;
    if (!pull) {
      if (!q_valid) EG_TRY(bitmap_to_queue(g, g->bfs_front, g->bfs_q, &qn, scan_tmp));
      if (qn == 0) break;  // no frontier vertex has out-edges: this is the last level
      EG_HIP(hipMemsetAsync(ctr, 0, 24, s));
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
      std::swap(g->bfs_q, g->bfs_nq);
      uint64_t h[3];
      EG_TRY(read_u64(g, ctr, 3, h));
      count = h[0];
      qdeg = h[1];
      qn = (uint32_t)h[2];
      q_valid = true;
      bm_valid = false;
    } else {
      pull_levels++;
      if (!bm_valid) {
        EG_HIP(hipMemsetAsync(g->bfs_front, 0, (size_t)nwords * 8, s));
        if (qn) {
          hipLaunchKernelGGL(k_queue_to_bitmap, dim3(blocks_for(qn, 256)), dim3(256), 0, s,
                             g->bfs_q, qn, g->bfs_front);
          EG_LAUNCHED();
        }
      }
      {
        EgProf prof(g, "bfs_pull");
        hipLaunchKernelGGL(k_bfs_pull, dim3(pull_blocks), dim3(kBlock), 0, s, n, nwords, iptr,
                           iidx, optr, g->bfs_front, g->bfs_next, g->bfs_visited, g->bfs_dist,
                           parent, newdist, g->bfs_part);
        EG_LAUNCHED();
      }
      hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, s, g->bfs_part, pull_blocks, 2u,
                         ctr);
      EG_LAUNCHED();
      std::swap(g->bfs_front, g->bfs_next);
      uint64_t h[2];
      EG_TRY(read_u64(g, ctr, 2, h));
      count = h[0];
      qdeg = h[1];
      bm_valid = true;
      q_valid = false;
    }
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
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(g->bfs_ctr);
  hipLaunchKernelGGL(k_bfs_init, dim3(blocks_for(n > nwords ? n : nwords, 256)), dim3(256), 0,
                     s, n, root, g->bfs_dist, parent, g->bfs_visited, g->bfs_front, nwords);
  EG_LAUNCHED();
  const uint32_t nb = stride_blocks(g->m);
  uint64_t count = 1;
  int levels = 0;
  while (count) {
    levels++;
    if ((uint64_t)levels > (uint64_t)n + 1)
      return eg_fail(EG_EDEVICE, "BFS did not terminate within n levels");
    EG_HIP(hipMemsetAsync(g->bfs_next, 0, (size_t)nwords * 8, s));
    {
      EgProf prof(g, g->cell_ptr ? "bfs_grid" : "bfs_edges");
      hipLaunchKernelGGL(k_bfs_edges, dim3(nb), dim3(kBlock), 0, s, src, dst, g->m, g->bfs_front,
                         g->bfs_visited, g->bfs_next, g->bfs_dist, parent, (uint32_t)levels + 1,
                         g->bfs_part);
      EG_LAUNCHED();
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, s, g->bfs_part, nb, 1u, ctr);
    EG_LAUNCHED();
    std::swap(g->bfs_front, g->bfs_next);
    EG_TRY(read_u64(g, ctr, 1, &count));
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
  return eg_bfs_host(g, root, start_pull, switch_mode, dist, parent, stats);
}

int eg_bfs_host(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
                uint32_t *dist, uint32_t *parent, eg_bfs_stats *stats) {
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
  const uint32_t nb = stride_blocks(g->n);
  hipLaunchKernelGGL(k_reached_edges, dim3(nb), dim3(kBlock), 0, s, g->n, g->bfs_dist, g->out_deg,
                     g->bfs_part);
  EG_LAUNCHED();
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(g->bfs_ctr);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, s, g->bfs_part, nb, 2u, ctr);
  EG_LAUNCHED();
  if (dist) EG_HIP(hipMemcpyAsync(dist, g->bfs_dist, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
  if (parent)
    EG_HIP(hipMemcpyAsync(parent, g->bfs_parent, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
  uint64_t h[2];
  EG_TRY(read_u64(g, ctr, 2, h));
  st.reached = h[1];
  st.edges_traversed = (g->flags & EG_UNDIRECTED) ? h[0] / 2 : h[0];
  if (stats) *stats = st;
  return EG_OK;
}

// ---------------------------------------------------------------------------
// SSSP driver: the iterator of sssp_pushpull.c:139-188 with the bfs_simple
// switch rule (:174-179), until no distance changes.
// ---------------------------------------------------------------------------
extern "C" int eg_sssp(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
                       int32_t weights, uint32_t *dist, uint32_t *parent, eg_sssp_stats *stats) {
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "SSSP needs adjacency lists (modes 2/3/4)");
  if (root >= g->n) return eg_fail(EG_ERANGE, "root %u >= n = %u", root, g->n);
  if (weights != EG_SSSP_WEIGHT_REF && weights != EG_SSSP_WEIGHT_UNIT)
    return eg_fail(EG_EINVAL, "unknown SSSP weight kind %d", weights);
  EG_TRY(bfs_state(g));
  const uint32_t n = g->n, nwords = (n + 63) / 64;
  const int unit = weights == EG_SSSP_WEIGHT_UNIT;
  hipStream_t s = g->stream;
  const uint32_t *optr = g->out_ptr, *ocol = g->out_col;
  const uint32_t *iptr, *iidx;
  const float *iw;
  PullPlan *pl;
  EG_TRY(eg_pull_orientation(g, &iptr, &iidx, &iw, &pl));
  if (g->sssp_segmin_cap < pl->nseg) {
    eg_dfree_t(&g->sssp_segmin);
    g->sssp_segmin_cap = 0;
    EG_TRY(eg_dalloc_t(&g->sssp_segmin, pl->nseg, "SSSP segment minima"));
    g->sssp_segmin_cap = pl->nseg;
  }
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(g->bfs_ctr);
  uint32_t *qoff = g->bfs_nq;
  void *scan_tmp = nullptr;
  EG_TRY(eg_scratch(g, eg_scan_tmp_bytes((uint64_t)n + 1), &scan_tmp));
  hipEvent_t e0, e1;
  EG_HIP(hipEventCreate(&e0));
  EG_HIP(hipEventCreate(&e1));
  EG_HIP(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_sssp_init, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, nwords, root,
                     g->bfs_dist, g->bfs_front, g->bfs_next);
  EG_LAUNCHED();
  uint64_t h[2];
  uint64_t count = 1;
  EG_HIP(hipMemcpyAsync(g->h_ctr, optr + root, 8, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  uint64_t degsum = (uint64_t)g->h_ctr[1] - g->h_ctr[0];
  bool pull = start_pull != 0;
  const double thr = (double)g->m / 20.0;
  const uint32_t sblocks = stride_blocks(nwords);
  int iters = 0, pull_iters = 0;
  int rc = EG_OK;
  while (count && rc == EG_OK) {
    if ((uint64_t)iters > (uint64_t)n + 1) {
      rc = eg_fail(EG_EDEVICE, "SSSP did not converge within n rounds");
      break;
    }
    iters++;
    EG_HIP(hipMemsetAsync(g->bfs_next, 0, (size_t)nwords * 8, s));
    if (!pull) {
      uint32_t qn = 0;
      EG_TRY(bitmap_to_queue(g, g->bfs_front, g->bfs_q, &qn, scan_tmp));
      if (qn && degsum) {
        hipLaunchKernelGGL(k_frontier_degrees, dim3(blocks_for(qn, 256)), dim3(256), 0, s, g->bfs_q,
                           qn, optr, qoff);
        EG_LAUNCHED();
        EG_TRY(eg_scan_excl(s, qoff, qoff, qn, scan_tmp));
        EgProf prof(g, "sssp_push");
        hipLaunchKernelGGL(k_sssp_push, dim3(blocks_for(degsum, kEdgesPerBlock)), dim3(kBlock), 0,
                           s, g->bfs_q, qn, qoff, optr, ocol, g->bfs_dist, g->bfs_next, unit);
        EG_LAUNCHED();
      }
    } else {
      pull_iters++;
      EgProf prof(g, "sssp_pull");
      hipLaunchKernelGGL(k_sssp_pull, dim3(pl->nseg + pl->nchunks), dim3(kBlock), 0, s, iptr,
                         iidx, pl->chunk_info, pl->seg_info, pl->long_rows, pl->nseg,
                         g->sssp_segmin, g->bfs_front, g->bfs_next, g->bfs_dist, unit);
      EG_LAUNCHED();
      if (pl->nlong) {
        hipLaunchKernelGGL(k_sssp_pull_finish, dim3(blocks_for(pl->nlong, 256)), dim3(256), 0, s,
                           pl->long_rows, pl->seg_base, pl->nlong, g->sssp_segmin, g->bfs_dist,
                           g->bfs_next);
        EG_LAUNCHED();
      }
    }
    std::swap(g->bfs_front, g->bfs_next);
    hipLaunchKernelGGL(k_front_stats, dim3(sblocks), dim3(kBlock), 0, s, g->bfs_front, nwords,
                       optr, g->bfs_part);
    EG_LAUNCHED();
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, s, g->bfs_part, sblocks, 2u, ctr);
    EG_LAUNCHED();
    EG_TRY(read_u64(g, ctr, 2, h));
    count = h[0];
    degsum = h[1];
    // bfs_simple.c:191-197 (the same rule, sssp_pushpull.c:174-179)
    const double load = (double)count + (double)degsum;
    pull = (switch_mode && !pull && load > thr) || (pull && load > thr);
  }
  if (rc == EG_OK) {
    const uint32_t fb = stride_blocks(n);
    hipLaunchKernelGGL(k_sssp_finish, dim3(fb), dim3(kBlock), 0, s, n, root, iptr, iidx,
                       g->bfs_dist, parent ? g->bfs_parent : nullptr, unit, g->bfs_part);
    EG_LAUNCHED();
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kBlock), 0, s, g->bfs_part, fb, 1u, ctr);
    EG_LAUNCHED();
  }
  (void)hipEventRecord(e1, s);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (dist) EG_HIP(hipMemcpyAsync(dist, g->bfs_dist, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  if (parent)
    EG_HIP(hipMemcpyAsync(parent, g->bfs_parent, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  EG_TRY(read_u64(g, ctr, 1, h));
  if (stats) {
    stats->iterations = iters;
    stats->pull_iterations = pull_iters;
    stats->reached = h[0];
    stats->device_ms = ms;
  }
  return EG_OK;
}
