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
//  * Adjacency layouts run one k_bfs_level launch per level that applies the
//    reference's switch rule on the device (below); the grid, edge-array and
//    SSSP paths read the level's counters (frontier size, its out-degree sum,
//    queue length) on the host and apply the rule there.
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

// ---------------------------------------------------------------------------
// Device-driven adjacency BFS (bfs_simple.c:136-229): one launch per level and
// no host round trip between levels.  Every workgroup of a level reads the
// frontier's statistics from the control words and applies the switch rule
// (bfs_simple.c:191-197) itself, so the direction is decided on the device.
//
//  * Push walks the frontier queue's edges, kLvlEdges per block (owners by a
//    wave-wide 64-ary search of the queue's edge offsets, staged in LDS).
//  * Pull walks tiles of kTileWords bitmap words (grid-stride over tiles), a
//    wave per word: a lane scans the first kPullSerial entries of its
//    vertex's in-list and the whole wave the rest of a longer list, stopping
//    at the first in-neighbour found in the frontier.
//  * Both emit the next frontier twice: as a bitmap (push: atomicOr into a
//    buffer cleared one level earlier; pull: whole words) and as a queue of
//    its vertices with out-edges plus their edge offsets.  A workgroup
//    collects the winners of a block (push) or tile (pull) in LDS and
//    reserves queue slots and edge offsets with ONE 64-bit atomicAdd of
//    (count << 32 | degree sum), so the offsets increase in queue order
//    without a scan.
//  * Three frontier bitmaps rotate: level L reads F[L % 3], writes
//    F[(L + 1) % 3] and clears F[(L + 2) % 3].
// ---------------------------------------------------------------------------
constexpr uint32_t kLvlEdges = 2048;   // push: frontier edges per block
constexpr uint32_t kTileWords = 32;    // pull: bitmap words per tile (2048 vertices)
constexpr uint32_t kEmitCap = 2048;    // winners per block or tile
constexpr uint32_t kPullSerial = 64;   // pull: entries a lane scans before its wave helps
constexpr int kCtlWords = 12;
// Control words (u64): [2s] = queue length << 32 | out-degree sum and
// [2s + 1] = vertex count of the frontier of the level L with L % 3 == s;
// [6 + s] = direction of that level (1 = pull); [9] = last active level;
// [10] = pull levels; [11] = done.

struct LevelArgs {
  uint32_t n, nwords, ntiles;
  uint64_t m;
  int start_pull, switch_mode;
  const uint32_t *optr, *ocol, *iptr, *iidx;
  // Ordered ids: the traversal runs on relabelled vertices; id_of maps a
  // vertex to its caller id for dist / parent, new_of_old maps the root
  // (both null: the layout's own ids).
  const uint32_t *id_of, *new_of_old;
  uint64_t *visited;
  uint64_t *F[3];
  uint32_t *dist, *parent;
  uint32_t *q[2], *qoff[2];
  unsigned long long *ctl;
};

__global__ void k_bfs_level_init(LevelArgs a, uint32_t root) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) {
    a.dist[i] = i == root ? 1u : 0u;
    if (a.parent) a.parent[i] = i == root ? root : a.n;
  }
  // dist / parent above are in caller ids, everything below in traversal ids
  const uint32_t rn = a.new_of_old ? a.new_of_old[root] : root;
  if (i < a.nwords) {
    const uint64_t w = (rn >> 6) == i ? (1ull << (rn & 63)) : 0ull;
    a.visited[i] = w;
    a.F[0][i] = 0ull;
    a.F[1][i] = w;
    a.F[2][i] = 0ull;
  }
  if (i == 0) {
    const uint32_t d = a.optr[rn + 1] - a.optr[rn];
    for (int k = 0; k < kCtlWords; k++) a.ctl[k] = 0ull;
    a.q[1][0] = rn;
    a.qoff[1][0] = 0u;
    a.ctl[2] = d ? (1ull << 32) | d : 0ull;  // level 1: slot 1
    a.ctl[3] = 1ull;
  }
}

// Smallest i in [lo, hi) with a[i] >= key (hi if none); a is non-decreasing.
// Called by a whole wave; each pass narrows the range 64-fold.
__device__ uint32_t wave_lower_bound(const uint32_t *__restrict__ a, uint32_t lo, uint32_t hi,
                                     uint32_t key) {
  const int lane = __lane_id();
  while (hi > lo) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint64_t p = (uint64_t)lo + (uint64_t)lane * step;
    const bool ge = p >= hi || a[p] >= key;
    const uint64_t mask = __ballot(ge);
    const uint32_t j = mask ? (uint32_t)(__ffsll((unsigned long long)mask) - 1) : 64u;
    if (j == 0) return lo;
    const uint64_t nhi = (uint64_t)lo + (uint64_t)j * step;
    lo = lo + (j - 1) * step + 1;
    hi = nhi < hi ? (uint32_t)nhi : hi;
  }
  return lo;
}

// Appends v for the lanes with put to the workgroup's buffer (one LDS atomic
// per wave).
__device__ __forceinline__ void wg_append(bool put, uint32_t v, uint32_t *s_w, uint32_t *s_n) {
  const uint64_t mask = __ballot(put);
  if (!mask) return;
  uint32_t base = 0;
  if (__lane_id() == 0) base = atomicAdd(s_n, (uint32_t)__popcll(mask));
  base = __shfl(base, 0);
  if (put) s_w[base + __popcll(mask & ((1ull << __lane_id()) - 1))] = v;
}

// Emits the workgroup's buffer to the next queue with one 64-bit atomicAdd.
// Called by the whole workgroup; leaves the buffer empty.
__device__ void wg_emit(const uint32_t *s_w, uint32_t *s_n, uint32_t *s_sum,
                        unsigned long long *s_base, const uint32_t *__restrict__ optr,
                        unsigned long long *slot, uint32_t *__restrict__ q_out,
                        uint32_t *__restrict__ qoff_out) {
  __syncthreads();
  const uint32_t cnt = *s_n;
  __syncthreads();  // every thread has the count before it is reset
  if (cnt == 0) return;
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  if (tid == 0) *s_n = 0;
  const uint32_t per = (cnt + kBlock - 1) / kBlock;
  const uint32_t b0 = tid * per < cnt ? tid * per : cnt;
  const uint32_t b1 = b0 + per < cnt ? b0 + per : cnt;
  uint32_t deg[kEmitCap / kBlock];
  uint32_t mine = 0;
#pragma unroll
  for (uint32_t j = 0; j < kEmitCap / kBlock; j++) {
    const uint32_t i = b0 + j;
    deg[j] = 0;
    if (i < b1) {
      const uint32_t v = s_w[i];
      deg[j] = optr[v + 1] - optr[v];
      mine += deg[j];
    }
  }
  uint32_t inc = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_sum[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    before += w < wave ? s_sum[w] : 0u;
    total += s_sum[w];
  }
  if (tid == 0) *s_base = atomicAdd(slot, ((unsigned long long)cnt << 32) | total);
  __syncthreads();
  const unsigned long long base = *s_base;
  const uint32_t qb = (uint32_t)(base >> 32);
  uint32_t off = (uint32_t)base + before + inc - mine;
#pragma unroll
  for (uint32_t j = 0; j < kEmitCap / kBlock; j++) {
    const uint32_t i = b0 + j;
    if (i < b1) {
      q_out[qb + i] = s_w[i];
      qoff_out[qb + i] = off;
      off += deg[j];
    }
  }
  __syncthreads();  // the buffer is read before the next appends overwrite it
}

__global__ __launch_bounds__(kBlock) void k_bfs_level(LevelArgs a, uint32_t L) {
  __shared__ uint32_t s_off[kLvlEdges + 1];
  __shared__ uint32_t s_u[kLvlEdges + 1];
  __shared__ uint32_t s_w[kEmitCap];
  __shared__ uint32_t s_n, s_first, s_cnt, s_sum[kWaves];
  __shared__ unsigned long long s_base;
  __shared__ unsigned long long s_tmp[kWaves];
  unsigned long long *ctl = a.ctl;
  if (ctl[11]) return;  // an earlier level found the frontier empty
  const uint32_t s0 = L % 3, s1 = (L + 1) % 3, s2 = (L + 2) % 3;
  const unsigned long long packed = ctl[2 * s0];
  const unsigned long long count = ctl[2 * s0 + 1];
  if (count == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[11] = 1ull;
    return;
  }
  const uint32_t qn = (uint32_t)(packed >> 32);
  const uint32_t total = (uint32_t)packed;
  bool pull;
  if (L == 1) {
    pull = a.start_pull != 0;
  } else {
    const bool prev = ctl[6 + (L - 1) % 3] != 0;
    const bool heavy = 20ull * (count + (unsigned long long)total) > a.m;  // load > m/20
    pull = (a.switch_mode && !prev && heavy) || (prev && heavy);
  }
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  if (blockIdx.x == 0 && tid == 0) {
    ctl[6 + s0] = pull ? 1ull : 0ull;
    ctl[9] = L;
    if (pull) ctl[10] += 1ull;
    ctl[2 * s2] = 0ull;  // statistics slot of level L + 2
    ctl[2 * s2 + 1] = 0ull;
  }
  if (tid == 0) s_n = 0;
  __syncthreads();
  const uint32_t newdist = L + 1;
  const uint64_t *F_in = a.F[s0];
  uint64_t *F_out = a.F[s1];
  uint64_t *F_clr = a.F[s2];
  uint32_t *q_out = a.q[(L + 1) & 1], *qoff_out = a.qoff[(L + 1) & 1];
  unsigned long long *slot = &ctl[2 * s1];
  for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const uint32_t w = t * kTileWords + (uint32_t)tid;
    if (tid < (int)kTileWords && w < a.nwords) F_clr[w] = 0ull;
  }
  unsigned long long found = 0;
  if (!pull) {
    const uint32_t *q = a.q[L & 1], *qoff = a.qoff[L & 1];
    const uint32_t nblk = (total + kLvlEdges - 1) / kLvlEdges;
    for (uint32_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
      const uint32_t e0 = blk * kLvlEdges;
      const uint32_t e1 = e0 + kLvlEdges < total ? e0 + kLvlEdges : total;
      __syncthreads();
      if (wave == 0) {
        // owners: the last entry with offset <= e0 up to the first with offset >= e1
        const uint32_t first = wave_lower_bound(qoff, 0, qn, e0 + 1) - 1;
        const uint32_t last = wave_lower_bound(qoff, first, qn, e1);
        if (lane == 0) {
          s_first = first;
          s_cnt = last - first;
        }
      }
      __syncthreads();
      const uint32_t first = s_first;
      const uint32_t cnt = s_cnt <= kLvlEdges + 1 ? s_cnt : kLvlEdges + 1;
      for (uint32_t i = tid; i < cnt; i += kBlock) {
        s_off[i] = qoff[first + i];
        s_u[i] = q[first + i];
      }
      __syncthreads();
      for (uint32_t base = e0; base < e0 + kLvlEdges; base += kBlock) {
        const uint32_t e = base + tid;
        bool put = false;
        uint32_t v = 0;
        if (e < e1) {
          uint32_t lo = 0, hi = cnt;  // last j with s_off[j] <= e
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s_off[mid] <= e) lo = mid; else hi = mid;
          }
          const uint32_t u = s_u[lo];
          v = a.ocol[a.optr[u] + (e - s_off[lo])];
          const uint64_t bit = 1ull << (v & 63);
          if (!(__atomic_load_n(&a.visited[v >> 6], __ATOMIC_RELAXED) & bit)) {
            const uint64_t old = atomicOr(
                reinterpret_cast<unsigned long long *>(&a.visited[v >> 6]), (unsigned long long)bit);
            if (!(old & bit)) {
              const uint32_t vo = a.id_of ? a.id_of[v] : v;
              a.dist[vo] = newdist;
              if (a.parent) a.parent[vo] = a.id_of ? a.id_of[u] : u;
              atomicOr(reinterpret_cast<unsigned long long *>(&F_out[v >> 6]),
                       (unsigned long long)bit);
              found++;
              put = a.optr[v + 1] != a.optr[v];
            }
          }
        }
        wg_append(put, v, s_w, &s_n);
      }
      wg_emit(s_w, &s_n, s_sum, &s_base, a.optr, slot, q_out, qoff_out);
    }
  } else {
    for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
      for (uint32_t j = (uint32_t)wave; j < kTileWords; j += kWaves) {
        const uint32_t w = t * kTileWords + j;
        if (w >= a.nwords) break;  // uniform over the wave
        const uint64_t vis = a.visited[w];
        const uint32_t v = w * 64 + lane;
        bool hit = false, longrow = false;
        uint32_t par = 0;
        if (v < a.n && !((vis >> lane) & 1ull)) {
          const uint32_t k0 = a.iptr[v], k1 = a.iptr[v + 1];
          const uint32_t ke = k1 - k0 > kPullSerial ? k0 + kPullSerial : k1;
          for (uint32_t k = k0; k < ke && !hit; k += 4) {
            uint32_t u[4];
#pragma unroll
            for (int x = 0; x < 4; x++) u[x] = k + x < ke ? a.iidx[k + x] : a.iidx[k];
            uint64_t fw[4];
#pragma unroll
            for (int x = 0; x < 4; x++) fw[x] = F_in[u[x] >> 6];
#pragma unroll
            for (int x = 0; x < 4; x++) {
              if (!hit && ((fw[x] >> (u[x] & 63)) & 1ull)) {
                hit = true;
                par = u[x];
              }
            }
          }
          longrow = !hit && ke < k1;
        }
        // The rest of an in-list longer than kPullSerial without a hit in its
        // first kPullSerial entries: the whole wave scans it, 256 entries per
        // pass, and stops at the first entry (in list order) in the frontier.
        uint64_t lm = __ballot(longrow);
        while (lm) {
          const int b = __ffsll((unsigned long long)lm) - 1;
          lm &= lm - 1;
          const uint32_t vv = w * 64 + (uint32_t)b;
          const uint32_t k0 = a.iptr[vv] + kPullSerial, k1 = a.iptr[vv + 1];
          bool got = false;
          uint32_t gp = 0;
          for (uint32_t base = k0; base < k1 && !got; base += 256) {
            uint32_t u[4];
            bool h[4];
#pragma unroll
            for (int x = 0; x < 4; x++) {
              const uint32_t k = base + (uint32_t)x * 64 + (uint32_t)lane;
              u[x] = k < k1 ? a.iidx[k] : 0u;
            }
#pragma unroll
            for (int x = 0; x < 4; x++) {
              const uint32_t k = base + (uint32_t)x * 64 + (uint32_t)lane;
              h[x] = k < k1 && ((F_in[u[x] >> 6] >> (u[x] & 63)) & 1ull);
            }
#pragma unroll
            for (int x = 0; x < 4; x++) {
              const uint64_t hm = __ballot(h[x]);
              if (!got && hm) {
                got = true;
                gp = __shfl(u[x], __ffsll((unsigned long long)hm) - 1);
              }
            }
          }
          if (got && lane == b) {
            hit = true;
            par = gp;
          }
        }
        if (hit) {
          const uint32_t vo = a.id_of ? a.id_of[v] : v;
          a.dist[vo] = newdist;
          if (a.parent) a.parent[vo] = a.id_of ? a.id_of[par] : par;
          found++;
        }
        const uint64_t nw = __ballot(hit);
        if (lane == 0) {
          F_out[w] = nw;
          if (nw) a.visited[w] = vis | nw;
        }
        wg_append(hit && a.optr[v + 1] != a.optr[v], v, s_w, &s_n);
      }
      wg_emit(s_w, &s_n, s_sum, &s_base, a.optr, slot, q_out, qoff_out);
    }
  }
  found = block_sum_u64(found, s_tmp);
  if (tid == 0 && found) atomicAdd(&ctl[2 * s1 + 1], found);
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

// key[k] = 2^24 - 1 - min(deg(col[k]), 2^24 - 1): ascending keys put the
// highest-degree neighbours first.
__global__ void k_deg_desc_keys(uint64_t m, const uint32_t *__restrict__ col,
                                const uint32_t *__restrict__ deg, uint32_t *__restrict__ key) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) key[k] = 0xFFFFFFu - min(deg[col[k]], 0xFFFFFFu);
}

// Symmetric graphs: a copy of the CSR columns with every row ordered by
// descending neighbour degree (stable), built on first use.  A pull level
// stops at the first neighbour in the frontier, and from the second level on
// the hubs are in it, so hubs first shortens the scans; levels do not depend
// on the order, and a parent is still an in-neighbour in the frontier.
static int bfs_sorted_cols(eg_graph *g) {
  if (g->bfs_col) return EG_OK;
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  uint32_t *key = nullptr, *tcol = nullptr, *tsrc = nullptr, *out = nullptr;
  int rc = EG_OK;
  do {
    if ((rc = eg_dalloc_t(&key, m, "bfs order keys")) ||
        (rc = eg_dalloc_t(&tcol, m, "bfs order columns")) ||
        (rc = eg_dalloc_t(&tsrc, m, "bfs order rows")) ||
        (rc = eg_dalloc_t(&out, m, "bfs columns")))
      break;
    hipLaunchKernelGGL(k_deg_desc_keys, dim3(blocks_for(m, 256)), dim3(256), 0, s, m, g->out_col,
                       g->out_deg, key);
    if (hipGetLastError() != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "bfs order keys launch");
      break;
    }
    void *tmp = nullptr;
    if ((rc = eg_scratch(g, eg_radix_tmp_bytes(m, 2), &tmp))) break;
    EgSortCols c1;
    c1.np = 2;
    c1.in[0] = g->out_col;
    c1.out[0] = tcol;
    c1.in[1] = g->out_src;
    c1.out[1] = tsrc;
    if ((rc = eg_radix_sort(s, m, 24, key, nullptr, c1, tmp))) break;
    EgSortCols c2;
    c2.np = 1;
    c2.in[0] = tcol;
    c2.out[0] = out;
    if ((rc = eg_radix_sort(s, m, eg_bits_for(n), tsrc, nullptr, c2, tmp))) break;
    if (hipStreamSynchronize(s) != hipSuccess) rc = eg_fail(EG_EDEVICE, "bfs order sort");
  } while (0);
  (void)hipStreamSynchronize(s);
  eg_dfree_t(&key);
  eg_dfree_t(&tcol);
  eg_dfree_t(&tsrc);
  if (rc) {
    eg_dfree_t(&out);
    return rc;
  }
  g->bfs_col = out;
  return EG_OK;
}

// row[k] = the row of entry k (a wave per row, grid-stride).
__global__ void k_entry_rows(uint32_t n, const uint32_t *__restrict__ ptr,
                             uint32_t *__restrict__ row) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const int lane = __lane_id();
  for (uint32_t v = wave; v < n; v += nwaves) {
    const uint32_t k1 = ptr[v + 1];
    for (uint32_t k = ptr[v] + lane; k < k1; k += 64) row[k] = v;
  }
}

// out[0] = the first i with deg[i] == 0 (n if none); deg is non-increasing.
__global__ void k_first_zero(uint32_t n, const uint32_t *__restrict__ deg,
                             uint32_t *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (deg[i] == 0 && (i == 0 || deg[i - 1] != 0)) out[0] = i;
  if (i == n - 1 && deg[i] != 0) out[0] = n;
}

// Symmetric graphs, ordered ids: the rows of the degree-ordered copy (new
// ids by descending degree) with every row's entries ascending, so hubs
// first and the hubs' frontier bits packed into the first bitmap words.
// Built on first use.
static int bfs_ordered_cols(eg_graph *g, EgOrdered *o) {
  if (g->bfs_ocol) return EG_OK;
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  uint32_t *row = nullptr, *tcol = nullptr, *trow = nullptr, *out = nullptr;
  int rc = EG_OK;
  do {
    if ((rc = eg_dalloc_t(&row, m, "bfs order rows")) ||
        (rc = eg_dalloc_t(&tcol, m, "bfs order columns")) ||
        (rc = eg_dalloc_t(&trow, m, "bfs order rows")) ||
        (rc = eg_dalloc_t(&out, m, "bfs ordered columns")))
      break;
    hipLaunchKernelGGL(k_entry_rows, dim3(2048), dim3(256), 0, s, n, o->ptr, row);
    if (hipGetLastError() != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "bfs order rows launch");
      break;
    }
    void *tmp = nullptr;
    if ((rc = eg_scratch(g, eg_radix_tmp_bytes(m, 2), &tmp))) break;
    // by column (new id), then stably by row
    EgSortCols c1;
    c1.np = 2;
    c1.in[0] = o->idx;
    c1.out[0] = tcol;
    c1.in[1] = row;
    c1.out[1] = trow;
    if ((rc = eg_radix_sort(s, m, eg_bits_for(n), o->idx, nullptr, c1, tmp))) break;
    EgSortCols c2;
    c2.np = 1;
    c2.in[0] = tcol;
    c2.out[0] = out;
    if ((rc = eg_radix_sort(s, m, eg_bits_for(n), trow, nullptr, c2, tmp))) break;
    // vertices with edges: a prefix of the ordered ids (degrees descend)
    uint32_t nz = 0;
    if (n) hipLaunchKernelGGL(k_first_zero, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, o->deg, row);
    if (n && (hipGetLastError() != hipSuccess ||
              hipMemcpyAsync(&nz, row, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess)) {
      rc = eg_fail(EG_EDEVICE, "bfs order sort");
      break;
    }
    g->bfs_onz = nz;
  } while (0);
  (void)hipStreamSynchronize(s);
  eg_dfree_t(&row);
  eg_dfree_t(&tcol);
  eg_dfree_t(&trow);
  if (rc) {
    eg_dfree_t(&out);
    return rc;
  }
  g->bfs_ocol = out;
  return EG_OK;
}

// The hubs-first rows, timed and counted as a one-time preparation.
static int bfs_rows_timed(eg_graph *g, int mode, EgOrdered *o) {
  if (mode == 2 ? g->bfs_ocol != nullptr : g->bfs_col != nullptr) return EG_OK;
  const double t0 = eg_wall_ms();
  EG_TRY(mode == 2 ? bfs_ordered_cols(g, o) : bfs_sorted_cols(g));
  g->prep_bfs_ms += eg_wall_ms() - t0;
  g->prep_bfs_bytes += 4ull * g->m;
  return EG_OK;
}

int eg_hubs_first_rows(eg_graph *g, const uint32_t **rows) {
  *rows = nullptr;
  if (!eg_reorder_enabled() || !(g->flags & (EG_UNDIRECTED | EG_SYMMETRIC))) return EG_OK;
  EgOrdered *o;
  EG_TRY(eg_ordered(g, &o));
  EG_TRY(bfs_rows_timed(g, 2, o));
  *rows = g->bfs_ocol;
  return EG_OK;
}

// Which rows a symmetric graph's BFS scans: 2 = ordered ids (needs the
// degree-ordered copy), 1 = hubs first in the layout's ids, 0 = as built.
static int bfs_mode(const eg_graph *g) {
  if (!g->out_ptr || !(g->flags & (EG_UNDIRECTED | EG_SYMMETRIC)) || !g->out_src) return 0;
  if (g->tune.bfs_order == 2 && !eg_reorder_enabled()) return 1;
  return g->tune.bfs_order;
}

// Adjacency BFS: k_bfs_level once per level.  The host only enqueues levels,
// in batches, and reads the control words after each batch.
int eg_bfs_prepare(eg_graph *g) {
  const int mode = bfs_mode(g);
  if (mode == 1) return bfs_rows_timed(g, 1, nullptr);
  if (mode == 2) {
    EgOrdered *o;
    EG_TRY(eg_ordered(g, &o));
    return bfs_rows_timed(g, 2, o);
  }
  return EG_OK;
}

static int bfs_adjacency(eg_graph *g, uint32_t root, int start_pull, int switch_mode,
                         bool want_parent, eg_bfs_stats *st) {
  const uint32_t n = g->n;
  const uint32_t nwords = (n + 63) / 64;
  hipStream_t s = g->stream;
  if (!g->bfs_f3) {
    EG_TRY(eg_dalloc_t(&g->bfs_f3, nwords, "frontier bitmap"));
    EG_TRY(eg_dalloc_t(&g->bfs_qoff, (uint64_t)n + 1, "queue offsets"));
    EG_TRY(eg_dalloc_t(&g->bfs_nqoff, (uint64_t)n + 1, "queue offsets"));
  }
  LevelArgs a;
  a.n = n;
  a.nwords = nwords;
  // bitmap words that pull levels scan (and levels clear)
  uint32_t scan_words = nwords;
  a.m = g->m;
  a.start_pull = start_pull;
  a.switch_mode = switch_mode;
  a.optr = g->out_ptr;
  a.ocol = g->out_col;
  a.iptr = g->in_ptr ? g->in_ptr : g->out_ptr;
  a.iidx = g->in_ptr ? g->in_idx : g->out_col;
  a.id_of = a.new_of_old = nullptr;
  const int mode = bfs_mode(g);
  if (mode == 1) {
    // in-lists are the out-lists: one copy, hubs first, serves both directions
    EG_TRY(bfs_rows_timed(g, 1, nullptr));
    a.ocol = a.iidx = g->bfs_col;
    a.iptr = g->out_ptr;
  } else if (mode == 2) {
    EgOrdered *o;
    EG_TRY(eg_ordered(g, &o));
    EG_TRY(bfs_rows_timed(g, 2, o));
    a.optr = a.iptr = o->ptr;
    a.ocol = a.iidx = g->bfs_ocol;
    a.id_of = o->old_of_new;
    a.new_of_old = o->new_of_old;
    // the vertices without edges are the tail of the ordered ids: no level
    // reaches them, so no frontier bit is ever set there (a root without
    // edges ends the traversal after its own level, and every traversal
    // starts by clearing the whole bitmaps)
    scan_words = (g->bfs_onz + 63) / 64;
  }
  a.ntiles = blocks_for(scan_words ? scan_words : 1, kTileWords);
  const uint32_t grid = a.ntiles < 2048 ? a.ntiles : 2048;
  a.visited = g->bfs_visited;
  a.F[0] = g->bfs_front;
  a.F[1] = g->bfs_next;
  a.F[2] = g->bfs_f3;
  a.dist = g->bfs_dist;
  a.parent = want_parent ? g->bfs_parent : nullptr;
  a.q[0] = g->bfs_q;
  a.q[1] = g->bfs_nq;
  a.qoff[0] = g->bfs_qoff;
  a.qoff[1] = g->bfs_nqoff;
  // control words after the counters that other passes use
  a.ctl = reinterpret_cast<unsigned long long *>(g->bfs_ctr) + 16;
  hipLaunchKernelGGL(k_bfs_level_init, dim3(blocks_for(n > nwords ? n : nwords, 256)), dim3(256),
                     0, s, a, root);
  EG_LAUNCHED();
  // the first batch covers the previous traversal's depth, later ones double
  uint64_t L = 1;
  uint32_t batch = g->bfs_depth_hint + 1 > 4 ? g->bfs_depth_hint + 1 : 4;
  if (batch > 64) batch = 64;
  bool first_batch = true;
  uint64_t h[kCtlWords];
  for (;;) {
    {
      EgProf prof(g, "bfs_level");
      for (uint32_t i = 0; i < batch; i++) {
        hipLaunchKernelGGL(k_bfs_level, dim3(grid), dim3(kBlock), 0, s, a, (uint32_t)(L + i));
        EG_LAUNCHED();
      }
    }
    L += batch;
    EG_TRY(read_u64(g, a.ctl, kCtlWords, h));
    if (h[11] || h[2 * (L % 3) + 1] == 0) break;
    if (L > (uint64_t)n + 1) return eg_fail(EG_EDEVICE, "BFS did not terminate within n levels");
    batch = first_batch ? 2 : (batch < 32 ? batch * 2 : 64);
    first_batch = false;
  }
  g->bfs_depth_hint = (uint32_t)h[9];
  st->levels = (int32_t)h[9];
  st->pull_levels = (int32_t)h[10];
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
  if (g->wide) return eg_wide_bfs(g, root, start_pull, switch_mode, want_parent, st);
  if (g->out_ptr) return bfs_adjacency(g, root, start_pull, switch_mode, want_parent, st);
  return bfs_edges(g, root, want_parent, st);
}

extern "C" int eg_bfs(eg_graph *g, uint32_t root, int32_t start_pull, int32_t switch_mode,
                      uint32_t *dist, uint32_t *parent, eg_bfs_stats *stats) {
  EG_TRY(eg_require_loaded_wide_ok(g));
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
