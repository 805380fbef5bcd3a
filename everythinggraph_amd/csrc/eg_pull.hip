// eg_pull.hip -- pull PageRank and SpMV over an adjacency orientation.
//
// Reference: pr_algo_pull (pagerank_simple.c:72-84), the PR iterator
// (:110-133) and spmv_algo (spmv.c:98-106).  One iteration computes, for every
// row v, a sum over its entries; PageRank sums contrib[u] = r[u]/outdeg(u),
// SpMV sums w * x[u].
//
// Work decomposition (built once per orientation, eg_plan_build):
//  * Rows are grouped by the merged coordinate key(r) = ptr[r] + r into chunks
//    of kPlanChunk units, so a chunk holds at most kPlanChunk rows and its
//    short rows at most kPlanChunk + kPlanLongRow entries.  Each chunk's row
//    and entry ranges are precomputed, so a workgroup learns its whole task
//    from one 16-byte load, then issues its row-pointer, index and prefetch
//    loads together, gathers the values into LDS (eight independent requests
//    in flight per lane), and reduces rows from LDS: rows of up to kSerialRow
//    entries by one lane each, longer ones by a whole wave.
//  * Rows with more than kPlanLongRow entries (R-MAT hubs) are cut into
//    segments of kPlanSegLen entries, one workgroup each; a finish kernel
//    merges the segments of a row in segment order.
// All sums are FP32 with Neumaier compensation in a fixed order, so results
// are deterministic.  PageRank's dangling mass and L1 delta are reduced in
// FP64 on the device (per-chunk partials, then a last-block reduction).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include "eg_internal.h"
#include "eg_pr.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kChunk = kPlanChunk;
constexpr uint32_t kLongRow = kPlanLongRow;
constexpr uint32_t kSegLen = kPlanSegLen;
constexpr uint32_t kSerialRow = 32;
constexpr uint32_t kLdsVals = kChunk + kLongRow;
constexpr int kFinBlocks = 128;

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

struct PlanDev {
  const uint4 *chunk_info;
  const uint4 *seg_info;
  const uint32_t *long_rows;
  const uint32_t *seg_base;
  float2 *seg_part;
  double2 *chunk_part;
  double2 *fin_part;
  uint32_t *fin_counter;
  uint32_t nchunks, nlong, nseg;
};

PlanDev plan_dev(const PullPlan &p) {
  PlanDev d;
  d.chunk_info = p.chunk_info;
  d.seg_info = p.seg_info;
  d.long_rows = p.long_rows;
  d.seg_base = p.seg_base;
  d.seg_part = p.seg_part;
  d.chunk_part = p.chunk_part;
  d.fin_part = p.fin_part;
  d.fin_counter = p.fin_counter;
  d.nchunks = p.nchunks;
  d.nlong = p.nlong;
  d.nseg = p.nseg;
  return d;
}

// ---------------------------------------------------------------------------
// Planning kernels.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t first_row_at(const uint32_t *__restrict__ ptr, uint32_t n,
                                                 uint64_t target) {
  uint32_t lo = 0, hi = n;  // first r in [0, n] with ptr[r] + r >= target
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)ptr[mid] + mid >= target) hi = mid; else lo = mid + 1;
  }
  return lo;
}

__global__ void k_chunk_info(const uint32_t *__restrict__ ptr, uint32_t n, uint32_t nchunks,
                             uint4 *__restrict__ info) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const uint32_t r0 = first_row_at(ptr, n, (uint64_t)c * kChunk);
  uint32_t r1 = c + 1 == nchunks ? n : first_row_at(ptr, n, (uint64_t)(c + 1) * kChunk);
  if (r1 > r0 && ptr[r1] - ptr[r1 - 1] > kLongRow) r1--;  // a trailing long row is segmented
  info[c] = make_uint4(r0, r1, ptr[r0], ptr[r1]);
}

__global__ void k_long_flags(const uint32_t *__restrict__ ptr, uint32_t n,
                             uint32_t *__restrict__ flags) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  flags[r] = (ptr[r + 1] - ptr[r]) > kLongRow ? 1u : 0u;
}

__global__ void k_long_compact(const uint32_t *__restrict__ ptr, uint32_t n,
                               const uint32_t *__restrict__ pos, uint32_t *__restrict__ long_rows,
                               uint32_t *__restrict__ segcnt) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t d = ptr[r + 1] - ptr[r];
  if (d > kLongRow) {
    long_rows[pos[r]] = r;
    segcnt[pos[r]] = (d + kSegLen - 1) / kSegLen;
  }
}

__global__ void k_seg_info(const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ long_rows,
                           const uint32_t *__restrict__ seg_base, uint32_t nlong,
                           uint4 *__restrict__ seg_info) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlong) return;
  const uint32_t r = long_rows[j];
  const uint32_t k0 = ptr[r], k1 = ptr[r + 1];
  uint32_t k = k0;
  for (uint32_t s = seg_base[j]; s < seg_base[j + 1]; s++, k += kSegLen)
    seg_info[s] = make_uint4(k, k + kSegLen < k1 ? k + kSegLen : k1, j, 0u);
}

// ---------------------------------------------------------------------------
// Iteration kernels.
// ---------------------------------------------------------------------------
// Values of entries k, k + kBlock, ..., k + (B-1)*kBlock (0 past kend).
// All index loads are issued before all gathers so that a lane keeps B
// independent memory requests in flight.  A lane meets its entries in the same
// ascending order for every B, so B changes no result.
template <bool WEIGHTED, int B, bool LDSHOT>
__device__ __forceinline__ void gather_batch(const uint32_t *__restrict__ idx,
                                             const float *__restrict__ w,
                                             const float *__restrict__ x, uint32_t k,
                                             uint32_t kend, uint32_t hot,
                                             const float *__restrict__ s_hot, uint32_t hot_n,
                                             float (&v)[B]) {
  uint32_t u[B];
  float wv[B];
#pragma unroll
  for (int q = 0; q < B; q++) {
    const uint32_t kk = k + q * kBlock;
    u[q] = kk < kend ? __builtin_nontemporal_load(idx + kk) : 0u;
    if (WEIGHTED) wv[q] = kk < kend ? __builtin_nontemporal_load(w + kk) : 0.f;
  }
#pragma unroll
  for (int q = 0; q < B; q++) {
    // In the degree-ordered copy the sources below `hot` take most gathers;
    // colder ones are loaded non-temporally so that they do not evict them.
    // With LDSHOT the first hot_n entries of x (the highest-degree sources of
    // the degree-ordered copy) are read from an LDS copy.
    const float xv = (LDSHOT && u[q] < hot_n) ? s_hot[u[q]]
                     : u[q] < hot             ? x[u[q]]
                                              : __builtin_nontemporal_load(x + u[q]);
    v[q] = (k + q * kBlock < kend) ? (WEIGHTED ? xv * wv[q] : xv) : 0.f;
  }
}

// B: entries per lane per gather pass (16 covers a whole chunk in one pass).
// Row offsets and the long-row list are u16 in LDS (a chunk has at most
// kLdsVals entries and kChunk rows), which leaves room for more workgroups.
// PERSIST: a workgroup first copies x[0, hot_n) into dynamic LDS and then takes
// work items blockIdx.x, blockIdx.x + gridDim.x, ...; otherwise one item per
// workgroup.  Thread-to-entry mapping and every summation order are the same
// in both forms, so they compute the same bits.
template <bool PR, bool WEIGHTED, int B, bool PERSIST>
__global__ __launch_bounds__(kBlock) void k_pull(const uint32_t *__restrict__ ptr,
                                                 const uint32_t *__restrict__ idx,
                                                 const float *__restrict__ w,
                                                 const float *__restrict__ x,
                                                 float *__restrict__ y, PlanDev pl, PrArgs a,
                                                 uint32_t hot_n) {
  static_assert(kLdsVals <= 65535 && kChunk <= 65535, "u16 LDS offsets");
  __shared__ float s_val[kLdsVals];
  __shared__ uint16_t s_ptr[kChunk + 1];
  __shared__ uint16_t s_list[kChunk];
  __shared__ uint32_t s_nlist;
  __shared__ float s_ks[kWaves], s_kc[kWaves];
  __shared__ double s_d[kWaves], s_l[kWaves];
  extern __shared__ float s_hot[];

  const uint32_t tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const uint32_t nitems = pl.nseg + pl.nchunks;
  if (PERSIST) {
    for (uint32_t i = tid; i < hot_n; i += kBlock) s_hot[i] = x[i];
    __syncthreads();
  }
  const uint32_t stride = PERSIST ? gridDim.x : nitems;
  for (uint32_t item = blockIdx.x; item < nitems; item += stride) {
    if (item < pl.nseg) {
      // ---- one segment of a long row ----
      const uint4 si = pl.seg_info[item];
      EgKahan acc;
      for (uint32_t base = si.x; base < si.y; base += kBlock * B) {
        float v[B];
        gather_batch<WEIGHTED, B, PERSIST>(idx, w, x, base + tid, si.y, a.hot, s_hot, hot_n, v);
#pragma unroll
        for (int q = 0; q < B; q++) acc.add(v[q]);
      }
      eg_wave_kahan_allreduce(acc);
      if (lane == 0) {
        s_ks[wave] = acc.s;
        s_kc[wave] = acc.c;
      }
      __syncthreads();
      if (tid == 0) {
        EgKahan t;
        t.s = s_ks[0];
        t.c = s_kc[0];
        for (int q = 1; q < kWaves; q++) t.merge(s_ks[q], s_kc[q]);
        pl.seg_part[item] = make_float2(t.s, t.c);
      }
      if (PERSIST) __syncthreads();
      continue;
    }

    // ---- a chunk of short rows ----
    const uint32_t c = item - pl.nseg;
    const uint4 ci = pl.chunk_info[c];
    const uint32_t r0 = ci.x, r1 = ci.y, k0 = ci.z, k1 = ci.w;
    const uint32_t nrows = r1 - r0;
    if (tid == 0) s_nlist = 0;
    // row pointers, relative to k0
    for (uint32_t i = tid; i <= nrows; i += kBlock) s_ptr[i] = (uint16_t)(ptr[r0 + i] - k0);
    // prefetch the per-row inputs of this lane's first row
    uint32_t deg0 = 0;
    float rold0 = 0.f;
    if (PR && tid < nrows) {
      deg0 = a.out_deg[r0 + tid];
      rold0 = a.rank_old[r0 + tid];
    }
    for (uint32_t base = k0; base < k1; base += kBlock * B) {
      float v[B];
      gather_batch<WEIGHTED, B, PERSIST>(idx, w, x, base + tid, k1, a.hot, s_hot, hot_n, v);
#pragma unroll
      for (int q = 0; q < B; q++) {
        const uint32_t k = base + tid + q * kBlock;
        if (k < k1) s_val[k - k0] = v[q];
      }
    }
    __syncthreads();

    const float basep = PR ? pr_base(a) : 0.f;
    double dang = 0.0, l1 = 0.0;
    for (uint32_t i = tid; i < nrows; i += kBlock) {
      const uint32_t r = r0 + i;
      const uint32_t ra = s_ptr[i], rb = s_ptr[i + 1];
      if (rb - ra <= kSerialRow) {
        EgKahan acc;
        for (uint32_t q = ra; q < rb; q++) acc.add(s_val[q]);
        uint32_t deg = deg0;
        float rold = rold0;
        if (PR && i != tid) {
          deg = a.out_deg[r];
          rold = a.rank_old[r];
        }
        finish_row<PR>(r, acc.value(), y, a, basep, deg, rold, dang, l1);
      } else {
        s_list[atomicAdd(&s_nlist, 1u)] = (uint16_t)i;
      }
    }
    __syncthreads();
    const uint32_t nlist = s_nlist;
    for (uint32_t li = wave; li < nlist; li += kWaves) {
      const uint32_t i = s_list[li];
      const uint32_t ra = s_ptr[i], rb = s_ptr[i + 1];
      EgKahan acc;
      for (uint32_t q = ra + lane; q < rb; q += 64) acc.add(s_val[q]);
      eg_wave_kahan_allreduce(acc);
      if (lane == 0) {
        const uint32_t r = r0 + i;
        const uint32_t deg = PR ? a.out_deg[r] : 0u;
        const float rold = PR ? a.rank_old[r] : 0.f;
        finish_row<PR>(r, acc.value(), y, a, basep, deg, rold, dang, l1);
      }
    }
    if (PR) {
      dang = eg_wave_sum_f64(dang);
      l1 = eg_wave_sum_f64(l1);
      if (lane == 0) {
        s_d[wave] = dang;
        s_l[wave] = l1;
      }
      __syncthreads();
      if (tid == 0) {
        double td = 0.0, tl = 0.0;
        for (int q = 0; q < kWaves; q++) {
          td += s_d[q];
          tl += s_l[q];
        }
        pl.chunk_part[c] = make_double2(td, tl);
      }
    }
    if (PERSIST) __syncthreads();  // LDS reuse by the next item
  }
}

// Long rows: merge segment partials in segment order.  PageRank: also reduce
// the FP64 partials of all chunks and long rows; the last block to finish
// writes the iteration's statistics and re-arms the counter.
template <bool PR>
__global__ __launch_bounds__(kBlock) void k_pull_finish(float *__restrict__ y, PlanDev pl,
                                                        PrArgs a) {
  __shared__ double s_d[kWaves], s_l[kWaves];
  __shared__ uint32_t s_last;
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const uint32_t gtid = blockIdx.x * kBlock + tid;
  const uint32_t gthreads = gridDim.x * kBlock;
  const float basep = PR ? pr_base(a) : 0.f;
  double dang = 0.0, l1 = 0.0;
  // Long rows of up to kWaveSegs segments: a thread each, merging the
  // segment partials in order.  Longer ones (the hubs): a wave each, lanes
  // merging strided partials, then the lanes' pairs in a fixed order.
  constexpr uint32_t kWaveSegs = 32;
  for (uint32_t j = gtid; j < pl.nlong; j += gthreads) {
    const uint32_t s0 = pl.seg_base[j], s1 = pl.seg_base[j + 1];
    if (s1 - s0 > kWaveSegs) continue;
    const uint32_t r = pl.long_rows[j];
    EgKahan acc;
    for (uint32_t s = s0; s < s1; s++) {
      const float2 p = pl.seg_part[s];
      acc.merge(p.x, p.y);
    }
    const uint32_t deg = PR ? a.out_deg[r] : 0u;
    const float rold = PR ? a.rank_old[r] : 0.f;
    finish_row<PR>(r, acc.value(), y, a, basep, deg, rold, dang, l1);
  }
  const uint32_t gwave = gtid >> 6, nwaves = gthreads >> 6;
  for (uint32_t j = gwave; j < pl.nlong; j += nwaves) {
    const uint32_t s0 = pl.seg_base[j], s1 = pl.seg_base[j + 1];
    if (s1 - s0 <= kWaveSegs) continue;  // the same for the whole wave
    const uint32_t r = pl.long_rows[j];
    EgKahan acc;
    for (uint32_t s = s0 + lane; s < s1; s += 64) {
      const float2 p = pl.seg_part[s];
      acc.merge(p.x, p.y);
    }
    eg_wave_kahan_allreduce(acc);
    if (lane == 0) {
      const uint32_t deg = PR ? a.out_deg[r] : 0u;
      const float rold = PR ? a.rank_old[r] : 0.f;
      finish_row<PR>(r, acc.value(), y, a, basep, deg, rold, dang, l1);
    }
  }
  if (!PR) return;
  for (uint32_t c = gtid; c < pl.nchunks; c += gthreads) {
    const double2 p = pl.chunk_part[c];
    dang += p.x;
    l1 += p.y;
  }
  dang = eg_wave_sum_f64(dang);
  l1 = eg_wave_sum_f64(l1);
  if (lane == 0) {
    s_d[wave] = dang;
    s_l[wave] = l1;
  }
  __syncthreads();
  if (tid == 0) {
    double td = 0.0, tl = 0.0;
    for (int q = 0; q < kWaves; q++) {
      td += s_d[q];
      tl += s_l[q];
    }
    pl.fin_part[blockIdx.x] = make_double2(td, tl);
    __threadfence();
    const uint32_t prev = atomicAdd(pl.fin_counter, 1u);
    s_last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return;  // the same for the whole block
  // the last block: every thread sums a strided share of the block partials,
  // then the waves and the block reduce them (a fixed order)
  __threadfence();
  double td = 0.0, tl = 0.0;
  for (uint32_t b = tid; b < gridDim.x; b += kBlock) {
    const double *pp = reinterpret_cast<const double *>(&pl.fin_part[b]);
    td += __builtin_nontemporal_load(pp);
    tl += __builtin_nontemporal_load(pp + 1);
  }
  td = eg_wave_sum_f64(td);
  tl = eg_wave_sum_f64(tl);
  __syncthreads();  // s_d / s_l reuse
  if (lane == 0) {
    s_d[wave] = td;
    s_l[wave] = tl;
  }
  __syncthreads();
  if (tid == 0) {
    double sd = 0.0, sl = 0.0;
    for (int q = 0; q < kWaves; q++) {
      sd += s_d[q];
      sl += s_l[q];
    }
    a.stats_out[0] = sd;
    a.stats_out[1] = sl;
    *pl.fin_counter = 0u;
  }
}

__global__ void k_pr_init(uint32_t n, float init, const uint32_t *__restrict__ out_deg,
                          float *__restrict__ rank, float *__restrict__ contrib, double dmass0,
                          double *__restrict__ stats) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) stats[0] = dmass0;
  if (v >= n) return;
  const uint32_t d = out_deg[v];
  rank[v] = init;
  contrib[v] = d ? init / (float)d : 0.f;
}

}  // namespace

void eg_plan_free(PullPlan *pl) {
  eg_dfree_t(&pl->chunk_info);
  eg_dfree_t(&pl->seg_info);
  eg_dfree_t(&pl->long_rows);
  eg_dfree_t(&pl->seg_base);
  eg_dfree_t(&pl->seg_part);
  eg_dfree_t(&pl->chunk_part);
  eg_dfree_t(&pl->fin_part);
  eg_dfree_t(&pl->fin_counter);
  *pl = PullPlan();
}

int eg_plan_build(eg_graph *g, PullPlan *pl, const uint32_t *ptr) {
  if (pl->built) return EG_OK;
  eg_plan_free(pl);
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  hipStream_t s = g->stream;
  const uint64_t nchunks = (m + n) / kChunk + 1;
  if (nchunks >= (1ull << 31)) return eg_fail(EG_ENOTSUP, "graph too large for the pull plan");
  pl->nchunks = (uint32_t)nchunks;
  EG_TRY(eg_dalloc_t(&pl->chunk_info, nchunks, "chunk info"));
  hipLaunchKernelGGL(k_chunk_info, dim3(blocks_for(nchunks, 256)), dim3(256), 0, s, ptr, n,
                     pl->nchunks, pl->chunk_info);
  EG_LAUNCHED();

  uint32_t *flags = nullptr, *segcnt = nullptr;
  EG_TRY(eg_dalloc_t(&flags, (uint64_t)n + 1, "long flags"));
  hipLaunchKernelGGL(k_long_flags, dim3(blocks_for(n, 256)), dim3(256), 0, s, ptr, n, flags);
  EG_LAUNCHED();
  void *tmp = nullptr;
  EG_TRY(eg_scratch(g, eg_scan_tmp_bytes(n), &tmp));
  EG_TRY(eg_scan_excl(s, flags, flags, n, tmp));
  uint32_t nlong = 0;
  EG_HIP(hipMemcpyAsync(&nlong, flags + n, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  pl->nlong = nlong;
  EG_TRY(eg_dalloc_t(&pl->long_rows, nlong ? nlong : 1, "long rows"));
  EG_TRY(eg_dalloc_t(&pl->seg_base, (uint64_t)nlong + 1, "segment base"));
  EG_TRY(eg_dalloc_t(&segcnt, (uint64_t)nlong + 1, "segment counts"));
  if (nlong) {
    hipLaunchKernelGGL(k_long_compact, dim3(blocks_for(n, 256)), dim3(256), 0, s, ptr, n, flags,
                       pl->long_rows, segcnt);
    EG_LAUNCHED();
  }
  EG_TRY(eg_scan_excl(s, segcnt, pl->seg_base, nlong, tmp));
  uint32_t nseg = 0;
  EG_HIP(hipMemcpyAsync(&nseg, pl->seg_base + nlong, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  eg_dfree_t(&flags);
  eg_dfree_t(&segcnt);
  pl->nseg = nseg;
  EG_TRY(eg_dalloc_t(&pl->seg_info, nseg ? nseg : 1, "segment info"));
  if (nlong) {
    hipLaunchKernelGGL(k_seg_info, dim3(blocks_for(nlong, 256)), dim3(256), 0, s, ptr,
                       pl->long_rows, pl->seg_base, nlong, pl->seg_info);
    EG_LAUNCHED();
  }
  EG_TRY(eg_dalloc_t(&pl->seg_part, nseg ? nseg : 1, "segment partials"));
  EG_TRY(eg_dalloc_t(&pl->chunk_part, nchunks, "chunk partials"));
  EG_TRY(eg_dalloc_t(&pl->fin_part, kFinBlocks, "finish partials"));
  EG_TRY(eg_dalloc_t(&pl->fin_counter, 1, "finish counter"));
  EG_HIP(hipMemsetAsync(pl->fin_counter, 0, 4, s));
  EG_HIP(hipStreamSynchronize(s));
  pl->built = true;
  return EG_OK;
}

// Orientation used by the pull kernels: in-edges when the CSC exists, else
// the CSR of a symmetric graph (-s).
int eg_pull_orientation(eg_graph *g, const uint32_t **ptr, const uint32_t **idx, const float **w,
                        PullPlan **pl) {
  if (g->in_ptr) {
    *ptr = g->in_ptr;
    *idx = g->in_idx;
    *w = g->in_w;
    *pl = &g->plan_in;
  } else if (g->out_ptr) {
    *ptr = g->out_ptr;
    *idx = g->out_col;
    *w = g->out_w;
    *pl = &g->plan_out;
  } else {
    return eg_fail(EG_ENOTSUP, "no adjacency layout");
  }
  return eg_plan_build(g, *pl, *ptr);
}

template <int B, bool PERSIST>
static void launch_k_pull(hipStream_t s, bool pr, uint32_t grid, uint32_t lds, const uint32_t *ptr,
                          const uint32_t *idx, const float *w, const float *x, float *y,
                          const PlanDev &d, const PrArgs &a, uint32_t hot_n) {
  if (pr)
    hipLaunchKernelGGL((k_pull<true, false, B, PERSIST>), dim3(grid), dim3(kBlock), lds, s, ptr,
                       idx, w, x, y, d, a, hot_n);
  else if (w)
    hipLaunchKernelGGL((k_pull<false, true, B, PERSIST>), dim3(grid), dim3(kBlock), lds, s, ptr,
                       idx, w, x, y, d, a, hot_n);
  else
    hipLaunchKernelGGL((k_pull<false, false, B, PERSIST>), dim3(grid), dim3(kBlock), lds, s, ptr,
                       idx, w, x, y, d, a, hot_n);
}

// Persistent form: per_cu workgroups per compute unit, each holding as many
// leading entries of x in LDS as the unit's 160 KB leave beside the static
// arrays of its workgroups.
template <int B>
static int launch_k_pull_persist(eg_graph *g, hipStream_t s, bool pr, uint32_t nitems,
                                 const uint32_t *ptr, const uint32_t *idx, const float *w,
                                 const float *x, float *y, const PlanDev &d, const PrArgs &a) {
  int cus = 0, lds_max = 0;
  EG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
  EG_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device));
  const void *fn = pr ? reinterpret_cast<const void *>(&k_pull<true, false, B, true>)
                   : w ? reinterpret_cast<const void *>(&k_pull<false, true, B, true>)
                       : reinterpret_cast<const void *>(&k_pull<false, false, B, true>);
  hipFuncAttributes attr;
  EG_HIP(hipFuncGetAttributes(&attr, fn));
  const uint32_t per_cu = (uint32_t)g->tune.pull_persist;
  const uint64_t unit_lds = 160u * 1024u;
  const uint64_t per_wg = unit_lds / per_cu;
  if (per_wg <= attr.sharedSizeBytes + 4096) return eg_fail(EG_EINVAL, "no LDS left for the hot set");
  uint64_t dyn = std::min<uint64_t>(per_wg - attr.sharedSizeBytes - 1024,
                                    (uint64_t)lds_max - attr.sharedSizeBytes);
  const uint32_t hot_n = (uint32_t)std::min<uint64_t>(g->n, dyn / 4);
  dyn = (uint64_t)hot_n * 4;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) {
    (void)hipGetLastError();
    return eg_fail(EG_EINVAL, "the device refuses %llu bytes of dynamic LDS",
                   (unsigned long long)dyn);
  }
  const uint32_t grid = std::min<uint32_t>(nitems, (uint32_t)cus * per_cu);
  launch_k_pull<B, true>(s, pr, grid, (uint32_t)dyn, ptr, idx, w, x, y, d, a, hot_n);
  return EG_OK;
}

static int launch_pull(eg_graph *g, bool pr, const uint32_t *ptr, const uint32_t *idx,
                       const float *w, const float *x, float *y, const PullPlan &pl,
                       const PrArgs &a_in, hipStream_t stream = nullptr) {
  hipStream_t s = stream ? stream : g->stream;
  const PlanDev d = plan_dev(pl);
  PrArgs a = a_in;
  const bool ordered = g->ord && &pl == &g->ord->plan;
  a.hot = ordered && g->tune.pull_hot > 0 ? (uint32_t)g->tune.pull_hot : 0xFFFFFFFFu;
  const uint32_t grid = pl.nseg + pl.nchunks;
  if (grid) {
    EgProf prof(g, pr ? "pr_pull" : "spmv_pull", s);
    if (ordered && g->tune.pull_persist > 0) {
      if (g->tune.pull_batch == 8)
        EG_TRY(launch_k_pull_persist<8>(g, s, pr, grid, ptr, idx, w, x, y, d, a));
      else
        EG_TRY(launch_k_pull_persist<16>(g, s, pr, grid, ptr, idx, w, x, y, d, a));
    } else if (g->tune.pull_batch == 8) {
      launch_k_pull<8, false>(s, pr, grid, 0, ptr, idx, w, x, y, d, a, 0);
    } else {
      launch_k_pull<16, false>(s, pr, grid, 0, ptr, idx, w, x, y, d, a, 0);
    }
    EG_LAUNCHED();
  }
  if (pr || pl.nlong) {
    EgProf prof(g, pr ? "pr_finish" : "spmv_finish", s);
    if (pr)
      hipLaunchKernelGGL((k_pull_finish<true>), dim3(kFinBlocks), dim3(kBlock), 0, s, y, d, a);
    else
      hipLaunchKernelGGL((k_pull_finish<false>), dim3(kFinBlocks), dim3(kBlock), 0, s, y, d, a);
    EG_LAUNCHED();
  }
  return EG_OK;
}

int eg_pagerank_grid(eg_graph *g, const eg_pr_opts *o, float *rank_out, double *dh, double *lh);
int eg_pagerank_edges(eg_graph *g, const eg_pr_opts *o, float *rank_out, double *dh, double *lh);
int eg_spmv_grid(eg_graph *g, const float *x, float *y);
int eg_spmv_edges(eg_graph *g, const float *x, float *y);

int eg_pr_state(eg_graph *g, int iters) {
  const uint32_t n = g->n;
  if (!g->rank_a) {
    EG_TRY(eg_dalloc_t(&g->rank_a, n, "rank"));
    EG_TRY(eg_dalloc_t(&g->rank_b, n, "rank"));
    EG_TRY(eg_dalloc_t(&g->contrib_a, n, "contrib"));
    EG_TRY(eg_dalloc_t(&g->contrib_b, n, "contrib"));
  }
  if (g->pr_stats_cap < iters + 1) {
    eg_dfree_t(&g->pr_stats);
    EG_TRY(eg_dalloc_t(&g->pr_stats, 2 * ((uint64_t)iters + 1), "pr stats"));
    g->pr_stats_cap = iters + 1;
  }
  if (!g->bfs_ctr) EG_TRY(eg_dalloc_t(&g->bfs_ctr, 64, "counters"));
  return EG_OK;
}

// Shared by the adjacency, grid and edge-array PageRank drivers: resolves the
// options into the recurrence's constants.
void eg_pr_resolve(const eg_graph *g, const eg_pr_opts *o, double *damping, int *iters,
                   double *init, int *uniform) {
  *damping = o ? o->damping : 0.85;
  *iters = o ? o->iterations : 10;
  *init = (o && o->init > 0.0) ? o->init : 1.0 / (double)g->n;
  *uniform = o ? o->dangling == EG_PR_DANGLING_UNIFORM : 0;
}

// First iterate: rank = init, contrib = init/outdeg, and the FP64 dangling
// mass of that iterate (init times the number of sinks) in pr_stats[0].
int eg_pr_start(eg_graph *g, double init, const uint32_t *deg, float *rank, float *contrib) {
  const uint32_t n = g->n;
  const double d0 = (double)g->ndangling * (double)(float)init;
  hipLaunchKernelGGL(k_pr_init, dim3(std::max(1u, blocks_for(n, 256))), dim3(256), 0, g->stream, n,
                     (float)init, deg, rank, contrib, d0, g->pr_stats);
  EG_LAUNCHED();
  return EG_OK;
}

// Downloads results and statistics after the iterations; keeps the final
// iterate in rank_a / contrib_a.  With unperm (new id of every vertex) the
// result is gathered back to the vertex ids of the layout through spare.
int eg_pr_finish(eg_graph *g, int iters, float *rank, float *spare, float *rank_out, double *dh,
                 double *lh, const uint32_t *unperm) {
  hipStream_t s = g->stream;
  if (rank_out) {
    const float *src = rank;
    if (unperm) {
      EG_TRY(eg_gather_vec(g, rank, unperm, spare));
      src = spare;
    }
    EG_HIP(hipMemcpyAsync(rank_out, src, (size_t)g->n * 4, hipMemcpyDeviceToHost, s));
  }
  std::vector<double> st(2 * ((size_t)iters + 1));
  if ((dh || lh) && iters > 0)
    EG_HIP(hipMemcpyAsync(st.data(), g->pr_stats, st.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  for (int it = 0; it < iters; it++) {
    if (dh) dh[it] = st[2 * it];
    if (lh) lh[it] = st[2 * (it + 1) + 1];
  }
  if (rank != g->rank_a) {
    std::swap(g->rank_a, g->rank_b);
    std::swap(g->contrib_a, g->contrib_b);
  }
  return EG_OK;
}

extern "C" int eg_pagerank(eg_graph *g, const eg_pr_opts *o, float *rank_out, double *dh,
                           double *lh) {
  EG_TRY(eg_require_loaded(g));
  double damping, init;
  int iters, uniform;
  eg_pr_resolve(g, o, &damping, &iters, &init, &uniform);
  if (iters < 0 || !(damping >= 0.0 && damping <= 1.0))
    return eg_fail(EG_EINVAL, "bad PageRank options");
  const int layout = o ? o->layout : EG_LAYOUT_AUTO;
  const bool adj = g->out_ptr != nullptr;
  if (layout == EG_LAYOUT_GRID || (layout == EG_LAYOUT_AUTO && g->cell_ptr))
    return eg_pagerank_grid(g, o, rank_out, dh, lh);
  if (layout == EG_LAYOUT_EDGES || (layout == EG_LAYOUT_AUTO && g->el_src))
    return eg_pagerank_edges(g, o, rank_out, dh, lh);
  if (!adj) return eg_fail(EG_ENOTSUP, "load mode %d has no adjacency lists", g->load_mode);

  const uint32_t *ptr, *idx, *deg = g->out_deg, *unperm = nullptr;
  const float *w;
  PullPlan *pl;
  if (eg_use_order(g, iters)) {
    EgOrdered *o;
    EG_TRY(eg_ordered(g, &o));
    ptr = o->ptr;
    idx = o->idx;
    w = o->w;
    pl = &o->plan;
    deg = o->deg;
    unperm = o->new_of_old;
  } else {
    EG_TRY(eg_pull_orientation(g, &ptr, &idx, &w, &pl));
  }
  EG_TRY(eg_pr_state(g, iters));
  const uint32_t n = g->n;
  // Binned form: every row sum is at most the total rank mass, which never
  // exceeds max(init * n, 1) (each iteration maps mass R to at most
  // (1 - d) + d R); the margin covers the FP32 rounding of the iterates.
  const bool binned = eg_pb_selected(g, iters, ptr);
  int S = 0;
  if (binned) {
    PbSource src;
    src.ptr = ptr;
    src.idx = idx;
    src.w = w;  // kept for SpMV over the same plan
    EG_TRY(eg_pb_build(g, src));
    S = eg_pb_scale(std::max(init * (double)n, 1.0) * 1.001);
  }
  float *rank = g->rank_a, *rank_next = g->rank_b;
  float *contrib = g->contrib_a, *contrib_next = g->contrib_b;
  EG_TRY(eg_pr_start(g, init, deg, rank, contrib));
  PrArgs a;
  a.out_deg = deg;
  a.base_vec = nullptr;
  a.base = (float)((1.0 - damping) / (double)n);
  a.damping = (float)damping;
  a.damping_over_n = damping / (double)n;
  a.uniform = uniform;
  for (int it = 0; it < iters; it++) {
    a.rank_old = rank;
    a.contrib_new = contrib_next;
    a.dmass_in = g->pr_stats + 2 * it;
    a.stats_out = g->pr_stats + 2 * (it + 1);
    if (binned)
      EG_TRY(eg_pb_iterate(g, contrib, rank_next, a, S, nullptr, !uniform && it >= 2));
    else
      EG_TRY(launch_pull(g, true, ptr, idx, nullptr, contrib, rank_next, *pl, a));
    std::swap(rank, rank_next);
    std::swap(contrib, contrib_next);
  }
  return eg_pr_finish(g, iters, rank, rank_next, rank_out, dh, lh, unperm);
}

// ---------------------------------------------------------------------------
// Single-vector personalized PageRank on the PageRank pull kernel: x0 = p,
// x' = alpha p + (1 - alpha) A^T D^-1 x with p uniform over the distinct
// seeds, iterated until the FP64 L1 change is <= tol or max_iter iterations.
// ---------------------------------------------------------------------------
namespace {
__global__ void k_ppr1_clear(uint32_t n, float *__restrict__ rank, float *__restrict__ contrib,
                             float *__restrict__ base) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  rank[v] = 0.f;
  contrib[v] = 0.f;
  base[v] = 0.f;
}

__global__ void k_ppr1_seed(const uint32_t *__restrict__ seeds, uint32_t ns,
                            const uint32_t *__restrict__ perm, const uint32_t *__restrict__ deg,
                            float w, float aw, float *__restrict__ rank,
                            float *__restrict__ contrib, float *__restrict__ base) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const uint32_t v = perm ? perm[seeds[i]] : seeds[i];
  rank[v] = w;
  contrib[v] = deg[v] ? w / (float)deg[v] : 0.f;
  base[v] = aw;
}
}  // namespace

// The orientation, vectors and (binned) plan of a single-vector query: built
// on first use, so eg_ppr1_prepare lets a caller pay for them up front.
static int ppr1_view(eg_graph *g, Ppr1 *st, int max_iter) {
  st->deg = g->out_deg;
  st->unperm = nullptr;
  if (eg_use_order(g, max_iter)) {
    EgOrdered *o;
    EG_TRY(eg_ordered(g, &o));
    st->ptr = o->ptr;
    st->idx = o->idx;
    st->w = o->w;
    st->pl = &o->plan;
    st->deg = o->deg;
    st->unperm = o->new_of_old;
  } else {
    EG_TRY(eg_pull_orientation(g, &st->ptr, &st->idx, &st->w, &st->pl));
  }
  EG_TRY(eg_pr_state(g, max_iter));
  // the iterate is a probability vector: row sums stay below 1
  st->binned = eg_pb_selected(g, max_iter, st->ptr);
  if (st->binned) {
    PbSource src;
    src.ptr = st->ptr;
    src.idx = st->idx;
    src.w = st->w;
    EG_TRY(eg_pb_build(g, src));
    st->scale = eg_pb_scale(1.001);
  }
  return EG_OK;
}

int eg_ppr1_prepare(eg_graph *g, int max_iter) {
  Ppr1 st;
  return ppr1_view(g, &st, max_iter);
}

int eg_ppr1_begin(eg_graph *g, Ppr1 *st, const uint32_t *seeds, uint32_t ns, double alpha,
                  int max_iter, hipStream_t stream, double *h_delta) {
  if (ns == 0 || max_iter < 0) return eg_fail(EG_EINVAL, "bad PPR query");
  EG_TRY(ppr1_view(g, st, max_iter));
  const uint32_t n = g->n;
  const uint32_t *perm = st->unperm;
  hipStream_t s = stream;
  st->stream = s;
  st->h_delta = h_delta;
  st->alpha = alpha;
  st->it = 0;
  st->max_iter = max_iter;
  if (!g->ppr_base) EG_TRY(eg_dalloc_t(&g->ppr_base, n, "ppr base"));
  if (g->ppr_seeds_cap < ns) {
    EG_HIP(hipStreamSynchronize(s));
    eg_dfree_t(&g->ppr_seeds);
    EG_TRY(eg_dalloc_t(&g->ppr_seeds, ns, "ppr seeds"));
    g->ppr_seeds_cap = ns;
  }
  uint32_t distinct = 0;
  for (uint32_t j = 0; j < ns; j++) {
    bool dup = false;
    for (uint32_t i = 0; i < j && !dup; i++) dup = seeds[i] == seeds[j];
    distinct += !dup;
  }
  st->rank = g->rank_a;
  st->rank_next = g->rank_b;
  st->contrib = g->contrib_a;
  st->contrib_next = g->contrib_b;
  hipLaunchKernelGGL(k_ppr1_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, st->rank,
                     st->contrib, g->ppr_base);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(g->ppr_seeds, seeds, (size_t)ns * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_ppr1_seed, dim3(blocks_for(ns, 64)), dim3(64), 0, s, g->ppr_seeds, ns, perm,
                     st->deg, (float)(1.0 / distinct), (float)(alpha / distinct), st->rank,
                     st->contrib, g->ppr_base);
  EG_LAUNCHED();
  EG_HIP(hipMemsetAsync(g->pr_stats, 0, 2 * sizeof(double), s));
  EG_HIP(hipStreamSynchronize(s));  // seeds come from the caller's memory
  return EG_OK;
}

int eg_ppr1_step(eg_graph *g, Ppr1 *st) {
  if (st->it >= st->max_iter) return eg_fail(EG_ESTATE, "PPR query already at its iteration cap");
  PrArgs a;
  a.out_deg = st->deg;
  a.base = 0.f;
  a.base_vec = g->ppr_base;
  a.damping = (float)(1.0 - st->alpha);
  a.damping_over_n = 0.0;
  a.uniform = 0;
  a.rank_old = st->rank;
  a.contrib_new = st->contrib_next;
  a.dmass_in = g->pr_stats + 2 * st->it;
  a.stats_out = g->pr_stats + 2 * (st->it + 1);
  if (st->binned)
    EG_TRY(eg_pb_iterate(g, st->contrib, st->rank_next, a, st->scale, st->stream));
  else
    EG_TRY(launch_pull(g, true, st->ptr, st->idx, nullptr, st->contrib, st->rank_next, *st->pl, a,
                       st->stream));
  std::swap(st->rank, st->rank_next);
  std::swap(st->contrib, st->contrib_next);
  st->it++;
  EG_HIP(hipMemcpyAsync(st->h_delta, a.stats_out + 1, sizeof(double), hipMemcpyDeviceToHost,
                        st->stream));
  return EG_OK;
}

int eg_ppr1_undo(eg_graph *, Ppr1 *st) {
  if (st->it == 0) return eg_fail(EG_ESTATE, "no PPR iteration to take back");
  std::swap(st->rank, st->rank_next);
  std::swap(st->contrib, st->contrib_next);
  st->it--;
  return EG_OK;
}

int eg_ppr1_output(eg_graph *g, Ppr1 *st, float *out) {
  if (out) {
    const float *src = st->rank;
    if (st->unperm) {
      EG_TRY(eg_gather_vec(g, st->rank, st->unperm, st->rank_next, st->stream));
      src = st->rank_next;
    }
    EG_HIP(hipMemcpyAsync(out, src, (size_t)g->n * 4, hipMemcpyDeviceToHost, st->stream));
  }
  EG_HIP(hipStreamSynchronize(st->stream));
  return EG_OK;
}

int eg_ppr_single(eg_graph *g, const uint32_t *seeds, uint32_t ns, double alpha, double tol,
                  int max_iter, float *out, int *iters_out, double *delta_out) {
  Ppr1 st;
  double *h_delta = reinterpret_cast<double *>(g->h_ctr);
  EG_TRY(eg_ppr1_begin(g, &st, seeds, ns, alpha, max_iter, g->stream, h_delta));
  double delta = 0.0;
  while (st.it < max_iter) {
    EG_TRY(eg_ppr1_step(g, &st));
    EG_HIP(hipStreamSynchronize(g->stream));
    delta = *h_delta;
    if (delta <= tol) break;
  }
  EG_TRY(eg_ppr1_output(g, &st, out));
  if (iters_out) *iters_out = st.it;
  if (delta_out) *delta_out = delta;
  return EG_OK;
}

extern "C" int eg_ppr(eg_graph *g, const uint32_t *seeds, uint32_t nseeds, double alpha,
                      double tol, int32_t max_iter, float *out, int32_t *iterations,
                      double *delta) {
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "PPR needs adjacency lists (modes 2/3/4)");
  if (nseeds && !seeds) return eg_fail(EG_EINVAL, "null seeds");
  if (!(alpha >= 0.0 && alpha <= 1.0) || !(tol >= 0.0)) return eg_fail(EG_EINVAL, "bad PPR options");
  for (uint32_t i = 0; i < nseeds; i++)
    if (seeds[i] >= g->n) return eg_fail(EG_ERANGE, "seed %u >= n = %u", seeds[i], g->n);
  int it = 0;
  EG_TRY(eg_ppr_single(g, seeds, nseeds, alpha, tol, max_iter, out, &it, delta));
  if (iterations) *iterations = it;
  return EG_OK;
}

extern "C" int eg_spmv(eg_graph *g, int32_t layout, const float *x, float *y) {
  EG_TRY(eg_require_loaded(g));
  if (!x) return eg_fail(EG_EINVAL, "null x");
  if (layout == EG_LAYOUT_GRID || (layout == EG_LAYOUT_AUTO && g->cell_ptr))
    return eg_spmv_grid(g, x, y);
  if (layout == EG_LAYOUT_EDGES || (layout == EG_LAYOUT_AUTO && g->el_src))
    return eg_spmv_edges(g, x, y);
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "load mode %d has no adjacency lists", g->load_mode);
  const uint32_t *ptr, *idx;
  const float *w;
  PullPlan *pl;
  EgOrdered *o = nullptr;
  if (eg_use_order(g, 1)) {
    EG_TRY(eg_ordered(g, &o));
    ptr = o->ptr;
    idx = o->idx;
    w = o->w;
    pl = &o->plan;
  } else {
    EG_TRY(eg_pull_orientation(g, &ptr, &idx, &w, &pl));
  }
  EG_TRY(eg_pr_state(g, 0));
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  float *xin = g->contrib_a, *yout = g->rank_a;
  if (o) {
    EG_HIP(hipMemcpyAsync(g->contrib_b, x, (size_t)n * 4, hipMemcpyHostToDevice, s));
    EG_TRY(eg_gather_vec(g, g->contrib_b, o->old_of_new, g->contrib_a));
  } else {
    EG_HIP(hipMemcpyAsync(g->contrib_a, x, (size_t)n * 4, hipMemcpyHostToDevice, s));
  }
  // Binned kernels sum exactly; their scale needs a finite bound on |x|.
  double xmax = 0.0;
  bool finite = true;
  const bool sel = eg_pb_selected(g, 1, ptr);
  if (sel)
    for (uint32_t i = 0; i < n && finite; i++) {
      const double ax = fabs((double)x[i]);
      finite = std::isfinite(ax);
      xmax = std::max(xmax, ax);
    }
  bool binned = sel && finite;
  if (binned) {
    PbSource src;
    src.ptr = ptr;
    src.idx = idx;
    src.w = w;
    EG_TRY(eg_pb_build(g, src));
    binned = eg_pb_spmv_exact(g, xmax);
  }
  if (binned) {
    EG_TRY(eg_pb_spmv(g, xin, yout, xmax));
  } else {
    PrArgs a = {};
    EG_TRY(launch_pull(g, false, ptr, idx, w, xin, yout, *pl, a));
  }
  if (o) {
    EG_TRY(eg_gather_vec(g, g->rank_a, o->new_of_old, g->rank_b));
    yout = g->rank_b;
  }
  if (y) EG_HIP(hipMemcpyAsync(y, yout, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

// Device-resident SpMV pass used by the benchmark (x, y in HBM, in the ids of
// the pull copy in use).
int eg_spmv_device(eg_graph *g, const float *x, float *y) {
  const uint32_t *ptr, *idx;
  const float *w;
  PullPlan *pl;
  if (eg_reorder_enabled()) {
    EgOrdered *o;
    EG_TRY(eg_ordered(g, &o));
    ptr = o->ptr;
    idx = o->idx;
    w = o->w;
    pl = &o->plan;
  } else {
    EG_TRY(eg_pull_orientation(g, &ptr, &idx, &w, &pl));
  }
  PrArgs a = {};
  return launch_pull(g, false, ptr, idx, w, x, y, *pl, a);
}
