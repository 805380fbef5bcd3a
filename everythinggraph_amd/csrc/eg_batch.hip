// eg_batch.hip -- batched form: many BFS sources and many personalized
// PageRank vectors per sweep over the graph.  No reference counterpart; the
// per-query semantics are those of eg_bfs (bfs_simple.c levels) and of the
// pull PageRank recurrence (pagerank_simple.c:72-84) with a personalization
// vector.
//
// MS-BFS: up to 64 sources share one traversal.  Every vertex carries three
// 64-bit lane masks (seen, frontier, next frontier); a level is one pull
// sweep in which vertex v ORs the frontier masks of its in-neighbours and
// stops as soon as every lane it has not seen is covered.  Levels of newly
// reached (vertex, lane) pairs are written vertex-major as u16 (root = 1).
//
// PPR: k <= 64 vectors stored row-major (64 FP32 columns per vertex).  One
// wave owns a row: it loads 64 in-edge ids with one coalesced read, then for
// each broadcasts the id and every lane adds its column of the neighbour's
// contribution row (a coalesced 256-byte gather) with Neumaier compensation.
// The seed term is added from a per-vertex 64-bit seed mask.
#include <math.h>
#include <string.h>

#include "eg_internal.h"

namespace {

constexpr int kBlock = 256;

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

__global__ void k_ms_init(uint32_t n, uint64_t *__restrict__ seen, uint64_t *__restrict__ visit,
                          uint16_t *__restrict__ lvl) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    seen[i] = 0;
    visit[i] = 0;
  }
  if (i < (uint64_t)n * 16) reinterpret_cast<uint64_t *>(lvl)[i] = 0;  // 64 u16 = 16 u64
}

__global__ void k_ms_seed(const uint32_t *__restrict__ roots, uint32_t k,
                          uint64_t *__restrict__ seen, uint64_t *__restrict__ visit,
                          uint16_t *__restrict__ lvl) {
  const uint32_t i = threadIdx.x;
  if (i >= k) return;
  const uint32_t r = roots[i];
  atomicOr(reinterpret_cast<unsigned long long *>(&seen[r]), 1ull << i);
  atomicOr(reinterpret_cast<unsigned long long *>(&visit[r]), 1ull << i);
  lvl[(uint64_t)r * 64 + i] = 1;
}

__global__ __launch_bounds__(kBlock) void k_ms_pull(uint32_t n, const uint32_t *__restrict__ ptr,
                                                    const uint32_t *__restrict__ idx,
                                                    uint64_t *__restrict__ seen,
                                                    const uint64_t *__restrict__ visit,
                                                    uint64_t *__restrict__ nvisit,
                                                    uint16_t *__restrict__ lvl, uint64_t all,
                                                    uint16_t newlvl, uint32_t *__restrict__ cnt) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t nw = 0;
  if (v < n) {
    const uint64_t s = seen[v];
    if (s != all) {
      uint64_t acc = 0;
      const uint32_t k1 = ptr[v + 1];
      for (uint32_t k = ptr[v]; k < k1; k += 4) {
        // four neighbours per step, then the early-exit test
        uint32_t u[4];
#pragma unroll
        for (int t = 0; t < 4; t++) u[t] = k + t < k1 ? idx[k + t] : idx[k];
#pragma unroll
        for (int t = 0; t < 4; t++) acc |= visit[u[t]];
        if ((acc | s) == all) break;
      }
      nw = acc & ~s;
      if (nw) {
        seen[v] = s | nw;
        uint64_t b = nw;
        while (b) {
          const int i = __ffsll((unsigned long long)b) - 1;
          lvl[(uint64_t)v * 64 + i] = newlvl;
          b &= b - 1;
        }
      }
    }
    nvisit[v] = nw;
  }
  // "some vertex was reached": only the first waves to see the flag clear write it
  if (__ballot(nw != 0) && __lane_id() == 0 && __atomic_load_n(cnt, __ATOMIC_RELAXED) == 0u)
    atomicOr(cnt, 1u);
}

__global__ void k_ms_extract(uint32_t n, uint32_t k, const uint16_t *__restrict__ lvl,
                             uint32_t *__restrict__ out) {
  // out[i * n + v] = lvl[v * 64 + i]
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * k) return;
  const uint32_t i = (uint32_t)(t / n), v = (uint32_t)(t % n);
  out[t] = lvl[(uint64_t)v * 64 + i];
}

// ---- PPR ----
__global__ void k_ppr_seedmask(uint32_t k, const uint32_t *__restrict__ seed_ptr,
                               const uint32_t *__restrict__ seeds, uint64_t *__restrict__ mask) {
  const uint32_t q = threadIdx.x;
  if (q >= k) return;
  for (uint32_t j = seed_ptr[q]; j < seed_ptr[q + 1]; j++)
    atomicOr(reinterpret_cast<unsigned long long *>(&mask[seeds[j]]), 1ull << q);
}

// x0 = p: X[v][q] = 1/|S_q| for seeds, contributions X/outdeg.
__global__ void k_ppr_init(uint32_t n, const uint64_t *__restrict__ mask,
                           const float *__restrict__ inv_seeds, const uint32_t *__restrict__ deg,
                           float *__restrict__ X, float *__restrict__ C) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * 64) return;
  const uint32_t v = (uint32_t)(t >> 6), q = (uint32_t)(t & 63);
  const float x = ((mask[v] >> q) & 1ull) ? inv_seeds[q] : 0.f;
  X[t] = x;
  const uint32_t d = deg[v];
  C[t] = d ? x / (float)d : 0.f;
}

// Adds the contribution rows of entries [k0, k1) to this lane's column.
__device__ __forceinline__ void spmm_range(const uint32_t *__restrict__ idx,
                                           const float *__restrict__ C, uint32_t k0, uint32_t k1,
                                           int lane, EgKahan &acc) {
  for (uint32_t base = k0; base < k1; base += 64) {
    const uint32_t cnt = k1 - base < 64 ? k1 - base : 64;
    const uint32_t my = lane < (int)cnt ? __builtin_nontemporal_load(idx + base + lane) : 0u;
    for (uint32_t j = 0; j < cnt; j += 8) {
      // eight independent 256-byte row gathers in flight per wave
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const uint32_t u = __shfl(my, (int)(j + t) & 63);
        v[t] = C[(uint64_t)u * 64 + lane];
      }
#pragma unroll
      for (int t = 0; t < 8; t++)
        if (j + t < cnt) acc.add(v[t]);
    }
  }
}

__device__ __forceinline__ void spmm_finish(uint32_t v, float sum, int lane,
                                            const uint64_t *__restrict__ mask, float seed_term,
                                            const uint32_t *__restrict__ deg, float one_minus_a,
                                            float *__restrict__ Y, float *__restrict__ Cn) {
  float y = one_minus_a * sum;
  if ((mask[v] >> lane) & 1ull) y += seed_term;
  Y[(uint64_t)v * 64 + lane] = y;
  const uint32_t d = deg[v];
  Cn[(uint64_t)v * 64 + lane] = d ? y / (float)d : 0.f;
}

// Work items of the pull plan: segments of long rows (the four waves split
// the segment, partials merged in wave order) and chunks of short rows (a
// wave per row).  Lane q owns query column q.
__global__ __launch_bounds__(kBlock) void k_ppr_spmm(
    const uint32_t *__restrict__ idx, const uint4 *__restrict__ chunk_info,
    const uint4 *__restrict__ seg_info, uint32_t nseg, float2 *__restrict__ seg_part,
    const uint32_t *__restrict__ ptr, const float *__restrict__ C,
    const uint64_t *__restrict__ mask, const float *__restrict__ alpha_inv_seeds,
    const uint32_t *__restrict__ deg, float one_minus_a, float *__restrict__ Y,
    float *__restrict__ Cn) {
  __shared__ float2 s_p[kBlock / 64][64];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  if (blockIdx.x < nseg) {
    const uint4 si = seg_info[blockIdx.x];
    const uint32_t len = si.y - si.x;
    const uint32_t part = (len + 4 * 64 - 1) / (4 * 64) * 64;  // multiple of 64 per wave
    const uint32_t a = si.x + wave * part < si.y ? si.x + wave * part : si.y;
    const uint32_t b = a + part < si.y ? a + part : si.y;
    EgKahan acc;
    spmm_range(idx, C, a, b, lane, acc);
    s_p[wave][lane] = make_float2(acc.s, acc.c);
    __syncthreads();
    if (wave == 0) {
      EgKahan t;
      t.s = s_p[0][lane].x;
      t.c = s_p[0][lane].y;
      for (int q = 1; q < kBlock / 64; q++) t.merge(s_p[q][lane].x, s_p[q][lane].y);
      seg_part[(uint64_t)blockIdx.x * 64 + lane] = make_float2(t.s, t.c);
    }
    return;
  }
  const uint4 ci = chunk_info[blockIdx.x - nseg];
  const float seed_term = alpha_inv_seeds[lane];
  for (uint32_t v = ci.x + wave; v < ci.y; v += kBlock / 64) {
    EgKahan acc;
    spmm_range(idx, C, ptr[v], ptr[v + 1], lane, acc);
    spmm_finish(v, acc.value(), lane, mask, seed_term, deg, one_minus_a, Y, Cn);
  }
}

// Long rows: a wave per row merges its segment partials in segment order.
__global__ __launch_bounds__(kBlock) void k_ppr_finish(
    const uint32_t *__restrict__ long_rows, const uint32_t *__restrict__ seg_base, uint32_t nlong,
    const float2 *__restrict__ seg_part, const uint64_t *__restrict__ mask,
    const float *__restrict__ alpha_inv_seeds, const uint32_t *__restrict__ deg,
    float one_minus_a, float *__restrict__ Y, float *__restrict__ Cn) {
  const int lane = __lane_id();
  const uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (j >= nlong) return;
  EgKahan acc;
  for (uint32_t s = seg_base[j]; s < seg_base[j + 1]; s++) {
    const float2 p = seg_part[(uint64_t)s * 64 + lane];
    acc.merge(p.x, p.y);
  }
  spmm_finish(long_rows[j], acc.value(), lane, mask, alpha_inv_seeds[lane], deg, one_minus_a, Y,
              Cn);
}

__global__ void k_ppr_extract(uint32_t n, uint32_t k, const float *__restrict__ X,
                              float *__restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * k) return;
  const uint32_t q = (uint32_t)(t / n), v = (uint32_t)(t % n);
  out[t] = X[(uint64_t)v * 64 + q];
}

}  // namespace

struct BatchState {
  uint64_t *seen = nullptr, *visit = nullptr, *nvisit = nullptr;
  uint16_t *lvl = nullptr;
  uint32_t *roots = nullptr;
  uint32_t *ext = nullptr;  // k x n extraction buffer
  uint64_t ext_cap = 0;
  float *X = nullptr, *Y = nullptr, *C = nullptr, *Cn = nullptr;
  uint64_t *mask = nullptr;
  float *inv = nullptr;
  uint32_t *seed_ptr = nullptr, *seeds = nullptr;
  uint32_t seeds_cap = 0;
  uint32_t *cnt = nullptr;
  float2 *seg_part = nullptr;  // [nseg * 64] long-row segment partials
  uint64_t seg_part_cap = 0;
};

void eg_batch_free(eg_graph *g) {
  BatchState *b = g->batch;
  if (!b) return;
  eg_dfree_t(&b->seen);
  eg_dfree_t(&b->visit);
  eg_dfree_t(&b->nvisit);
  eg_dfree_t(&b->lvl);
  eg_dfree_t(&b->roots);
  eg_dfree_t(&b->ext);
  eg_dfree_t(&b->X);
  eg_dfree_t(&b->Y);
  eg_dfree_t(&b->C);
  eg_dfree_t(&b->Cn);
  eg_dfree_t(&b->mask);
  eg_dfree_t(&b->inv);
  eg_dfree_t(&b->seed_ptr);
  eg_dfree_t(&b->seeds);
  eg_dfree_t(&b->cnt);
  eg_dfree_t(&b->seg_part);
  delete b;
  g->batch = nullptr;
}

static BatchState *batch_state(eg_graph *g) {
  if (!g->batch) g->batch = new BatchState();
  return g->batch;
}

static void in_orientation(eg_graph *g, const uint32_t **ptr, const uint32_t **idx) {
  if (g->in_ptr) {
    *ptr = g->in_ptr;
    *idx = g->in_idx;
  } else {
    *ptr = g->out_ptr;
    *idx = g->out_col;
  }
}

// One group of up to 64 sources; levels stay in b->lvl (vertex-major).
int eg_msbfs_group(eg_graph *g, uint32_t k, const uint32_t *d_roots, int *levels_out) {
  BatchState *b = batch_state(g);
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  if (!b->seen) {
    EG_TRY(eg_dalloc_t(&b->seen, n, "ms seen"));
    EG_TRY(eg_dalloc_t(&b->visit, n, "ms visit"));
    EG_TRY(eg_dalloc_t(&b->nvisit, n, "ms next"));
    EG_TRY(eg_dalloc_t(&b->lvl, (uint64_t)n * 64, "ms levels"));
    EG_TRY(eg_dalloc_t(&b->cnt, 16, "ms counters"));
  }
  const uint32_t *ptr, *idx;
  in_orientation(g, &ptr, &idx);
  hipLaunchKernelGGL(k_ms_init, dim3(blocks_for((uint64_t)n * 16, 256)), dim3(256), 0, s, n,
                     b->seen, b->visit, b->lvl);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_ms_seed, dim3(1), dim3(64), 0, s, d_roots, k, b->seen, b->visit, b->lvl);
  EG_LAUNCHED();
  const uint64_t all = k == 64 ? ~0ull : ((1ull << k) - 1ull);
  int levels = 1;
  for (;;) {
    if (levels >= 65534 || (uint64_t)levels > (uint64_t)n + 1)
      return eg_fail(EG_ENOTSUP, "BFS depth exceeds the u16 level store");
    EG_HIP(hipMemsetAsync(b->cnt, 0, 4, s));
    {
      EgProf prof(g, "msbfs_pull");
      hipLaunchKernelGGL(k_ms_pull, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, n, ptr, idx,
                         b->seen, b->visit, b->nvisit, b->lvl, all, (uint16_t)(levels + 1),
                         b->cnt);
      EG_LAUNCHED();
    }
    std::swap(b->visit, b->nvisit);
    EG_HIP(hipMemcpyAsync(g->h_ctr + 32, b->cnt, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    if (g->h_ctr[32] == 0) break;
    levels++;
  }
  if (levels_out) *levels_out = levels;
  return EG_OK;
}

extern "C" int eg_bfs_batch(eg_graph *g, uint32_t k, const uint32_t *roots, uint32_t *dist,
                            double *device_ms) {
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "batched BFS needs adjacency lists (modes 2/3/4)");
  if (k && !roots) return eg_fail(EG_EINVAL, "null roots");
  for (uint32_t i = 0; i < k; i++)
    if (roots[i] >= g->n) return eg_fail(EG_ERANGE, "root %u >= n", roots[i]);
  BatchState *b = batch_state(g);
  hipStream_t s = g->stream;
  const uint32_t n = g->n;
  if (!b->roots) EG_TRY(eg_dalloc_t(&b->roots, 64, "roots"));
  if (dist && b->ext_cap < (uint64_t)n * 64) {
    eg_dfree_t(&b->ext);
    EG_TRY(eg_dalloc_t(&b->ext, (uint64_t)n * 64, "extract"));
    b->ext_cap = (uint64_t)n * 64;
  }
  double total_ms = 0.0;
  for (uint32_t g0 = 0; g0 < k; g0 += 64) {
    const uint32_t kk = k - g0 < 64 ? k - g0 : 64;
    EG_HIP(hipMemcpyAsync(b->roots, roots + g0, kk * 4, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    EG_HIP(hipEventCreate(&e0));
    EG_HIP(hipEventCreate(&e1));
    EG_HIP(hipEventRecord(e0, s));
    int lv = 0;
    int rc = eg_msbfs_group(g, kk, b->roots, &lv);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    total_ms += ms;
    if (dist) {
      hipLaunchKernelGGL(k_ms_extract, dim3(blocks_for((uint64_t)n * kk, 256)), dim3(256), 0, s,
                         n, kk, b->lvl, b->ext);
      EG_LAUNCHED();
      EG_HIP(hipMemcpyAsync(dist + (uint64_t)g0 * n, b->ext, (uint64_t)n * kk * 4,
                            hipMemcpyDeviceToHost, s));
      EG_HIP(hipStreamSynchronize(s));
    }
  }
  if (device_ms) *device_ms = total_ms;
  return EG_OK;
}

// k <= 64 queries; results stay in b->X (row-major, 64 columns).
int eg_ppr_group(eg_graph *g, uint32_t k, const uint32_t *seed_ptr, const uint32_t *seeds,
                 double alpha, int iterations) {
  BatchState *b = batch_state(g);
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  if (!b->X) {
    EG_TRY(eg_dalloc_t(&b->X, (uint64_t)n * 64, "ppr X"));
    EG_TRY(eg_dalloc_t(&b->Y, (uint64_t)n * 64, "ppr Y"));
    EG_TRY(eg_dalloc_t(&b->C, (uint64_t)n * 64, "ppr C"));
    EG_TRY(eg_dalloc_t(&b->Cn, (uint64_t)n * 64, "ppr Cn"));
    EG_TRY(eg_dalloc_t(&b->mask, n, "ppr seed mask"));
    EG_TRY(eg_dalloc_t(&b->inv, 128, "ppr seed weights"));
    EG_TRY(eg_dalloc_t(&b->seed_ptr, 65, "seed ptr"));
  }
  const uint32_t nseeds = seed_ptr[k];
  if (b->seeds_cap < nseeds + 1) {
    eg_dfree_t(&b->seeds);
    EG_TRY(eg_dalloc_t(&b->seeds, (uint64_t)nseeds + 1, "seeds"));
    b->seeds_cap = nseeds + 1;
  }
  // 1/|distinct seeds| per query, and alpha times it.
  float inv[128];
  memset(inv, 0, sizeof(inv));
  for (uint32_t q = 0; q < k; q++) {
    uint32_t distinct = 0;
    for (uint32_t j = seed_ptr[q]; j < seed_ptr[q + 1]; j++) {
      bool dup = false;
      for (uint32_t i = seed_ptr[q]; i < j; i++) dup |= seeds[i] == seeds[j];
      distinct += !dup;
    }
    if (distinct == 0) return eg_fail(EG_EINVAL, "query %u has no seeds", q);
    inv[q] = (float)(1.0 / distinct);
    inv[64 + q] = (float)(alpha / distinct);
  }
  std::vector<uint32_t> sp(seed_ptr, seed_ptr + k + 1);
  for (auto &x : sp) x -= seed_ptr[0];
  EG_HIP(hipMemcpyAsync(b->inv, inv, sizeof(inv), hipMemcpyHostToDevice, s));
  EG_HIP(hipMemcpyAsync(b->seed_ptr, sp.data(), (k + 1) * 4, hipMemcpyHostToDevice, s));
  EG_HIP(hipMemcpyAsync(b->seeds, seeds + seed_ptr[0], (uint64_t)sp[k] * 4, hipMemcpyHostToDevice,
                        s));
  EG_HIP(hipMemsetAsync(b->mask, 0, (size_t)n * 8, s));
  hipLaunchKernelGGL(k_ppr_seedmask, dim3(1), dim3(64), 0, s, k, b->seed_ptr, b->seeds, b->mask);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_ppr_init, dim3(blocks_for((uint64_t)n * 64, 256)), dim3(256), 0, s, n,
                     b->mask, b->inv, g->out_deg, b->X, b->C);
  EG_LAUNCHED();
  const uint32_t *ptr, *idx;
  const float *w;
  PullPlan *pl;
  EG_TRY(eg_pull_orientation(g, &ptr, &idx, &w, &pl));
  if (b->seg_part_cap < (uint64_t)pl->nseg * 64) {
    eg_dfree_t(&b->seg_part);
    EG_TRY(eg_dalloc_t(&b->seg_part, (uint64_t)(pl->nseg ? pl->nseg : 1) * 64, "ppr partials"));
    b->seg_part_cap = (uint64_t)(pl->nseg ? pl->nseg : 1) * 64;
  }
  const float oma = (float)(1.0 - alpha);
  for (int it = 0; it < iterations; it++) {
    {
      EgProf prof(g, "ppr_spmm");
      hipLaunchKernelGGL(k_ppr_spmm, dim3(pl->nseg + pl->nchunks), dim3(kBlock), 0, s, idx,
                         pl->chunk_info, pl->seg_info, pl->nseg, b->seg_part, ptr, b->C, b->mask,
                         b->inv + 64, g->out_deg, oma, b->Y, b->Cn);
      EG_LAUNCHED();
    }
    if (pl->nlong) {
      hipLaunchKernelGGL(k_ppr_finish, dim3(blocks_for((uint64_t)pl->nlong * 64, kBlock)),
                         dim3(kBlock), 0, s, pl->long_rows, pl->seg_base, pl->nlong, b->seg_part,
                         b->mask, b->inv + 64, g->out_deg, oma, b->Y, b->Cn);
      EG_LAUNCHED();
    }
    std::swap(b->X, b->Y);
    std::swap(b->C, b->Cn);
  }
  return EG_OK;
}

extern "C" int eg_ppr_batch(eg_graph *g, uint32_t k, const uint32_t *seed_ptr,
                            const uint32_t *seeds, double alpha, int32_t iterations, float *out,
                            double *device_ms) {
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "batched PPR needs adjacency lists (modes 2/3/4)");
  if (k && (!seed_ptr || !seeds)) return eg_fail(EG_EINVAL, "null seeds");
  if (!(alpha >= 0.0 && alpha <= 1.0) || iterations < 0)
    return eg_fail(EG_EINVAL, "bad PPR options");
  for (uint32_t q = 0; q < k; q++)
    for (uint32_t j = seed_ptr[q]; j < seed_ptr[q + 1]; j++)
      if (seeds[j] >= g->n) return eg_fail(EG_ERANGE, "seed %u >= n", seeds[j]);
  BatchState *b = batch_state(g);
  hipStream_t s = g->stream;
  const uint32_t n = g->n;
  double total_ms = 0.0;
  for (uint32_t q0 = 0; q0 < k; q0 += 64) {
    const uint32_t kk = k - q0 < 64 ? k - q0 : 64;
    hipEvent_t e0, e1;
    EG_HIP(hipEventCreate(&e0));
    EG_HIP(hipEventCreate(&e1));
    EG_HIP(hipEventRecord(e0, s));
    int rc = eg_ppr_group(g, kk, seed_ptr + q0, seeds, alpha, iterations);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    total_ms += ms;
    if (out) {
      if (b->ext_cap < (uint64_t)n * 64) {
        eg_dfree_t(&b->ext);
        EG_TRY(eg_dalloc_t(&b->ext, (uint64_t)n * 64, "extract"));
        b->ext_cap = (uint64_t)n * 64;
      }
      float *ext = reinterpret_cast<float *>(b->ext);
      hipLaunchKernelGGL(k_ppr_extract, dim3(blocks_for((uint64_t)n * kk, 256)), dim3(256), 0, s,
                         n, kk, b->X, ext);
      EG_LAUNCHED();
      EG_HIP(hipMemcpyAsync(out + (uint64_t)q0 * n, ext, (uint64_t)n * kk * 4,
                            hipMemcpyDeviceToHost, s));
      EG_HIP(hipStreamSynchronize(s));
    }
  }
  if (device_ms) *device_ms = total_ms;
  return EG_OK;
}
