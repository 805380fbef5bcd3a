// eg_prims.hip -- device primitives of the layout builders: exclusive scan,
// vertex histogram, stable LSD radix sort, record unpacking and the synthetic
// generator.
//
// The reference builds every layout with Ligra's stable radix sort
// (intSort::iSort, radixSort_ligra.h:355) and sequential/blocked scans
// (utils_ligra.h:185-214).  Here the same contracts (stable by key; exclusive
// prefix sums) are met with wave64 kernels: ballot-based stable ranking inside
// a tile, LDS staging so that every tile writes contiguous runs, and a small
// per-range count matrix scanned on the device.
#include <stdarg.h>
#include <stdio.h>

#include "eg_internal.h"

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;  // per thread; a wave scans 1024 items
constexpr uint32_t kScanTile = kScanBlock * kScanItems;

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// Per-tile sums.
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(
    const uint32_t *__restrict__ in, uint64_t count, uint32_t *__restrict__ part) {
  __shared__ uint32_t s_w[kScanBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t acc = 0;
  for (uint32_t j = threadIdx.x; j < kScanTile; j += kScanBlock) {
    const uint64_t i = base + j;
    if (i < count) acc += in[i];
  }
  acc = eg_wave_sum_u32(acc);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kScanBlock / 64; w++) t += s_w[w];
    part[blockIdx.x] = t;
  }
}

// Exclusive scan of one tile plus the scanned tile offset.  Wave w scans the
// contiguous items [w*1024, (w+1)*1024) of the tile in 16 steps of 64.
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(
    const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t count,
    const uint32_t *__restrict__ part_scanned) {
  __shared__ uint32_t s_w[kScanBlock / 64];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)wave * 64 * kScanItems;
  uint32_t v[kScanItems];
  uint32_t carry = 0;
#pragma unroll
  for (int s = 0; s < kScanItems; s++) {
    const uint64_t i = base + (uint64_t)s * 64 + lane;
    const uint32_t x = i < count ? in[i] : 0u;
    const uint32_t inc = wave_incl_scan(x);
    v[s] = carry + inc - x;
    carry += __shfl(inc, 63);
  }
  if (lane == 0) s_w[wave] = carry;
  __syncthreads();
  uint32_t off = part_scanned ? part_scanned[blockIdx.x] : 0u;
  for (int w = 0; w < wave; w++) off += s_w[w];
#pragma unroll
  for (int s = 0; s < kScanItems; s++) {
    const uint64_t i = base + (uint64_t)s * 64 + lane;
    if (i < count) out[i] = off + v[s];
  }
}

__global__ void k_store_total(const uint32_t *__restrict__ last_in,
                              const uint32_t *__restrict__ last_out,
                              uint32_t *__restrict__ total) {
  // total = exclusive[count-1] + in[count-1]; written by one thread after the
  // apply kernel (in may alias out, so the input value was saved first).
  *total = *last_out + *last_in;
}

__global__ void k_copy_u32(const uint32_t *__restrict__ a, uint32_t *__restrict__ b) {
  *b = *a;
}

// ---------------------------------------------------------------------------
// Vertex histogram with LDS privatisation of the low ids, where R-MAT hubs
// live; higher ids go straight to global atomics.
// ---------------------------------------------------------------------------
constexpr int kHistBlock = 256;
constexpr int kHistItems = 32;
constexpr uint32_t kHistLds = 4096;

__global__ __launch_bounds__(kHistBlock) void k_histogram(
    const uint32_t *__restrict__ keys, uint64_t count, uint32_t *__restrict__ deg,
    uint32_t nbins) {
  __shared__ uint32_t s_h[kHistLds];
  const uint32_t lds_bins = nbins < kHistLds ? nbins : kHistLds;
  for (uint32_t j = threadIdx.x; j < lds_bins; j += kHistBlock) s_h[j] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kHistBlock * kHistItems;
#pragma unroll 4
  for (int s = 0; s < kHistItems; s++) {
    const uint64_t i = base + (uint64_t)s * kHistBlock + threadIdx.x;
    if (i < count) {
      const uint32_t k = keys[i];
      if (k < lds_bins)
        atomicAdd(&s_h[k], 1u);
      else
        atomicAdd(&deg[k], 1u);
    }
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < lds_bins; j += kHistBlock)
    if (s_h[j]) atomicAdd(&deg[j], s_h[j]);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, range-blocked.  The items are cut into at most
// kRbMaxRanges contiguous ranges of whole tiles (kRbTile items).  One pass =
//  * k_rb_hist: the digit histogram of every range (one read of the keys),
//    laid out digit-major, range-minor;
//  * an exclusive scan of that small matrix: the start of every (digit,
//    range) run in the output;
//  * k_rb_scatter: one workgroup per range walks its tiles in order.  A tile
//    is ranked stably (lanes with equal digits find each other with one
//    ballot per digit bit, and a per-wave LDS row carries the running
//    counts), reordered through LDS a column at a time so that each digit's
//    run is written contiguously, and written at the range's running digit
//    offsets, which then advance by the tile's counts.
// No workgroup waits on another; stability follows from the range order, the
// tile order inside a range and the rank order inside a tile.
// ---------------------------------------------------------------------------
constexpr int kRbBlock = 512;
constexpr int kRbItems = 16;
constexpr uint32_t kRbTile = kRbBlock * kRbItems;  // 8192
constexpr int kRbWaves = kRbBlock / 64;
constexpr uint32_t kRbMaxRanges = 2048;
constexpr int kRbHistUnroll = 8;

// Digit histogram of range r = [r * span, min((r + 1) * span, count)) into
// counts[d * nranges + r]; with limit != 0 also flags keys >= limit.
__global__ __launch_bounds__(kRbBlock) void k_rb_hist(
    const uint32_t *__restrict__ keys, uint64_t count, uint64_t span, int shift, uint32_t mask,
    uint32_t nranges, uint32_t *__restrict__ counts, uint32_t limit, uint32_t *__restrict__ flag) {
  __shared__ uint32_t s_h[kRbWaves][256];
  const int wave = threadIdx.x >> 6;
  for (int j = threadIdx.x; j < kRbWaves * 256; j += kRbBlock) (&s_h[0][0])[j] = 0;
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * span;
  const uint64_t hi = lo + span < count ? lo + span : count;
  bool bad = false;
  for (uint64_t base = lo; base < hi; base += (uint64_t)kRbBlock * kRbHistUnroll) {
    uint32_t k[kRbHistUnroll];
#pragma unroll
    for (int u = 0; u < kRbHistUnroll; u++) {
      const uint64_t i = base + (uint64_t)u * kRbBlock + threadIdx.x;
      k[u] = i < hi ? keys[i] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < kRbHistUnroll; u++) {
      const uint64_t i = base + (uint64_t)u * kRbBlock + threadIdx.x;
      if (i < hi) {
        bad |= limit && k[u] >= limit;
        atomicAdd(&s_h[wave][(k[u] >> shift) & mask], 1u);
      }
    }
  }
  if (limit && __ballot(bad) && __lane_id() == 0) atomicOr(flag, 1u);
  __syncthreads();
  for (int d = threadIdx.x; d <= (int)mask; d += kRbBlock) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kRbWaves; w++) c += s_h[w][d];
    counts[(uint64_t)d * nranges + blockIdx.x] = c;
  }
}

// One range per workgroup: tiles in order, each ranked, reordered through LDS
// and written at the running offsets of its digits (offs: the scanned
// counts).  With limit != 0 payload column 0 is checked against it.
template <int NP, bool WRITE_KEYS>
__global__ __launch_bounds__(kRbBlock) void k_rb_scatter(
    const uint32_t *__restrict__ kin, uint32_t *__restrict__ kout,
    const uint32_t *__restrict__ p0i, uint32_t *__restrict__ p0o,
    const uint32_t *__restrict__ p1i, uint32_t *__restrict__ p1o,
    const uint32_t *__restrict__ p2i, uint32_t *__restrict__ p2o, uint64_t count,
    uint64_t span, int shift, int nbits, uint32_t nranges, const uint32_t *__restrict__ offs,
    uint32_t limit, uint32_t *__restrict__ flag) {
  __shared__ uint32_t s_wh[kRbWaves][256];
  __shared__ uint32_t s_tstart[256];
  __shared__ uint32_t s_gofs[256];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_buf[kRbTile];

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const uint32_t mask = (1u << nbits) - 1u;
  const uint32_t *pin[3] = {p0i, p1i, p2i};
  uint32_t *pout[3] = {p0o, p1o, p2o};
  const uint64_t lt = eg_lanemask_lt();
  const uint64_t lo = (uint64_t)blockIdx.x * span;
  const uint64_t hi = lo + span < count ? lo + span : count;
  if (tid < 256) s_gofs[tid] = tid <= (int)mask ? offs[(uint64_t)tid * nranges + blockIdx.x] : 0u;
  bool bad = false;

  for (uint64_t tbase = lo; tbase < hi; tbase += kRbTile) {
    for (int j = tid; j < kRbWaves * 256; j += kRbBlock) (&s_wh[0][0])[j] = 0;
    __syncthreads();
    uint32_t key[kRbItems], pos[kRbItems];
    uint32_t pay[NP > 0 ? NP : 1][kRbItems];
    const uint64_t wbase = tbase + (uint64_t)wave * 64 * kRbItems;
#pragma unroll
    for (int s = 0; s < kRbItems; s++) {
      const uint64_t i = wbase + (uint64_t)s * 64 + lane;
      const bool valid = i < hi;
      key[s] = valid ? kin[i] : 0u;
#pragma unroll
      for (int p = 0; p < NP; p++) pay[p][s] = valid ? pin[p][i] : 0u;
      if (NP > 0 && limit && valid) bad |= pay[0][s] >= limit;
    }
    // stable rank of each item among the wave's items of its digit
#pragma unroll
    for (int s = 0; s < kRbItems; s++) {
      const uint64_t i = wbase + (uint64_t)s * 64 + lane;
      const bool valid = i < hi;
      const uint32_t d = (key[s] >> shift) & mask;
      uint64_t peers = __ballot(valid);
      for (int b = 0; b < nbits; b++) {
        const int bit = (d >> b) & 1;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      const uint32_t r = __popcll(peers & lt);
      pos[s] = 0;
      if (valid) {
        const uint32_t base = s_wh[wave][d];
        pos[s] = base + r;
        if (r == 0) s_wh[wave][d] = base + (uint32_t)__popcll(peers);
      }
    }
    __syncthreads();
    // per digit (threads 0..255): prefix over waves and the tile's scan over digits
    uint32_t run = 0, excl = 0;
    if (tid < 256) {
#pragma unroll
      for (int w = 0; w < kRbWaves; w++) {
        const uint32_t c = s_wh[w][tid];
        s_wh[w][tid] = run;
        run += c;
      }
      const uint32_t inc = wave_incl_scan(run);
      excl = inc - run;
      if (lane == 63) s_wsum[wave] = inc;
    }
    __syncthreads();
    if (tid < 256) {
      uint32_t off = 0;
      for (int w = 0; w < wave; w++) off += s_wsum[w];
      s_tstart[tid] = off + excl;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kRbItems; s++) {
      const uint64_t i = wbase + (uint64_t)s * 64 + lane;
      if (i < hi) {
        const uint32_t d = (key[s] >> shift) & mask;
        pos[s] += s_tstart[d] + s_wh[wave][d];
        s_buf[pos[s]] = key[s];
      }
    }
    __syncthreads();
    const uint64_t rem = hi - tbase;
    const uint32_t nvalid = rem < kRbTile ? (uint32_t)rem : kRbTile;
    uint32_t dst[kRbItems];
#pragma unroll
    for (int r = 0; r < kRbItems; r++) {
      const uint32_t j = tid + (uint32_t)r * kRbBlock;
      dst[r] = 0;
      if (j < nvalid) {
        const uint32_t k = s_buf[j];
        const uint32_t d = (k >> shift) & mask;
        dst[r] = s_gofs[d] + (j - s_tstart[d]);
        if (WRITE_KEYS) kout[dst[r]] = k;
      }
    }
#pragma unroll
    for (int p = 0; p < NP; p++) {
      __syncthreads();
#pragma unroll
      for (int s = 0; s < kRbItems; s++) {
        const uint64_t i = wbase + (uint64_t)s * 64 + lane;
        if (i < hi) s_buf[pos[s]] = pay[p][s];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < kRbItems; r++) {
        const uint32_t j = tid + (uint32_t)r * kRbBlock;
        if (j < nvalid) pout[p][dst[r]] = s_buf[j];
      }
    }
    __syncthreads();  // every thread has read s_gofs / s_tstart / s_buf of this tile
    if (tid < 256) s_gofs[tid] += run;
  }
  if (limit && __ballot(bad) && lane == 0) atomicOr(flag, 1u);
}

template <int NP>
void launch_rb_scatter(hipStream_t s, bool wk, uint32_t nranges, const uint32_t *kin,
                       uint32_t *kout, const uint32_t *const *pi, uint32_t *const *po,
                       uint64_t count, uint64_t span, int shift, int nbits, const uint32_t *offs,
                       uint32_t limit, uint32_t *flag) {
  if (wk)
    hipLaunchKernelGGL((k_rb_scatter<NP, true>), dim3(nranges), dim3(kRbBlock), 0, s, kin, kout,
                       pi[0], po[0], pi[1], po[1], pi[2], po[2], count, span, shift, nbits,
                       nranges, offs, limit, flag);
  else
    hipLaunchKernelGGL((k_rb_scatter<NP, false>), dim3(nranges), dim3(kRbBlock), 0, s, kin, kout,
                       pi[0], po[0], pi[1], po[1], pi[2], po[2], count, span, shift, nbits,
                       nranges, offs, limit, flag);
}

// ---------------------------------------------------------------------------
// Generator (same streams as the oracle's ego_gen_rmat / ego_gen_er /
// ego_gen_weights).
// ---------------------------------------------------------------------------
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// Record kernels stride over the grid: counts reach past 2^32 (the threads of
// one launch do not).
__device__ __forceinline__ uint64_t first_index() {
  return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ uint64_t grid_threads() { return (uint64_t)gridDim.x * blockDim.x; }

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += kGolden;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t mix64_host(uint64_t z) { return mix64(z); }

__global__ void k_gen_rmat(uint64_t m_in, int scale, uint64_t skey, uint32_t ta,
                           uint32_t tab, uint32_t tabc, uint32_t *__restrict__ src,
                           uint32_t *__restrict__ dst) {
  for (uint64_t e = first_index(); e < m_in; e += grid_threads()) {
    const uint64_t ekey = mix64(skey ^ e);
    uint32_t s = 0, d = 0;
    uint64_t h = 0;
    for (int l = 0; l < scale; l++) {
      uint32_t r;
      if ((l & 1) == 0) {
        h = mix64(ekey + (uint64_t)(l / 2 + 1) * kGolden);
        r = (uint32_t)(h >> 32);
      } else {
        r = (uint32_t)h;
      }
      const uint32_t qs = r >= tab ? 1u : 0u;
      const uint32_t qd = (r >= ta && r < tab) || r >= tabc ? 1u : 0u;
      s = (s << 1) | qs;
      d = (d << 1) | qd;
    }
    src[e] = s;
    dst[e] = d;
  }
}

__global__ void k_gen_er(uint64_t m_in, uint32_t n, uint64_t skey,
                         uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
  for (uint64_t e = first_index(); e < m_in; e += grid_threads()) {
    const uint64_t h = mix64(mix64(skey ^ e) + kGolden);
    src[e] = (uint32_t)(((h >> 32) * (uint64_t)n) >> 32);
    dst[e] = (uint32_t)(((h & 0xFFFFFFFFull) * (uint64_t)n) >> 32);
  }
}

__global__ void k_gen_weights(uint64_t m_in, uint64_t wkey, float *__restrict__ w) {
  for (uint64_t e = first_index(); e < m_in; e += grid_threads()) {
    const uint64_t h = mix64(wkey ^ e);
    w[e] = (float)((h >> 40) + 1) * 0x1p-24f;
  }
}

__global__ void k_unpack(const uint32_t *__restrict__ rec, uint64_t m_in, int words,
                         uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                         float *__restrict__ w) {
  for (uint64_t e = first_index(); e < m_in; e += grid_threads()) {
    const uint32_t *r = rec + e * (uint64_t)words;
    src[e] = r[0];
    dst[e] = r[1];
    if (words == 3 && w) w[e] = __uint_as_float(r[2]);
  }
}

// preload_graph's reversed copies (init_all.c:102-106) at slot e + m_in.
__global__ void k_append_reverse(uint64_t m_in, uint32_t *__restrict__ src,
                                 uint32_t *__restrict__ dst, float *__restrict__ w) {
  for (uint64_t e = first_index(); e < m_in; e += grid_threads()) {
    src[e + m_in] = dst[e];
    dst[e + m_in] = src[e];
    if (w) w[e + m_in] = w[e];
  }
}

__global__ void k_check_ids(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                            uint64_t count, uint32_t n, uint32_t *__restrict__ flag) {
  bool bad = false;
  for (uint64_t i = first_index(); i < count; i += grid_threads()) bad |= a[i] >= n || b[i] >= n;
  if (__ballot(bad) && __lane_id() == 0) atomicOr(flag, 1u);
}

__global__ void k_cell_keys(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            uint64_t count, uint32_t n, uint32_t P,
                            uint32_t *__restrict__ keys) {
  for (uint64_t i = first_index(); i < count; i += grid_threads())
    keys[i] = eg_partition_id(src[i], n, P) * P + eg_partition_id(dst[i], n, P);
}

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

// Workgroups of 256 for a grid-stride record kernel over count items.
inline uint32_t record_grid(uint64_t count) {
  const uint64_t b = (count + 255) / 256;
  return b < 65536 ? (b ? (uint32_t)b : 1u) : 65536u;
}

// ptr[v] = first k with keys[k] >= v over ascending keys, v in [0, n]: the
// boundary detection + scanIBack(min) of init_all.c:675-691 as one binary
// search per vertex.  deg[v] = ptr[v+1] - ptr[v] when deg is given.
__global__ void k_ptr_from_sorted(const uint32_t *__restrict__ keys, uint64_t m, uint32_t n,
                                  uint32_t *__restrict__ ptr) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n) return;
  uint64_t lo = 0, hi = m;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (keys[mid] >= v) hi = mid; else lo = mid + 1;
  }
  ptr[v] = (uint32_t)lo;
}

__global__ void k_deg_from_ptr(const uint32_t *__restrict__ ptr, uint32_t n,
                               uint32_t *__restrict__ deg) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) deg[v] = ptr[v + 1] - ptr[v];
}

// Number of zero entries: per-workgroup counts, then one workgroup sums them.
constexpr uint32_t kCountBlocks = 1024;
__global__ __launch_bounds__(256) void k_count_zero(const uint32_t *__restrict__ a, uint32_t n,
                                                    uint32_t *__restrict__ part) {
  __shared__ uint32_t s_w[4];
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    c += a[i] == 0;
  c = eg_wave_sum_u32(c);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(256) void k_sum_u32(const uint32_t *__restrict__ part, uint32_t np,
                                                 uint32_t *__restrict__ out) {
  __shared__ uint32_t s_w[4];
  uint32_t c = 0;
  for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) c += part[i];
  c = eg_wave_sum_u32(c);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) *out = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

}  // namespace

int eg_ptr_from_sorted(hipStream_t s, const uint32_t *keys, uint64_t m, uint32_t n, uint32_t *ptr,
                       uint32_t *deg) {
  hipLaunchKernelGGL(k_ptr_from_sorted, dim3(blocks_for((uint64_t)n + 1, 256)), dim3(256), 0, s,
                     keys, m, n, ptr);
  EG_LAUNCHED();
  if (deg) {
    hipLaunchKernelGGL(k_deg_from_ptr, dim3(blocks_for(n, 256)), dim3(256), 0, s, ptr, n, deg);
    EG_LAUNCHED();
  }
  return EG_OK;
}

int eg_count_zero(hipStream_t s, const uint32_t *a, uint32_t n, uint32_t *d_tmp,
                  uint32_t *result) {
  const uint32_t nb = blocks_for(n, 256) < kCountBlocks ? blocks_for(n, 256) : kCountBlocks;
  hipLaunchKernelGGL(k_count_zero, dim3(nb ? nb : 1), dim3(256), 0, s, a, n, d_tmp);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_sum_u32, dim3(1), dim3(256), 0, s, d_tmp, nb ? nb : 1, d_tmp + kCountBlocks);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(result, d_tmp + kCountBlocks, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Host entry points.
// ---------------------------------------------------------------------------
size_t eg_scan_tmp_bytes(uint64_t count) {
  size_t bytes = 256;  // saved last input value
  uint64_t c = count;
  while (c > kScanTile) {
    c = (c + kScanTile - 1) / kScanTile;
    bytes += ((c + 1) * sizeof(uint32_t) + 255) / 256 * 256;
  }
  return bytes;
}

static int scan_level(hipStream_t s, const uint32_t *in, uint32_t *out,
                      uint64_t count, uint8_t *tmp) {
  if (count == 0) return EG_OK;
  const uint32_t tiles = blocks_for(count, kScanTile);
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kScanBlock), 0, s, in, out, count,
                       (const uint32_t *)nullptr);
    EG_LAUNCHED();
    return EG_OK;
  }
  uint32_t *part = reinterpret_cast<uint32_t *>(tmp);
  const size_t part_bytes = ((tiles + 1) * sizeof(uint32_t) + 255) / 256 * 256;
  hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kScanBlock), 0, s, in, count, part);
  EG_LAUNCHED();
  EG_TRY(scan_level(s, part, part, tiles, tmp + part_bytes));
  hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanBlock), 0, s, in, out, count,
                     (const uint32_t *)part);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_scan_excl(hipStream_t s, const uint32_t *in, uint32_t *out, uint64_t count,
                 void *tmp) {
  uint8_t *t = reinterpret_cast<uint8_t *>(tmp);
  uint32_t *saved = reinterpret_cast<uint32_t *>(t);
  if (count == 0) {
    EG_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), s));
    return EG_OK;
  }
  hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(1), 0, s, in + count - 1, saved);
  EG_LAUNCHED();
  EG_TRY(scan_level(s, in, out, count, t + 256));
  hipLaunchKernelGGL(k_store_total, dim3(1), dim3(1), 0, s, saved, out + count - 1,
                     out + count);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_histogram(hipStream_t s, const uint32_t *keys, uint64_t count, uint32_t *deg,
                 uint32_t nbins) {
  if (count == 0) return EG_OK;
  const uint32_t blocks = blocks_for(count, (uint64_t)kHistBlock * kHistItems);
  hipLaunchKernelGGL(k_histogram, dim3(blocks), dim3(kHistBlock), 0, s, keys, count, deg,
                     nbins);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_bits_for(uint64_t nkeys) {
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) < nkeys) b++;
  return b;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Range geometry of a sort of count items: nranges ranges of span items.
static void rb_geometry(uint64_t count, uint32_t *nranges, uint64_t *span) {
  const uint64_t ntiles = (count + kRbTile - 1) / kRbTile;
  const uint64_t per = (ntiles + kRbMaxRanges - 1) / kRbMaxRanges;  // tiles per range
  const uint64_t tp = per ? per : 1;
  *span = tp * kRbTile;
  *nranges = (uint32_t)((ntiles + tp - 1) / tp);
  if (*nranges == 0) *nranges = 1;
}

size_t eg_radix_tmp_bytes(uint64_t count, int np) {
  uint32_t nr;
  uint64_t span;
  rb_geometry(count, &nr, &span);
  const uint64_t cells = 256ull * nr;
  size_t b = 2 * align256(count * 4) * (size_t)(1 + np);  // ping-pong keys + payloads
  b += align256((cells + 1) * 4);                         // range counts (scanned in place)
  b += align256(eg_scan_tmp_bytes(cells));
  return b;
}

int eg_radix_sort(hipStream_t s, uint64_t count, int key_bits, const uint32_t *kin,
                  uint32_t *kout, const EgSortCols &cols, void *tmp, const EgIdCheck *chk) {
  if (count == 0) return EG_OK;
  if (count >= (1ull << 32))
    return eg_fail(EG_ENOTSUP, "radix sort: %llu items exceed 2^32",
                   (unsigned long long)count);
  const int np = cols.np;
  if (np < 0 || np > 3) return eg_fail(EG_EINVAL, "radix sort: %d payload columns", np);
  uint8_t *t = reinterpret_cast<uint8_t *>(tmp);
  uint32_t *kb[2];
  uint32_t *pb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  for (int h = 0; h < 2; h++) {
    kb[h] = reinterpret_cast<uint32_t *>(t);
    t += align256(count * 4);
    for (int p = 0; p < np; p++) {
      pb[h][p] = reinterpret_cast<uint32_t *>(t);
      t += align256(count * 4);
    }
  }
  uint32_t nranges;
  uint64_t span;
  rb_geometry(count, &nranges, &span);
  uint32_t *counts = reinterpret_cast<uint32_t *>(t);
  t += align256(((uint64_t)256 * nranges + 1) * 4);
  void *scan_tmp = t;
  // ids: keys in the first histogram, vals as payload column 0 of the first
  // scatter (or by a separate pass when vals is not a payload column)
  const uint32_t limit = chk ? chk->limit : 0u;
  const bool vals_in_scatter = chk && chk->vals && np > 0 && chk->vals == cols.in[0];
  if (limit && chk->vals && !vals_in_scatter) {
    hipLaunchKernelGGL(k_check_ids, dim3(record_grid(count)), dim3(256), 0, s, kin, chk->vals,
                       count, limit, chk->flag);
    EG_LAUNCHED();
  }

  int bits = key_bits < 1 ? 1 : key_bits;
  const int passes = (bits + 7) / 8;
  const int bpp = (bits + passes - 1) / passes;
  const uint32_t *cur_k = kin;
  const uint32_t *cur_p[3] = {cols.in[0], cols.in[1], cols.in[2]};
  for (int pass = 0; pass < passes; pass++) {
    const int shift = pass * bpp;
    const int nb = (bits - shift) < bpp ? (bits - shift) : bpp;
    const uint32_t mask = (1u << nb) - 1u;
    const bool last = pass == passes - 1;
    uint32_t *nk = last ? kout : kb[pass & 1];
    uint32_t *np_out[3] = {nullptr, nullptr, nullptr};
    for (int p = 0; p < np; p++) np_out[p] = last ? cols.out[p] : pb[pass & 1][p];
    const uint32_t klimit = pass == 0 ? limit : 0u;
    const uint32_t plimit = pass == 0 && vals_in_scatter ? limit : 0u;
    hipLaunchKernelGGL(k_rb_hist, dim3(nranges), dim3(kRbBlock), 0, s, cur_k, count, span,
                       shift, mask, nranges, counts, klimit, chk ? chk->flag : nullptr);
    EG_LAUNCHED();
    EG_TRY(eg_scan_excl(s, counts, counts, (uint64_t)(mask + 1) * nranges, scan_tmp));
    const bool wk = nk != nullptr;
    uint32_t *fl = chk ? chk->flag : nullptr;
    switch (np) {
      case 0: launch_rb_scatter<0>(s, wk, nranges, cur_k, nk, cur_p, np_out, count, span, shift, nb, counts, plimit, fl); break;
      case 1: launch_rb_scatter<1>(s, wk, nranges, cur_k, nk, cur_p, np_out, count, span, shift, nb, counts, plimit, fl); break;
      case 2: launch_rb_scatter<2>(s, wk, nranges, cur_k, nk, cur_p, np_out, count, span, shift, nb, counts, plimit, fl); break;
      default: launch_rb_scatter<3>(s, wk, nranges, cur_k, nk, cur_p, np_out, count, span, shift, nb, counts, plimit, fl); break;
    }
    EG_LAUNCHED();
    if (!last && nk == nullptr) return eg_fail(EG_EINVAL, "radix sort: key buffer missing");
    cur_k = nk;
    for (int p = 0; p < np; p++) cur_p[p] = np_out[p];
  }
  return EG_OK;
}

int eg_gen_device(hipStream_t s, const eg_gen_params *p, uint32_t n, uint32_t *src,
                  uint32_t *dst, float *w) {
  if (p->m_in == 0) return EG_OK;
  const uint32_t blocks = record_grid(p->m_in);
  const uint64_t skey = mix64_host(p->seed);
  if (p->kind == 0) {
    const double two32 = 4294967296.0;
    const uint32_t ta = (uint32_t)(p->a * two32);
    const uint32_t tab = (uint32_t)((p->a + p->b) * two32);
    const uint32_t tabc = (uint32_t)((p->a + p->b + p->c) * two32);
    hipLaunchKernelGGL(k_gen_rmat, dim3(blocks), dim3(256), 0, s, p->m_in, p->scale, skey,
                       ta, tab, tabc, src, dst);
  } else {
    hipLaunchKernelGGL(k_gen_er, dim3(blocks), dim3(256), 0, s, p->m_in, n, skey, src, dst);
  }
  EG_LAUNCHED();
  if (w) {
    const uint64_t wkey = mix64_host(p->weight_seed ^ 0x5745494748545321ull);
    hipLaunchKernelGGL(k_gen_weights, dim3(blocks), dim3(256), 0, s, p->m_in, wkey, w);
    EG_LAUNCHED();
  }
  return EG_OK;
}

int eg_unpack_records(hipStream_t s, const void *rec, uint64_t m_in, int weighted,
                      uint32_t *src, uint32_t *dst, float *w) {
  if (m_in == 0) return EG_OK;
  hipLaunchKernelGGL(k_unpack, dim3(record_grid(m_in)), dim3(256), 0, s,
                     reinterpret_cast<const uint32_t *>(rec), m_in, weighted ? 3 : 2, src,
                     dst, w);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_append_reverse(hipStream_t s, uint64_t m_in, uint32_t *src, uint32_t *dst, float *w) {
  if (m_in == 0) return EG_OK;
  hipLaunchKernelGGL(k_append_reverse, dim3(record_grid(m_in)), dim3(256), 0, s, m_in,
                     src, dst, w);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_check_ids(hipStream_t s, const uint32_t *a, const uint32_t *b, uint64_t count,
                 uint32_t n, uint32_t *d_flag, int *bad) {
  *bad = 0;
  if (count == 0) return EG_OK;
  EG_HIP(hipMemsetAsync(d_flag, 0, sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_check_ids, dim3(record_grid(count)), dim3(256), 0, s, a, b, count,
                     n, d_flag);
  EG_LAUNCHED();
  uint32_t h = 0;
  EG_HIP(hipMemcpyAsync(&h, d_flag, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  *bad = h != 0;
  return EG_OK;
}

int eg_cell_keys(hipStream_t s, const uint32_t *src, const uint32_t *dst, uint64_t count,
                 uint32_t n, uint32_t P, uint32_t *keys) {
  if (count == 0) return EG_OK;
  hipLaunchKernelGGL(k_cell_keys, dim3(record_grid(count)), dim3(256), 0, s, src, dst,
                     count, n, P, keys);
  EG_LAUNCHED();
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Exclusive scan of 64-bit counts (row offsets past 2^32 entries):
// kScan64Parts contiguous ranges are reduced, their totals scanned by one
// workgroup, and every range rescanned from its offset in tiles of 256.
// ---------------------------------------------------------------------------
namespace {

constexpr uint32_t kScan64Parts = 1024;

__device__ __forceinline__ uint64_t block_reduce_u64(uint64_t v, uint64_t *s_w) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  const uint64_t t = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(256) void k_scan64_reduce(const uint64_t *__restrict__ a,
                                                       uint64_t count, uint64_t per,
                                                       uint64_t *__restrict__ part) {
  __shared__ uint64_t s_w[4];
  const uint64_t lo = (uint64_t)blockIdx.x * per;
  const uint64_t hi = lo + per < count ? lo + per : count;
  uint64_t v = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) v += a[i];
  v = block_reduce_u64(v, s_w);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Serial exclusive scan of at most kScan64Parts totals (one thread).
__global__ void k_scan64_parts(uint64_t *__restrict__ part, uint32_t np) {
  if (threadIdx.x != 0) return;
  uint64_t run = 0;
  for (uint32_t i = 0; i < np; i++) {
    const uint64_t v = part[i];
    part[i] = run;
    run += v;
  }
  part[np] = run;
}

__global__ __launch_bounds__(256) void k_scan64_apply(uint64_t *__restrict__ a, uint64_t count,
                                                      uint64_t per,
                                                      const uint64_t *__restrict__ part) {
  __shared__ uint64_t s_w[4];
  const uint64_t lo = (uint64_t)blockIdx.x * per;
  const uint64_t hi = lo + per < count ? lo + per : count;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  uint64_t run = part[blockIdx.x];
  for (uint64_t c = lo; c < hi; c += 256) {
    const uint64_t i = c + threadIdx.x;
    const uint64_t x = i < hi ? a[i] : 0;
    uint64_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint64_t before = 0;
    for (int w = 0; w < wave; w++) before += s_w[w];
    const uint64_t tile = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    if (i < hi) a[i] = run + before + inc - x;
    run += tile;
    __syncthreads();
  }
  if (hi == count && threadIdx.x == 0) a[count] = run;
}

}  // namespace

size_t eg_scan64_tmp_bytes() { return (kScan64Parts + 1) * sizeof(uint64_t); }

int eg_scan_excl_u64(hipStream_t s, uint64_t *a, uint64_t count, void *tmp) {
  if (count == 0) {
    EG_HIP(hipMemsetAsync(a, 0, sizeof(uint64_t), s));
    return EG_OK;
  }
  uint64_t *part = static_cast<uint64_t *>(tmp);
  const uint64_t per = (count + kScan64Parts - 1) / kScan64Parts;
  const uint32_t np = (uint32_t)((count + per - 1) / per);
  hipLaunchKernelGGL(k_scan64_reduce, dim3(np), dim3(256), 0, s, a, count, per, part);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_scan64_parts, dim3(1), dim3(64), 0, s, part, np);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_scan64_apply, dim3(np), dim3(256), 0, s, a, count, per,
                     static_cast<const uint64_t *>(part));
  EG_LAUNCHED();
  return EG_OK;
}
