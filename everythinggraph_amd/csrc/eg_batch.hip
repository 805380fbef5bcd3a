// eg_batch.hip -- batched form: many BFS sources and many personalized
// PageRank vectors per sweep over the graph.  No reference counterpart; the
// per-query semantics are those of eg_bfs (bfs_simple.c levels) and of the
// pull PageRank recurrence (pagerank_simple.c:72-84) with a personalization
// vector.
//
// MS-BFS: up to 64 or 128 sources share one traversal.  Every vertex carries
// lane masks of one or two 64-bit words (seen, frontier, next frontier); with
// two, a pull gather loads both words with one 16-byte load.  A level is either
//  * a pull sweep over the plan of the pull orientation (the chunk / segment
//    decomposition of eg_pull.hip): a row ORs the frontier masks of its
//    in-neighbours, gathered eight at a time into LDS, or
//  * on symmetric graphs with a small frontier, a push over the frontier's
//    edges (one workgroup per 2048-edge slice, owners found by binary search
//    in LDS) into a pending mask, followed by an apply pass.
// The direction follows the reference BFS rule (bfs_simple.c:191-197) on the
// frontier of all lanes.  Levels of newly reached (vertex, lane) pairs are
// written lane-major as u16 from a per-lane table, so lanes that started at
// different steps (the serving engine) keep their own level numbering.
//
// PPR: k <= 64 vectors stored row-major (64 FP32 columns per vertex).  One
// wave owns a row: it loads 64 in-edge ids with one coalesced read, then for
// each broadcasts the id and every lane adds its column of the neighbour's
// contribution row (a coalesced 256-byte gather) with Neumaier compensation.
// Long rows are split into segments as in eg_pull.hip.  The seed term is
// added from a per-vertex 64-bit seed mask.
#include <math.h>
#include <string.h>

#include <algorithm>

#include "eg_batch.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kChunk = kPlanChunk;
constexpr uint32_t kLds = kPlanChunk + kPlanLongRow;
constexpr uint32_t kSerialRow = 32;
constexpr uint32_t kEdgesPerBlock = 2048;
constexpr uint32_t kStatBlocks = 1024;
constexpr uint32_t kMaxLanes = 256;  // four lane words
// Grids of the device-driven level kernels (grid-stride): the kernels of the
// direction a level does not take are dispatched and return at once.
constexpr uint32_t kDevGrid = 2048;
constexpr uint32_t kDevPullGrid = 4096;

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}
// Grid of a grid-stride kernel of 256-thread blocks over count elements:
// at most 2^20 blocks (2^28 threads), far below the 2^32 threads of a launch.
inline uint32_t stride_grid(uint64_t count) {
  const uint64_t b = (count + 255) / 256;
  return (uint32_t)(b < (1u << 20) ? (b ? b : 1) : (1u << 20));
}

__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// The lane masks of a vertex: NW 64-bit words (64 or 128 lanes), stored
// vertex-major (words [v * NW, v * NW + NW)); lane l is bit l % 64 of word
// l / 64.
template <int NW>
struct Lanes {
  uint64_t w[NW];
  __device__ __forceinline__ static Lanes zero() {
    Lanes r;
#pragma unroll
    for (int j = 0; j < NW; j++) r.w[j] = 0ull;
    return r;
  }
  __device__ __forceinline__ bool any() const {
    uint64_t a = 0;
#pragma unroll
    for (int j = 0; j < NW; j++) a |= w[j];
    return a != 0;
  }
  __device__ __forceinline__ Lanes operator|(const Lanes &o) const {
    Lanes r;
#pragma unroll
    for (int j = 0; j < NW; j++) r.w[j] = w[j] | o.w[j];
    return r;
  }
  __device__ __forceinline__ Lanes operator&(const Lanes &o) const {
    Lanes r;
#pragma unroll
    for (int j = 0; j < NW; j++) r.w[j] = w[j] & o.w[j];
    return r;
  }
  __device__ __forceinline__ Lanes operator~() const {
    Lanes r;
#pragma unroll
    for (int j = 0; j < NW; j++) r.w[j] = ~w[j];
    return r;
  }
  // every lane of need is set here
  __device__ __forceinline__ bool covers(const Lanes &need) const {
    uint64_t miss = 0;
#pragma unroll
    for (int j = 0; j < NW; j++) miss |= need.w[j] & ~w[j];
    return miss == 0;
  }
};

template <int NW>
__device__ __forceinline__ Lanes<NW> ld_lanes(const uint64_t *__restrict__ a, uint64_t v) {
  Lanes<NW> r;
  if constexpr (NW == 2 || NW == 4) {
#pragma unroll
    for (int h = 0; h < NW / 2; h++) {
      const ulonglong2 x = reinterpret_cast<const ulonglong2 *>(a)[v * (NW / 2) + h];
      r.w[2 * h] = x.x;
      r.w[2 * h + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < NW; j++) r.w[j] = a[v * NW + j];
  }
  return r;
}

template <int NW>
__device__ __forceinline__ void st_lanes(uint64_t *__restrict__ a, uint64_t v, const Lanes<NW> &x) {
  if constexpr (NW == 2 || NW == 4) {
#pragma unroll
    for (int h = 0; h < NW / 2; h++)
      reinterpret_cast<ulonglong2 *>(a)[v * (NW / 2) + h] = make_ulonglong2(x.w[2 * h], x.w[2 * h + 1]);
  } else {
#pragma unroll
    for (int j = 0; j < NW; j++) a[v * NW + j] = x.w[j];
  }
}

template <int NW>
__device__ __forceinline__ Lanes<NW> wave_or(Lanes<NW> v) {
#pragma unroll
  for (int j = 0; j < NW; j++) v.w[j] = wave_or(v.w[j]);
  return v;
}

// ctr word holding the progress bits of lane word j
__host__ __device__ constexpr int progress_slot(int j) { return j == 0 ? 0 : 3 + j; }
// ctr words 8 and 9: count and degree sum of the frontier the running level
// builds (device-driven levels); the partials of the passes over n start at 10.
constexpr int kCtrFront = 8;
constexpr int kCtrPart = 10;

// Device-driven levels (eg_bfs_batch): control words written by the level
// kernels themselves, so that the host enqueues levels without reading
// anything back between them.
constexpr int kCtlDone = 0;        // the traversal has ended (every kernel returns at once)
constexpr int kCtlPush = 1;        // this level pushes (1) or pulls (0)
constexpr int kCtlLevel = 2;       // level value written for newly reached (vertex, lane) pairs
constexpr int kCtlOverflow = 3;    // the level would pass the u16 store
constexpr int kCtlLevels = 4;      // levels run
constexpr int kCtlPushLevels = 5;  // of which pushed
constexpr int kCtlWords = 8;
constexpr uint32_t kMaxLevelValue = 65534;

// True when a level kernel has nothing to do: the traversal ended, or the
// level runs the other direction (want_push < 0: either).  ctl is null for
// levels the host drives.
__device__ __forceinline__ bool ms_skip(const uint32_t *__restrict__ ctl, int want_push) {
  if (!ctl) return false;
  return ctl[kCtlDone] != 0u || (want_push >= 0 && ctl[kCtlPush] != (uint32_t)want_push);
}

// Adds this workgroup's new lanes to the progress words of ctr (only bits not
// yet recorded).
template <int NW>
__device__ __forceinline__ void publish_progress(Lanes<NW> nw, uint64_t *s_or,
                                                 unsigned long long *ctr) {
  nw = wave_or(nw);
  if (__lane_id() == 0)
#pragma unroll
    for (int j = 0; j < NW; j++) s_or[(threadIdx.x >> 6) * NW + j] = nw.w[j];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NW; j++) {
      uint64_t b = 0;
      for (int w = 0; w < kWaves; w++) b |= s_or[w * NW + j];
      unsigned long long *slot = &ctr[progress_slot(j)];
      const uint64_t known = __atomic_load_n(slot, __ATOMIC_RELAXED);
      if (b & ~known) atomicOr(slot, (unsigned long long)b);
    }
  }
}

// Out-degree statistics of the vertices a level reaches, added to fs for the
// next level's direction decision: fs[0] those with out-edges, fs[1] their
// degree sum.  One pair of atomics per workgroup that reached any.
__device__ __forceinline__ void publish_front_stats(unsigned long long c, unsigned long long d,
                                                    unsigned long long *s_cd,
                                                    unsigned long long *__restrict__ fs) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    c += __shfl_xor(c, o);
    d += __shfl_xor(d, o);
  }
  if (__lane_id() == 0) {
    s_cd[2 * (threadIdx.x >> 6)] = c;
    s_cd[2 * (threadIdx.x >> 6) + 1] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    c = s_cd[0] + s_cd[2] + s_cd[4] + s_cd[6];
    d = s_cd[1] + s_cd[3] + s_cd[5] + s_cd[7];
    if (c) {
      atomicAdd(&fs[0], c);
      atomicAdd(&fs[1], d);
    }
  }
}

// Frontier bitmap of the device-driven levels (bit v of word v / 64: v is in
// the frontier and has out-edges), kept by the kernels that build the frontier
// so that the push queue is compacted from n / 8 bytes instead of the masks.
// Sets the bits of rows r + i (bit i of mask) to bit i of bits and leaves the
// others.  Every bit has one writer per level; rows of one word may belong to
// two workgroups, hence the atomics.
__device__ __forceinline__ void fbits_put(uint64_t *fbits, uint32_t r, uint64_t mask,
                                          uint64_t bits) {
  unsigned long long *f = reinterpret_cast<unsigned long long *>(fbits);
  bits &= mask;
  const uint32_t w = r >> 6, o = r & 63u;
  const uint64_t m0 = mask << o, b0 = bits << o;
  if (m0 == ~0ull) {
    f[w] = b0;
  } else {
    atomicAnd(&f[w], ~m0);
    if (b0) atomicOr(&f[w], b0);
  }
  if (o) {
    const uint64_t m1 = mask >> (64u - o), b1 = bits >> (64u - o);
    if (m1) {
      atomicAnd(&f[w + 1], ~m1);
      if (b1) atomicOr(&f[w + 1], b1);
    }
  }
}

// Level store: lane-major u16, lvl[lane * n + v].
template <int NW>
__device__ __forceinline__ void store_levels(uint32_t n, uint32_t r, const Lanes<NW> &nw,
                                             uint16_t *__restrict__ lvl, const uint16_t *tab) {
#pragma unroll
  for (int j = 0; j < NW; j++) {
    uint64_t b = nw.w[j];
    while (b) {
      const int i = 64 * j + __ffsll((unsigned long long)b) - 1;
      lvl[(uint64_t)i * n + r] = tab[i];
      b &= b - 1;
    }
  }
}

// Level store of a whole wave (every lane active; a lane without new lanes
// passes zero): the wave walks the union of its new lanes, so that the stores
// of one lane index go out together, 64 rows in one instruction, instead of
// one scattered 2-byte store per (row, lane) pair.  The same values as
// store_levels.
template <int NW>
__device__ __forceinline__ void store_levels_wave(uint32_t n, uint32_t r, const Lanes<NW> &nw,
                                                  uint16_t *__restrict__ lvl,
                                                  const uint16_t *tab) {
  const Lanes<NW> u = wave_or(nw);
#pragma unroll
  for (int j = 0; j < NW; j++) {
    uint64_t b = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u.w[j] >> 32))
                  << 32) |
                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u.w[j]);
    while (b) {
      const int k = __ffsll((unsigned long long)b) - 1;
      if ((nw.w[j] >> k) & 1ull) lvl[(uint64_t)(64 * j + k) * n + r] = tab[64 * j + k];
      b &= b - 1ull;
    }
  }
}

// LV: store the levels here (else the caller stores them for its wave)
template <int NW, bool LV = true>
__device__ __forceinline__ Lanes<NW> finish_ms_row(uint32_t n, uint32_t r, const Lanes<NW> &acc,
                                                   const Lanes<NW> &active,
                                                   uint64_t *__restrict__ seen,
                                                   uint64_t *__restrict__ nvisit,
                                                   uint16_t *__restrict__ lvl,
                                                   const uint16_t *tab) {
  const Lanes<NW> s = ld_lanes<NW>(seen, r);
  const Lanes<NW> nw = acc & ~s & active;
  st_lanes<NW>(nvisit, r, nw);
  if (nw.any()) {
    st_lanes<NW>(seen, r, s | nw);
    if (LV) store_levels<NW>(n, r, nw, lvl, tab);
  }
  return nw;
}

// The kernels over n x 64 arrays are grid-stride loops: at n = 2^26 one
// thread per element would exceed the 2^32 threads of a launch.
template <int NW>
__global__ void k_ms_reset(uint32_t n, uint64_t *__restrict__ seen, uint64_t *__restrict__ visit,
                           uint64_t *__restrict__ nvisit, uint64_t *__restrict__ pend) {
  // the level store is not cleared: its readers take a level only where the
  // lane's seen bit is set
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)n * NW;
       i += stride) {
    seen[i] = 0;
    visit[i] = 0;
    nvisit[i] = 0;
    pend[i] = 0;
  }
}

template <int NW>
__global__ void k_ms_seed(uint32_t n, uint32_t k, const uint32_t *__restrict__ lanes,
                          const uint32_t *__restrict__ roots, const uint32_t *__restrict__ perm,
                          uint64_t *__restrict__ seen, uint64_t *__restrict__ visit,
                          uint16_t *__restrict__ lvl) {
  const uint32_t i = threadIdx.x;
  if (i >= k) return;
  const uint32_t r = perm ? perm[roots[i]] : roots[i];
  const uint32_t lane = lanes[i];
  const uint64_t word = (uint64_t)r * NW + (lane >> 6);
  atomicOr(reinterpret_cast<unsigned long long *>(&seen[word]), 1ull << (lane & 63));
  atomicOr(reinterpret_cast<unsigned long long *>(&visit[word]), 1ull << (lane & 63));
  lvl[(uint64_t)lane * n + r] = 1;
}

// Rows whose new lanes nw are final: record them (only this thread writes
// row r in this level).
template <int NW, bool LV = true>
__device__ __forceinline__ Lanes<NW> finish_ms_new(uint32_t n, uint32_t r, const Lanes<NW> &nw,
                                                   uint64_t *__restrict__ seen,
                                                   uint64_t *__restrict__ nvisit,
                                                   uint16_t *__restrict__ lvl,
                                                   const uint16_t *tab) {
  st_lanes<NW>(nvisit, r, nw);
  if (nw.any()) {
    st_lanes<NW>(seen, r, ld_lanes<NW>(seen, r) | nw);
    if (LV) store_levels<NW>(n, r, nw, lvl, tab);
  }
  return nw;
}

// Pull level over the plan.  A row needs the active lanes it has not seen;
// it ORs the frontier masks of its in-neighbours until it holds all of them
// (OR is exact, so stopping early changes nothing).
//  * Segments of long rows OR into seg_or (skipped when the row needs nothing).
//  * A chunk first checks whether any of its rows needs a lane; if none does it
//    only clears their next-frontier words.  Otherwise it stages the chunk's
//    entries in LDS (B per lane per pass, coalesced) and each needy row gathers
//    from them: rows of up to kSerialRow entries by one lane, four gathers in
//    flight, longer rows by a wave, 128 entries per check.
// With NW = 2 a gather loads the 16 bytes of both lane words at once.
template <int B, int NW>
__device__ __forceinline__ void ms_pull_item(
    uint32_t item, uint32_t n, const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
    const uint4 *__restrict__ chunk_info, const uint4 *__restrict__ seg_info,
    const uint32_t *__restrict__ long_rows, uint32_t nseg, uint64_t *__restrict__ seg_or,
    uint64_t *__restrict__ seen, const uint64_t *__restrict__ visit,
    uint64_t *__restrict__ nvisit, uint16_t *__restrict__ lvl, const uint16_t *__restrict__ tab,
    const Lanes<NW> &active, unsigned long long *__restrict__ ctr, unsigned long long &fc,
    unsigned long long &fd, uint64_t *__restrict__ fbits) {
  static_assert(kLds <= 65535 && kChunk <= 65535, "u16 LDS offsets");
  __shared__ uint32_t s_idx[kLds];
  __shared__ uint16_t s_ptr[kChunk + 1];
  __shared__ uint16_t s_list[kChunk];
  __shared__ uint32_t s_nlist;
  __shared__ uint16_t s_tab[64 * NW];
  __shared__ uint64_t s_or[kWaves * NW];
  const uint32_t tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  if (item < nseg) {
    const uint4 si = seg_info[item];
    const Lanes<NW> need = active & ~ld_lanes<NW>(seen, long_rows[si.z]);
    Lanes<NW> acc = Lanes<NW>::zero();
    if (need.any()) {
      for (uint32_t base = si.x; base < si.y; base += kBlock * B) {
        uint32_t u[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
          const uint32_t k = base + tid + q * kBlock;
          u[q] = k < si.y ? __builtin_nontemporal_load(idx + k) : idx[si.x];
        }
#pragma unroll
        for (int q = 0; q < B; q++) acc = acc | ld_lanes<NW>(visit, u[q]);
      }
    }
    acc = wave_or(acc);
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < NW; j++) s_or[wave * NW + j] = acc.w[j];
    __syncthreads();
    if (tid == 0)
#pragma unroll
      for (int j = 0; j < NW; j++)
        seg_or[(uint64_t)item * NW + j] =
            s_or[0 * NW + j] | s_or[1 * NW + j] | s_or[2 * NW + j] | s_or[3 * NW + j];
    return;
  }
  const uint4 ci = chunk_info[item - nseg];
  const uint32_t r0 = ci.x, r1 = ci.y, k0 = ci.z, k1 = ci.w;
  const uint32_t nrows = r1 - r0;
  if (tid == 0) s_nlist = 0;
  for (uint32_t i = tid; i < 64 * NW; i += kBlock) s_tab[i] = tab[i];
  int any = 0;
  for (uint32_t i = tid; i < nrows; i += kBlock)
    any |= (active & ~ld_lanes<NW>(seen, r0 + i)).any();
  for (uint32_t i = tid; i <= nrows; i += kBlock) s_ptr[i] = (uint16_t)(ptr[r0 + i] - k0);
  any = __syncthreads_or(any);
  if (!any) {
    for (uint32_t i = tid; i < nrows; i += kBlock) {
      st_lanes<NW>(nvisit, r0 + i, Lanes<NW>::zero());
      // lane 0 holds the wave's first row
      if (fbits) {
        const uint64_t m = __ballot(1);  // the wave's rows, from lane 0's
        if (lane == 0) fbits_put(fbits, r0 + i, m, 0ull);
      }
    }
    return;
  }
  for (uint32_t base = k0; base < k1; base += kBlock * B) {
    uint32_t u[B];
#pragma unroll
    for (int q = 0; q < B; q++) {
      const uint32_t k = base + tid + q * kBlock;
      u[q] = k < k1 ? __builtin_nontemporal_load(idx + k) : 0u;
    }
#pragma unroll
    for (int q = 0; q < B; q++) {
      const uint32_t k = base + tid + q * kBlock;
      if (k < k1) s_idx[k - k0] = u[q];
    }
  }
  __syncthreads();
  Lanes<NW> nwall = Lanes<NW>::zero();
  // every lane of a wave runs every step (the level store is the wave's)
  for (uint32_t ib = tid - lane; ib < nrows; ib += kBlock) {
    const uint32_t i = ib + lane;
    const bool live = i < nrows;
    const uint32_t ra = live ? s_ptr[i] : 0u, rb = live ? s_ptr[i + 1] : 0u;
    bool hit = false;  // a new frontier row with entries (listed rows: their own bit below)
    const bool serial = live && rb - ra <= kSerialRow;
    Lanes<NW> nwr = Lanes<NW>::zero();
    if (serial) {
      const uint32_t r = r0 + i;
      const Lanes<NW> need = active & ~ld_lanes<NW>(seen, r);
      Lanes<NW> acc = Lanes<NW>::zero();
      if (need.any()) {
        for (uint32_t q = ra; q < rb; q += 4) {
          const Lanes<NW> v0 = ld_lanes<NW>(visit, s_idx[q]);
          const Lanes<NW> v1 = q + 1 < rb ? ld_lanes<NW>(visit, s_idx[q + 1]) : Lanes<NW>::zero();
          const Lanes<NW> v2 = q + 2 < rb ? ld_lanes<NW>(visit, s_idx[q + 2]) : Lanes<NW>::zero();
          const Lanes<NW> v3 = q + 3 < rb ? ld_lanes<NW>(visit, s_idx[q + 3]) : Lanes<NW>::zero();
          acc = acc | ((v0 | v1) | (v2 | v3));
          if (acc.covers(need)) break;
        }
      }
      const Lanes<NW> nw = finish_ms_new<NW, false>(n, r, acc & need, seen, nvisit, lvl, s_tab);
      if (nw.any()) {
        fc += rb > ra;
        fd += rb - ra;
        hit = rb > ra;
      }
      nwall = nwall | nw;
      nwr = nw;
    } else if (live) {
      s_list[atomicAdd(&s_nlist, 1u)] = (uint16_t)i;
    }
    if (fbits) {
      const uint64_t m = __ballot(serial), b = __ballot(hit);
      if (lane == 0) fbits_put(fbits, r0 + ib, m, b);
    }
    store_levels_wave<NW>(n, r0 + i, nwr, lvl, s_tab);
  }
  __syncthreads();
  const uint32_t nlist = s_nlist;
  for (uint32_t li = wave; li < nlist; li += kWaves) {
    const uint32_t i = s_list[li];
    const uint32_t r = r0 + i;
    const uint32_t ra = s_ptr[i], rb = s_ptr[i + 1];
    const Lanes<NW> need = active & ~ld_lanes<NW>(seen, r);
    Lanes<NW> acc = Lanes<NW>::zero();
    if (need.any()) {
      for (uint32_t base = ra; base < rb; base += 128) {
        const uint32_t qa = base + lane, qb = base + 64 + lane;
        const Lanes<NW> va = qa < rb ? ld_lanes<NW>(visit, s_idx[qa]) : Lanes<NW>::zero();
        const Lanes<NW> vb = qb < rb ? ld_lanes<NW>(visit, s_idx[qb]) : Lanes<NW>::zero();
        acc = wave_or(acc | va | vb);
        if (acc.covers(need)) break;
      }
    }
    if (lane == 0) {
      const Lanes<NW> nw = finish_ms_new<NW>(n, r, acc & need, seen, nvisit, lvl, s_tab);
      if (nw.any()) {
        fc += 1;  // a listed row has more than kSerialRow entries
        fd += rb - ra;
      }
      if (fbits) {
        unsigned long long *f = reinterpret_cast<unsigned long long *>(&fbits[r >> 6]);
        if (nw.any())
          atomicOr(f, 1ull << (r & 63u));
        else
          atomicAnd(f, ~(1ull << (r & 63u)));
      }
      nwall = nwall | nw;
    }
  }
  publish_progress<NW>(nwall, s_or, ctr);
}

// The pull level's work items (segments of long rows, then chunks) for
// item = blockIdx.x, blockIdx.x + gridDim.x, ... < nitems: one item per
// workgroup when the host drives the levels, a fixed grid when the device
// does (a push level then dispatches few idle workgroups).
template <int B, int NW>
__global__ __launch_bounds__(kBlock) void k_ms_pull(
    uint32_t n, const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ idx,
    const uint4 *__restrict__ chunk_info, const uint4 *__restrict__ seg_info,
    const uint32_t *__restrict__ long_rows, uint32_t nseg, uint32_t nitems,
    uint64_t *__restrict__ seg_or, uint64_t *__restrict__ seen, const uint64_t *__restrict__ visit,
    uint64_t *__restrict__ nvisit, uint16_t *__restrict__ lvl, const uint16_t *__restrict__ tab,
    Lanes<NW> active, unsigned long long *__restrict__ ctr, const uint32_t *__restrict__ ctl,
    unsigned long long *__restrict__ fs, uint64_t *__restrict__ fbits) {
  __shared__ unsigned long long s_cd[2 * kWaves];
  if (ms_skip(ctl, 0)) return;
  unsigned long long fc = 0, fd = 0;
  for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    ms_pull_item<B, NW>(item, n, ptr, idx, chunk_info, seg_info, long_rows, nseg, seg_or, seen,
                        visit, nvisit, lvl, tab, active, ctr, fc, fd, fbits);
    __syncthreads();  // the item's LDS is reused by the next one
  }
  if (fs) publish_front_stats(fc, fd, s_cd, fs);
}

template <int NW>
__global__ __launch_bounds__(kBlock) void k_ms_pull_finish(
    uint32_t n, const uint32_t *__restrict__ ptr, const uint32_t *__restrict__ long_rows,
    const uint32_t *__restrict__ seg_base, uint32_t nlong,
    const uint64_t *__restrict__ seg_or, uint64_t *__restrict__ seen,
    uint64_t *__restrict__ nvisit, uint16_t *__restrict__ lvl, const uint16_t *__restrict__ tab,
    Lanes<NW> active, unsigned long long *__restrict__ ctr, const uint32_t *__restrict__ ctl,
    unsigned long long *__restrict__ fs, uint64_t *__restrict__ fbits) {
  __shared__ uint64_t s_or[kWaves * NW];
  __shared__ unsigned long long s_cd[2 * kWaves];
  if (ms_skip(ctl, 0)) return;
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  Lanes<NW> nw = Lanes<NW>::zero();
  unsigned long long fc = 0, fd = 0;
  if (j < nlong) {
    Lanes<NW> acc = Lanes<NW>::zero();
    for (uint32_t s = seg_base[j]; s < seg_base[j + 1]; s++) acc = acc | ld_lanes<NW>(seg_or, s);
    const uint32_t r = long_rows[j];
    nw = finish_ms_row<NW>(n, r, acc, active, seen, nvisit, lvl, tab);
    if (fs && nw.any()) {
      fc = 1;  // a long row has entries
      fd = ptr[r + 1] - ptr[r];
    }
    if (fbits) {  // no chunk holds a long row: its bit is set here only
      unsigned long long *f = reinterpret_cast<unsigned long long *>(&fbits[r >> 6]);
      if (nw.any())
        atomicOr(f, 1ull << (r & 63u));
      else
        atomicAnd(f, ~(1ull << (r & 63u)));
    }
  }
  publish_progress<NW>(nw, s_or, ctr);
  if (fs) publish_front_stats(fc, fd, s_cd, fs);
}

// Frontier statistics of visit: [1] vertices with out-edges, [2] their degree
// sum (per-workgroup partials, summed by k_ms_stats_sum into out[0..1]).
template <int NW>
__global__ __launch_bounds__(kBlock) void k_ms_stats(uint32_t n, const uint64_t *__restrict__ visit,
                                                     const uint32_t *__restrict__ ptr,
                                                     unsigned long long *__restrict__ part,
                                                     const uint32_t *__restrict__ ctl,
                                                     uint64_t *__restrict__ fbits) {
  __shared__ unsigned long long s_c[kWaves], s_d[kWaves];
  if (ms_skip(ctl, -1)) return;
  unsigned long long c = 0, d = 0;
  // a wave takes 64 consecutive vertices (one bitmap word) per step
  for (uint32_t vb = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); vb < n;
       vb += gridDim.x * blockDim.x) {
    const uint32_t v = vb + __lane_id();
    bool hit = false;
    if (v < n && ld_lanes<NW>(visit, v).any()) {
      const uint32_t dv = ptr[v + 1] - ptr[v];
      c += dv > 0;
      d += dv;
      hit = dv > 0;
    }
    if (fbits) {
      const uint64_t b = __ballot(hit);
      if (__lane_id() == 0) fbits[vb >> 6] = b;
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    c += __shfl_xor(c, o);
    d += __shfl_xor(d, o);
  }
  if (__lane_id() == 0) {
    s_c[threadIdx.x >> 6] = c;
    s_d[threadIdx.x >> 6] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    part[2 * blockIdx.x + 1] = s_d[0] + s_d[1] + s_d[2] + s_d[3];
  }
}

__global__ __launch_bounds__(kBlock) void k_ms_stats_sum(const unsigned long long *__restrict__ part,
                                                         uint32_t np,
                                                         unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_c[kWaves], s_d[kWaves];
  unsigned long long c = 0, d = 0;
  for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
    c += part[2 * i];
    d += part[2 * i + 1];
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    c += __shfl_xor(c, o);
    d += __shfl_xor(d, o);
  }
  if (__lane_id() == 0) {
    s_c[threadIdx.x >> 6] = c;
    s_d[threadIdx.x >> 6] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    out[1] = s_d[0] + s_d[1] + s_d[2] + s_d[3];
  }
}

// Frontier vertices with out-edges -> queue (run only for small frontiers:
// one atomic per workgroup that holds some).
template <int NW>
__global__ __launch_bounds__(kBlock) void k_ms_compact(uint32_t n, const uint64_t *__restrict__ visit,
                                                       const uint32_t *__restrict__ ptr,
                                                       uint32_t *__restrict__ q,
                                                       unsigned long long *__restrict__ qlen,
                                                       const uint32_t *__restrict__ ctl) {
  __shared__ uint32_t s_q[kBlock];
  __shared__ uint32_t s_n, s_base;
  if (ms_skip(ctl, 1)) return;
  for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const uint32_t v = b0 + threadIdx.x;
    if (v < n && ld_lanes<NW>(visit, v).any() && ptr[v + 1] > ptr[v])
      s_q[atomicAdd(&s_n, 1u)] = v;
    __syncthreads();
    if (threadIdx.x == 0 && s_n) s_base = (uint32_t)atomicAdd(qlen, (unsigned long long)s_n);
    __syncthreads();
    if (threadIdx.x < s_n) q[s_base + threadIdx.x] = s_q[threadIdx.x];
    __syncthreads();
  }
}

__global__ void k_ms_qdeg(const uint32_t *__restrict__ q, uint32_t qn,
                          const uint32_t *__restrict__ ptr, uint32_t *__restrict__ deg) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < qn) deg[i] = ptr[q[i] + 1] - ptr[q[i]];
}

// Push level: the frontier's edges [b*2048, (b+1)*2048) of the degree scan,
// blocks b = blockIdx.x, blockIdx.x + gridDim.x, ...; the queue length is
// *qn_dev when given (device-driven levels), else qn.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_ms_push(
    const uint32_t *__restrict__ q, uint32_t qn_host, const unsigned long long *__restrict__ qn_dev,
    const uint32_t *__restrict__ qoff, const uint32_t *__restrict__ ptr,
    const uint32_t *__restrict__ idx, const uint64_t *__restrict__ visit,
    const uint64_t *__restrict__ seen, uint64_t *__restrict__ pend, Lanes<NW> active,
    const uint32_t *__restrict__ ctl) {
  __shared__ uint32_t s_off[kEdgesPerBlock + 1];
  __shared__ uint32_t s_u[kEdgesPerBlock + 1];
  __shared__ uint64_t s_vis[(kEdgesPerBlock + 1) * NW];
  __shared__ uint32_t s_first, s_cnt;
  if (ms_skip(ctl, 1)) return;
  const uint32_t qn = qn_dev ? (uint32_t)*qn_dev : qn_host;
  if (qn == 0) return;
  const uint32_t total = qoff[qn];
  for (uint32_t blk = blockIdx.x; (uint64_t)blk * kEdgesPerBlock < total; blk += gridDim.x) {
    const uint32_t e0 = blk * kEdgesPerBlock;
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
    const uint32_t cnt = s_cnt <= kEdgesPerBlock + 1 ? s_cnt : kEdgesPerBlock + 1;
    for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
      s_off[i] = qoff[s_first + i];
      const uint32_t u = q[s_first + i];
      s_u[i] = u;
      const Lanes<NW> vis = ld_lanes<NW>(visit, u) & active;
#pragma unroll
      for (int j = 0; j < NW; j++) s_vis[i * NW + j] = vis.w[j];
    }
    __syncthreads();
    for (uint32_t base = e0; base < e1; base += kBlock) {
      const uint32_t e = base + threadIdx.x;
      if (e < e1) {
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) / 2;
          if (s_off[mid] <= e) lo = mid; else hi = mid;
        }
        const uint32_t w = idx[ptr[s_u[lo]] + (e - s_off[lo])];
#pragma unroll
        for (int j = 0; j < NW; j++) {
          const uint64_t word = (uint64_t)w * NW + j;
          const uint64_t bits =
              s_vis[lo * NW + j] & ~__atomic_load_n(&seen[word], __ATOMIC_RELAXED);
          if (bits && (__atomic_load_n(&pend[word], __ATOMIC_RELAXED) & bits) != bits)
            atomicOr(reinterpret_cast<unsigned long long *>(&pend[word]), (unsigned long long)bits);
        }
      }
    }
    __syncthreads();  // the LDS tables are rebuilt for the next block of edges
  }
}

template <int NW>
__global__ __launch_bounds__(kBlock) void k_ms_apply(uint32_t n, const uint32_t *__restrict__ ptr,
                                                     uint64_t *__restrict__ pend,
                                                     uint64_t *__restrict__ seen,
                                                     uint64_t *__restrict__ nvisit,
                                                     uint16_t *__restrict__ lvl,
                                                     const uint16_t *__restrict__ tab,
                                                     Lanes<NW> active,
                                                     unsigned long long *__restrict__ ctr,
                                                     const uint32_t *__restrict__ ctl,
                                                     unsigned long long *__restrict__ fs,
                                                     uint64_t *__restrict__ fbits) {
  __shared__ uint64_t s_or[kWaves * NW];
  __shared__ uint16_t s_tab[64 * NW];
  __shared__ unsigned long long s_cd[2 * kWaves];
  if (ms_skip(ctl, 1)) return;
  for (uint32_t i = threadIdx.x; i < 64 * NW; i += blockDim.x) s_tab[i] = tab[i];
  __syncthreads();
  Lanes<NW> all = Lanes<NW>::zero();
  unsigned long long fc = 0, fd = 0;
  // a wave takes 64 consecutive vertices (one bitmap word) per step
  for (uint32_t vb = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); vb < n;
       vb += gridDim.x * blockDim.x) {
    const uint32_t v = vb + __lane_id();
    bool hit = false;
    Lanes<NW> nwr = Lanes<NW>::zero();
    if (v < n) {
      const Lanes<NW> p = ld_lanes<NW>(pend, v);
      if (p.any()) {
        st_lanes<NW>(pend, v, Lanes<NW>::zero());
        const Lanes<NW> nw = finish_ms_row<NW, false>(n, v, p, active, seen, nvisit, lvl, s_tab);
        nwr = nw;
        if (fs && nw.any()) {
          const uint32_t dv = ptr[v + 1] - ptr[v];
          fc += dv > 0;
          fd += dv;
          hit = dv > 0;
        }
        all = all | nw;
      } else {
        st_lanes<NW>(nvisit, v, Lanes<NW>::zero());
      }
    }
    if (fbits) {
      const uint64_t b = __ballot(hit);
      if (__lane_id() == 0) fbits[vb >> 6] = b;
    }
    store_levels_wave<NW>(n, v, nwr, lvl, s_tab);
  }
  publish_progress<NW>(all, s_or, ctr);
  if (fs) publish_front_stats(fc, fd, s_cd, fs);
}

// ---- device-driven levels (eg_bfs_batch) ----
// Every level enqueues the same kernels with the same grids: the direction
// decision (from the frontier statistics that the last level's kernels added
// up as they reached its vertices; k_ms_stats counts the seeds), the push kernels (queue, degrees,
// their scan, the edge push, the apply pass) and the pull kernels; the ones
// of the other direction return at once.  The decision applies the
// reference's rule (bfs_simple.c:191-197) to the frontier of all lanes, as the
// host-driven levels do.
__global__ void k_msd_init(uint32_t *__restrict__ ctl) {
  if (threadIdx.x < kCtlWords) ctl[threadIdx.x] = threadIdx.x == kCtlLevel ? 2u : 0u;
}

// Takes the statistics of the frontier the last level built (fs, zeroed for
// this level's kernels), picks the direction, arms the level: level table,
// zeroed progress words and queue length.
template <int NW>
__global__ __launch_bounds__(kBlock) void k_msd_decide(unsigned long long *__restrict__ fs,
                                                       double thr, int sym,
                                                       uint32_t *__restrict__ ctl,
                                                       unsigned long long *__restrict__ ctr,
                                                       uint16_t *__restrict__ tab) {
  if (ctl[kCtlDone]) return;
  const uint32_t level = ctl[kCtlLevel];
  __syncthreads();  // orders every thread's read of the level before its update
  for (uint32_t i = threadIdx.x; i < 64 * NW; i += blockDim.x) tab[i] = (uint16_t)level;
  if (threadIdx.x == 0) {
    const unsigned long long cnt = fs[0];
    const unsigned long long deg = fs[1];
    fs[0] = 0;
    fs[1] = 0;
    ctr[1] = cnt;
    ctr[2] = deg;
    ctr[3] = 0;
#pragma unroll
    for (int j = 0; j < NW; j++) ctr[progress_slot(j)] = 0;
    if (level > kMaxLevelValue) {
      ctl[kCtlOverflow] = 1u;
      ctl[kCtlDone] = 1u;
    } else {
      const uint32_t push = sym && (double)(cnt + deg) <= thr ? 1u : 0u;
      ctl[kCtlPush] = push;
      ctl[kCtlLevel] = level + 1;
      ctl[kCtlLevels] += 1u;
      ctl[kCtlPushLevels] += push;
    }
  }
}

// Ends the traversal after a level that reached nothing.
template <int NW>
__global__ void k_msd_end(uint32_t *__restrict__ ctl, const unsigned long long *__restrict__ ctr) {
  if (ctl[kCtlDone]) return;
  unsigned long long p = 0;
#pragma unroll
  for (int j = 0; j < NW; j++) p |= ctr[progress_slot(j)];
  if (!p) ctl[kCtlDone] = 1u;
}

// Push queue of a device-driven level from the frontier bitmap: a thread
// expands one word, a workgroup reserves its slots with one atomic.
__global__ __launch_bounds__(kBlock) void k_msd_compact_bits(uint32_t nwords,
                                                            const uint64_t *__restrict__ fbits,
                                                            uint32_t *__restrict__ q,
                                                            unsigned long long *__restrict__ qlen,
                                                            const uint32_t *__restrict__ ctl) {
  __shared__ uint32_t s_w[kWaves];
  __shared__ uint32_t s_base;
  if (ms_skip(ctl, 1)) return;
  const int lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < nwords; b0 += gridDim.x * blockDim.x) {
    const uint32_t wi = b0 + threadIdx.x;
    uint64_t word = wi < nwords ? fbits[wi] : 0ull;
    const uint32_t cnt = (uint32_t)__popcll(word);
    uint32_t x = cnt;  // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (int w = 0; w < kWaves; w++) t += s_w[w];
      s_base = t ? (uint32_t)atomicAdd(qlen, (unsigned long long)t) : 0u;
    }
    __syncthreads();
    uint32_t pos = s_base + x - cnt;
    for (uint32_t w = 0; w < wave; w++) pos += s_w[w];
    while (word) {
      q[pos++] = 64u * wi + (uint32_t)(__ffsll((unsigned long long)word) - 1);
      word &= word - 1ull;
    }
    __syncthreads();  // s_w, s_base reuse
  }
}

// Out-degrees of the queue (length *qn), grid-stride.
__global__ void k_msd_qdeg(const uint32_t *__restrict__ q, const unsigned long long *__restrict__ qn,
                           const uint32_t *__restrict__ ptr, uint32_t *__restrict__ deg,
                           const uint32_t *__restrict__ ctl) {
  if (ms_skip(ctl, 1)) return;
  const uint32_t n = (uint32_t)*qn;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    deg[i] = ptr[q[i] + 1] - ptr[q[i]];
}

// Exclusive scan of a[0..*qn) in place, a[*qn] = total, with a fixed grid:
// kScanParts slices of ceil(qn / kScanParts) entries; slice sums, their scan,
// then each slice rescanned from its offset.
constexpr uint32_t kScanParts = 256;

__device__ __forceinline__ void scan_slice(uint32_t qn, uint32_t &lo, uint32_t &hi) {
  const uint32_t per = (uint32_t)(((uint64_t)qn + kScanParts - 1) / kScanParts);
  lo = (uint32_t)min((uint64_t)qn, (uint64_t)blockIdx.x * per);
  hi = (uint32_t)min((uint64_t)qn, (uint64_t)lo + per);
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t x, uint32_t *s_w) {
  x = eg_wave_sum_u32(x);
  if (__lane_id() == 0) s_w[threadIdx.x >> 6] = x;
  __syncthreads();
  const uint32_t t = s_w[0] + s_w[1] + s_w[2] + s_w[3];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(kBlock) void k_msd_scan_a(const uint32_t *__restrict__ a,
                                                       const unsigned long long *__restrict__ qn,
                                                       uint32_t *__restrict__ part,
                                                       const uint32_t *__restrict__ ctl) {
  __shared__ uint32_t s_w[kWaves];
  if (ms_skip(ctl, 1)) return;
  uint32_t lo, hi;
  scan_slice((uint32_t)*qn, lo, hi);
  uint32_t x = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) x += a[i];
  x = block_sum_u32(x, s_w);
  if (threadIdx.x == 0) part[blockIdx.x] = x;
}

__global__ __launch_bounds__(kScanParts) void k_msd_scan_b(uint32_t *__restrict__ a,
                                                           const unsigned long long *__restrict__ qn,
                                                           uint32_t *__restrict__ part,
                                                           const uint32_t *__restrict__ ctl) {
  __shared__ uint32_t s_w[kScanParts / 64];
  if (ms_skip(ctl, 1)) return;
  const uint32_t t = threadIdx.x;
  const uint32_t x = part[t];
  uint32_t inc = x;
  const int lane = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) s_w[t >> 6] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t w = 0; w < (t >> 6); w++) off += s_w[w];
  __syncthreads();
  part[t] = off + inc - x;
  if (t == kScanParts - 1) a[(uint32_t)*qn] = off + inc;
}

__global__ __launch_bounds__(kBlock) void k_msd_scan_c(uint32_t *__restrict__ a,
                                                       const unsigned long long *__restrict__ qn,
                                                       const uint32_t *__restrict__ part,
                                                       const uint32_t *__restrict__ ctl) {
  __shared__ uint32_t s_w[kWaves];
  if (ms_skip(ctl, 1)) return;
  uint32_t lo, hi;
  scan_slice((uint32_t)*qn, lo, hi);
  uint32_t run = part[blockIdx.x];
  const int lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  for (uint32_t base = lo; base < hi; base += kBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t x = i < hi ? a[i] : 0u;
    uint32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o);
      if (lane >= o) inc += y;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w = 0; w < wave; w++) off += s_w[w];
    const uint32_t total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
    if (i < hi) a[i] = run + off + inc - x;
    run += total;
  }
}

__global__ void k_ms_extract(uint32_t n, uint32_t k, uint32_t lane0,
                             const uint16_t *__restrict__ lvl, const uint64_t *__restrict__ seen,
                             uint32_t nw, const uint32_t *__restrict__ perm,
                             uint32_t *__restrict__ out) {
  // out[i * n + v] = level of vertex v (layout id) in lane lane0 + i, 0 where
  // the lane did not reach it
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * k;
       t += stride) {
    const uint32_t i = (uint32_t)(t / n), v = (uint32_t)(t % n);
    const uint32_t vv = perm ? perm[v] : v;
    const uint32_t L = lane0 + i;
    const bool reached = (seen[(uint64_t)vv * nw + (L >> 6)] >> (L & 63u)) & 1ull;
    out[t] = reached ? lvl[(uint64_t)L * n + vv] : 0u;
  }
}

// Retired lanes: drop their bits from every mask (their levels then no
// longer count: a level holds only where the seen bit is set).  A lane
// retires after a level in which it reached nothing, so its frontier bits are
// already clear; nvisit and pend are rewritten by every level.
template <int NW>
__global__ void k_ms_clear_lanes(uint32_t n, Lanes<NW> lanes, uint64_t *__restrict__ seen,
                                 uint64_t *__restrict__ visit) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  st_lanes<NW>(seen, v, ld_lanes<NW>(seen, v) & ~lanes);
  st_lanes<NW>(visit, v, ld_lanes<NW>(visit, v) & ~lanes);
}

// Reached vertices of lane lanes[blockIdx.y]: per-workgroup partial counts.
__global__ __launch_bounds__(kBlock) void k_ms_reached(uint32_t n, const uint64_t *__restrict__ seen,
                                                       uint32_t nw,
                                                       const uint32_t *__restrict__ lanes,
                                                       unsigned long long *__restrict__ part) {
  __shared__ unsigned long long s_c[kWaves];
  const uint32_t L = lanes[blockIdx.y];
  const uint64_t *col = seen + (L >> 6);
  unsigned long long c = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    c += (col[(uint64_t)v * nw] >> (L & 63u)) & 1ull;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(uint64_t)blockIdx.y * gridDim.x + blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
}

__global__ void k_ms_reached_sum(const unsigned long long *__restrict__ part, uint32_t nb,
                                 unsigned long long *__restrict__ out) {
  unsigned long long c = 0;
  for (uint32_t b = threadIdx.x; b < nb; b += 64) c += part[(uint64_t)blockIdx.x * nb + b];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
  if (threadIdx.x == 0) out[blockIdx.x] = c;
}

// ---- PPR ----
__global__ void k_ppr_zero(uint64_t count, float *__restrict__ X, float *__restrict__ C,
                           float *__restrict__ Y, float *__restrict__ Cn) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += stride) {
    X[t] = 0.f;
    C[t] = 0.f;
    Y[t] = 0.f;
    Cn[t] = 0.f;
  }
}

__global__ void k_ppr_col_seed(uint32_t q, const uint32_t *__restrict__ seeds, uint32_t ns,
                               const uint32_t *__restrict__ perm, const float *__restrict__ w,
                               const uint32_t *__restrict__ deg, uint64_t *__restrict__ mask,
                               float *__restrict__ X, float *__restrict__ C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const uint32_t v = perm ? perm[seeds[i]] : seeds[i];
  atomicOr(reinterpret_cast<unsigned long long *>(&mask[v]), 1ull << q);
  X[(uint64_t)v * 64 + q] = w[q];
  const uint32_t d = deg[v];
  C[(uint64_t)v * 64 + q] = d ? w[q] / (float)d : 0.f;
}

// Zeroes columns `cols` of all four iterate buffers and drops them from the
// seed masks: a wave per row, lane q owns column q.  Columns at or above the
// active width are neither read nor written by k_ppr_spmm, so every buffer
// must hold zeros there.
__global__ void k_ppr_cols_clear(uint32_t n, uint64_t cols, uint64_t *__restrict__ mask,
                                 float *__restrict__ X, float *__restrict__ C,
                                 float *__restrict__ Y, float *__restrict__ Cn) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * 64;
       t += stride) {
    const uint32_t v = (uint32_t)(t >> 6);
    const int q = (int)(t & 63);
    if (q == 0) {
      const uint64_t m = mask[v];
      if (m & cols) mask[v] = m & ~cols;
    }
    if ((cols >> q) & 1ull) {
      X[(uint64_t)v * 64 + q] = 0.f;
      C[(uint64_t)v * 64 + q] = 0.f;
      Y[(uint64_t)v * 64 + q] = 0.f;
      Cn[(uint64_t)v * 64 + q] = 0.f;
    }
  }
}

// Adds the contribution rows of entries [k0, k1) to column q, for the W-lane
// slot `slot` of the wave (each slot walks its own row).  D row gathers are in
// flight per slot; entries are added in ascending order whatever W and D, so
// a column's sum never depends on how many columns are active.  Lanes with
// on = false issue no gathers.
template <int D, int W>
__device__ __forceinline__ void spmm_slot(const uint32_t *__restrict__ idx,
                                          const float *__restrict__ C, uint32_t k0, uint32_t k1,
                                          int slot, int q, bool on, EgKahan &acc) {
  const uint32_t len = k1 - k0;
  uint32_t lmax = len;
#pragma unroll
  for (int o = W; o < 64; o <<= 1) lmax = max(lmax, (uint32_t)__shfl_xor((int)lmax, o));
  for (uint32_t blk = 0; blk < lmax; blk += W) {
    const uint32_t cnt = len > blk ? (len - blk < (uint32_t)W ? len - blk : (uint32_t)W) : 0u;
    const uint32_t my = (uint32_t)q < cnt ? __builtin_nontemporal_load(idx + k0 + blk + q) : 0u;
    uint32_t cmax = cnt;
#pragma unroll
    for (int o = W; o < 64; o <<= 1) cmax = max(cmax, (uint32_t)__shfl_xor((int)cmax, o));
    for (uint32_t j = 0; j < cmax; j += D) {
      float v[D];
#pragma unroll
      for (int t = 0; t < D; t++) {
        const uint32_t u = __shfl(my, slot * W + (int)((j + t) & (W - 1)));
        v[t] = (on && j + t < cnt) ? C[(uint64_t)u * 64 + q] : 0.f;
      }
#pragma unroll
      for (int t = 0; t < D; t++)
        if (j + t < cnt) acc.add(v[t]);
    }
  }
}

// Writes row v of Cn, and of Y unless write_y is 0 (a step whose Y no later
// step reads: the gathers read Cn only, and X only for the DELTA change);
// returns |y - x| (x = the row of X) when DELTA, which needs write_y.
template <bool DELTA>
__device__ __forceinline__ double spmm_finish(uint32_t v, float sum, int lane, bool on,
                                              const uint64_t *__restrict__ mask, float seed_term,
                                              const uint32_t *__restrict__ deg, float one_minus_a,
                                              uint32_t write_y, const float *__restrict__ X,
                                              float *__restrict__ Y, float *__restrict__ Cn) {
  if (!on) return 0.0;
  float y = one_minus_a * sum;
  if ((mask[v] >> lane) & 1ull) y += seed_term;
  if (DELTA || write_y) Y[(uint64_t)v * 64 + lane] = y;
  const uint32_t d = deg[v];
  Cn[(uint64_t)v * 64 + lane] = d ? y / (float)d : 0.f;
  if (!DELTA) return 0.0;
  return fabs((double)y - (double)X[(uint64_t)v * 64 + lane]);
}

// Column sums of this workgroup's per-wave deltas, merged in wave order.
template <bool DELTA>
__device__ __forceinline__ void store_delta(double d, double *__restrict__ dpart, uint32_t slot) {
  if (!DELTA) return;
  __shared__ double s_d[kWaves][64];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  s_d[wave][lane] = d;
  __syncthreads();
  if (wave == 0) {
    double t = s_d[0][lane];
    for (int q = 1; q < kWaves; q++) t += s_d[q][lane];
    dpart[(uint64_t)slot * 64 + lane] = t;
  }
}

// Work items of the pull plan, for W = 16, 32 or 64 gathered columns (a wave
// works on 64 / W rows at once, one W-lane slot per row; lane slot*W + q owns
// column q).  Everything that is summed uses an order independent of W:
//  * a chunk row is summed by one slot over all its entries;
//  * a long-row segment is cut into kSegParts parts, each summed by one slot,
//    and the part sums are merged in part order;
//  * a chunk's per-column L1 change is binned by row index mod kDeltaBins and
//    the bins are merged in bin order.
// A workgroup takes items blockIdx.x, blockIdx.x + gridDim.x, ...: a grid
// smaller than the item count leaves wave slots to concurrent kernels.
constexpr int kSegParts = 16;
constexpr int kDeltaBins = 16;

template <bool DELTA, int D, int W>
__global__ __launch_bounds__(kBlock) void k_ppr_spmm(
    const uint32_t *__restrict__ idx, const uint4 *__restrict__ chunk_info,
    const uint4 *__restrict__ seg_info, uint32_t nseg, uint32_t nitems,
    float2 *__restrict__ seg_part, const uint32_t *__restrict__ ptr,
    const float *__restrict__ C, const uint64_t *__restrict__ mask,
    const float *__restrict__ alpha_inv_seeds, const uint32_t *__restrict__ deg,
    float one_minus_a, uint32_t width, uint32_t write_y, const float *__restrict__ X,
    float *__restrict__ Y, float *__restrict__ Cn, double *__restrict__ dpart) {
  constexpr int R = 64 / W;                       // rows per wave
  constexpr int kUnits = kWaves * R;              // row slots per workgroup
  constexpr int kBinsPerUnit = kDeltaBins / kUnits;
  static_assert(kSegParts % kUnits == 0 && kDeltaBins % kUnits == 0, "slot layout");
  __shared__ float2 s_p[kSegParts][64];
  __shared__ double s_d[kDeltaBins][64];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const int slot = lane / W;
  const int q = lane % W;
  const int unit = wave * R + slot;
  const bool on = (uint32_t)q < width;
  for (uint32_t item = blockIdx.x; item < nitems; item += gridDim.x) {
    if (item < nseg) {
      const uint4 si = seg_info[item];
      const uint32_t part = (si.y - si.x + kSegParts - 1) / kSegParts;
      for (int p = unit; p < kSegParts; p += kUnits) {
        const uint32_t a = si.x + p * part < si.y ? si.x + p * part : si.y;
        const uint32_t b = a + part < si.y ? a + part : si.y;
        EgKahan acc;
        spmm_slot<D, W>(idx, C, a, b, slot, q, on, acc);
        s_p[p][q] = make_float2(acc.s, acc.c);
      }
      __syncthreads();
      if (wave == 0) {
        float2 out = make_float2(0.f, 0.f);
        if (lane < W) {
          EgKahan t;
          t.s = s_p[0][lane].x;
          t.c = s_p[0][lane].y;
          for (int p = 1; p < kSegParts; p++) t.merge(s_p[p][lane].x, s_p[p][lane].y);
          out = make_float2(t.s, t.c);
        }
        seg_part[(uint64_t)item * 64 + lane] = out;
      }
    } else {
      const uint4 ci = chunk_info[item - nseg];
      const float seed_term = alpha_inv_seeds[q];
      double dbin[kBinsPerUnit];
#pragma unroll
      for (int b = 0; b < kBinsPerUnit; b++) dbin[b] = 0.0;
      uint32_t j = 0;
      for (uint32_t v0 = ci.x + wave * R; v0 < ci.y; v0 += kUnits, j++) {
        const uint32_t v = v0 + slot;
        const bool row = v < ci.y;
        EgKahan acc;
        spmm_slot<D, W>(idx, C, row ? ptr[v] : 0u, row ? ptr[v + 1] : 0u, slot, q, on, acc);
        if (row) {
          const double d = spmm_finish<DELTA>(v, acc.value(), q, on, mask, seed_term, deg,
                                              one_minus_a, write_y, X, Y, Cn);
          if (DELTA) dbin[j % kBinsPerUnit] += d;
        }
      }
      if (DELTA) {
#pragma unroll
        for (int b = 0; b < kBinsPerUnit; b++) s_d[unit + b * kUnits][q] = dbin[b];
        __syncthreads();
        if (wave == 0) {
          double t = 0.0;
          if (lane < W)
            for (int b = 0; b < kDeltaBins; b++) t += s_d[b][lane];
          dpart[(uint64_t)(item - nseg) * 64 + lane] = t;
        }
      }
    }
    __syncthreads();  // LDS reuse by the next item
  }
}

// Long rows: a wave per row merges its segment partials in segment order.
// Delta partials go to slots [slot0, slot0 + gridDim.x).
template <bool DELTA>
__global__ __launch_bounds__(kBlock) void k_ppr_finish(
    const uint32_t *__restrict__ long_rows, const uint32_t *__restrict__ seg_base, uint32_t nlong,
    const float2 *__restrict__ seg_part, const uint64_t *__restrict__ mask,
    const float *__restrict__ alpha_inv_seeds, const uint32_t *__restrict__ deg,
    float one_minus_a, uint32_t width, uint32_t write_y, const float *__restrict__ X,
    float *__restrict__ Y, float *__restrict__ Cn, double *__restrict__ dpart, uint32_t slot0) {
  const int lane = __lane_id();
  const bool on = (uint32_t)lane < width;
  const uint32_t j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  double d = 0.0;
  if (j < nlong) {
    EgKahan acc;
    for (uint32_t s = seg_base[j]; s < seg_base[j + 1]; s++) {
      const float2 p = seg_part[(uint64_t)s * 64 + lane];
      acc.merge(p.x, p.y);
    }
    d = spmm_finish<DELTA>(long_rows[j], acc.value(), lane, on, mask, alpha_inv_seeds[lane], deg,
                           one_minus_a, write_y, X, Y, Cn);
  }
  store_delta<DELTA>(d, dpart, slot0 + blockIdx.x);
}

// Column sums of dpart[0..np) in a fixed order: workgroup b sums a
// contiguous slice (four quarter-slices, merged in order) into part2[b];
// k_ppr_delta_final sums part2 in order.
constexpr uint32_t kDeltaBlocks = 512;
// Partials [skip_lo, skip_hi) belong to skipped static chunks, whose deltas
// are exactly 0: they add nothing, which gives the same bits as adding them.
__global__ __launch_bounds__(kBlock) void k_ppr_delta_partial(const double *__restrict__ dpart,
                                                              uint32_t np, uint32_t skip_lo,
                                                              uint32_t skip_hi,
                                                              double *__restrict__ part2) {
  __shared__ double s_d[kWaves][64];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const uint32_t per = (np + kDeltaBlocks - 1) / kDeltaBlocks;
  const uint32_t b0 = blockIdx.x * per < np ? blockIdx.x * per : np;
  const uint32_t b1 = b0 + per < np ? b0 + per : np;
  const uint32_t q = (b1 - b0 + kWaves - 1) / kWaves;
  const uint32_t a = b0 + wave * q < b1 ? b0 + wave * q : b1;
  const uint32_t e = a + q < b1 ? a + q : b1;
  double t = 0.0;
  for (uint32_t i = a; i < e; i++)
    if (i < skip_lo || i >= skip_hi) t += dpart[(uint64_t)i * 64 + lane];
  s_d[wave][lane] = t;
  __syncthreads();
  if (wave == 0) {
    double u = s_d[0][lane];
    for (int w = 1; w < kWaves; w++) u += s_d[w][lane];
    part2[(uint64_t)blockIdx.x * 64 + lane] = u;
  }
}

__global__ void k_ppr_delta_final(const double *__restrict__ part2, double *__restrict__ delta) {
  const int lane = threadIdx.x;
  double t = 0.0;
  for (uint32_t b = 0; b < kDeltaBlocks; b++) t += part2[(uint64_t)b * 64 + lane];
  delta[lane] = t;
}

__global__ void k_ppr_extract(uint32_t n, uint32_t k, uint32_t col0, const float *__restrict__ X,
                              const uint32_t *__restrict__ perm, float *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * k;
       t += stride) {
    const uint32_t q = (uint32_t)(t / n), v = (uint32_t)(t % n);
    const uint32_t vv = perm ? perm[v] : v;
    out[t] = X[(uint64_t)vv * 64 + col0 + q];
  }
}

// static_from = 1 + the last row with entries or out-edges.
__global__ void k_static_from(uint32_t n, const uint32_t *__restrict__ ptr,
                              const uint32_t *__restrict__ deg, uint32_t *__restrict__ out) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t last = 0;
  if (r < n && (ptr[r + 1] > ptr[r] || deg[r] > 0)) last = r + 1;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) last = max(last, (uint32_t)__shfl_xor(last, o));
  if (__lane_id() == 0 && last) atomicMax(out, last);
}

// out[1] = first chunk whose first row is >= out[0] (nchunks if none).
__global__ void k_first_chunk_from(const uint4 *__restrict__ chunk_info, uint32_t nchunks,
                                   uint32_t *__restrict__ out) {
  const uint32_t from = out[0];
  uint32_t lo = 0, hi = nchunks;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (chunk_info[mid].x >= from) hi = mid; else lo = mid + 1;
  }
  out[1] = lo;
}

}  // namespace

// First chunk of the plan whose rows have neither entries nor out-edges
// (computed once per plan).
static int static_chunk0(eg_graph *g, const uint32_t *ptr, const uint32_t *deg, PullPlan *pl) {
  if (pl->static_chunk0 != 0xFFFFFFFFu) return EG_OK;
  hipStream_t s = g->stream;
  uint32_t *d = nullptr;
  EG_TRY(eg_dalloc_t(&d, 2, "static rows"));
  uint32_t h[2] = {0, 0};
  int rc = EG_OK;
  if (hipMemsetAsync(d, 0, 8, s) != hipSuccess) rc = eg_fail(EG_EDEVICE, "static rows clear");
  if (!rc && g->n) {
    hipLaunchKernelGGL(k_static_from, dim3((g->n + 255) / 256), dim3(256), 0, s, g->n, ptr, deg,
                       d);
    hipLaunchKernelGGL(k_first_chunk_from, dim3(1), dim3(1), 0, s, pl->chunk_info, pl->nchunks, d);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = eg_fail(EG_EDEVICE, "static rows");
  }
  eg_dfree_t(&d);
  EG_TRY(rc);
  pl->static_chunk0 = g->n ? h[1] : pl->nchunks;
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Views and state.
// ---------------------------------------------------------------------------
int eg_batch_view(eg_graph *g, BatchView *v) {
  v->sym = (g->flags & (EG_UNDIRECTED | EG_SYMMETRIC)) != 0 || !g->in_ptr;
  if (eg_reorder_enabled()) {
    EgOrdered *o;
    EG_TRY(eg_ordered(g, &o));
    v->ptr = o->ptr;
    v->idx = o->idx;
    v->deg = o->deg;
    v->perm = o->new_of_old;
    v->pl = &o->plan;
    v->ms_idx = o->idx;
    if (g->flags & (EG_UNDIRECTED | EG_SYMMETRIC)) {
      const uint32_t *rows = nullptr;
      EG_TRY(eg_hubs_first_rows(g, &rows));
      if (rows) v->ms_idx = rows;
    }
  } else {
    const float *w;
    EG_TRY(eg_pull_orientation(g, &v->ptr, &v->idx, &w, &v->pl));
    v->deg = g->out_deg;
    v->perm = nullptr;
    v->ms_idx = v->idx;
  }
  EG_TRY(static_chunk0(g, v->ptr, v->deg, v->pl));
  v->static_chunk0 = v->pl->static_chunk0;
  return EG_OK;
}

int eg_ms_alloc(eg_graph *g, MsState *st, uint32_t nw) {
  if (nw != 1 && nw != 2 && nw != 4) return eg_fail(EG_EINVAL, "MS-BFS lane words must be 1, 2 or 4");
  const uint32_t n = g->n;
  st->nw = nw;
  const uint64_t nn = (uint64_t)n * nw;
  EG_TRY(eg_dalloc_t(&st->seen, nn, "ms seen"));
  EG_TRY(eg_dalloc_t(&st->visit, nn, "ms visit"));
  EG_TRY(eg_dalloc_t(&st->nvisit, nn, "ms next"));
  EG_TRY(eg_dalloc_t(&st->pend, nn, "ms pending"));
  EG_TRY(eg_dalloc_t(&st->lvl, (uint64_t)n * 64 * nw, "ms levels"));
  EG_TRY(eg_dalloc_t(&st->queue, (uint64_t)n + 64, "ms queue"));
  EG_TRY(eg_dalloc_t(&st->qoff, (uint64_t)n + 65, "ms queue offsets"));
  EG_TRY(eg_dalloc_t(&st->fbits, (uint64_t)n / 64 + 2, "ms frontier bitmap"));
  // [0] and [4..6] progress words, [1] count, [2] degree sum, [3] queue
  // length, [8..9] the statistics of the frontier a device-driven level
  // builds; [10..) holds the frontier statistics partials (2 per workgroup) or
  // the per-lane reached partials (64 nw x 64) and sums (64 nw), never both at once.
  EG_TRY(eg_dalloc_t(&st->ctr, kCtrPart + 2 * (uint64_t)kStatBlocks + 64 * 64 * nw + 64 * nw,
                     "ms counters"));
  EG_TRY(eg_dalloc_t(&st->lvl_tab, 64 * nw, "ms level table"));
  EG_TRY(eg_dalloc_t(&st->roots, 2 * kMaxLanes, "ms roots"));
  EG_TRY(eg_dalloc_t(&st->ctl, kCtlWords, "ms control words"));
  // the host-driven queue scan, or the kScanParts slice sums of the
  // device-driven one
  const size_t scan_bytes = std::max<size_t>(eg_scan_tmp_bytes((uint64_t)n + 64), 4 * kScanParts);
  EG_TRY(eg_dalloc_t(reinterpret_cast<uint8_t **>(&st->scan_tmp), scan_bytes,
                     "ms scan workspace"));
  if (hipHostMalloc(reinterpret_cast<void **>(&st->h), (8 + 64 * nw) * 8, hipHostMallocDefault) !=
      hipSuccess)
    return eg_fail(EG_ENOMEM, "pinned counters");
  return EG_OK;
}

void eg_ms_free(MsState *st) {
  eg_dfree_t(&st->seen);
  eg_dfree_t(&st->visit);
  eg_dfree_t(&st->nvisit);
  eg_dfree_t(&st->pend);
  eg_dfree_t(&st->fbits);
  eg_dfree_t(&st->lvl);
  eg_dfree_t(&st->seg_or);
  eg_dfree_t(&st->queue);
  eg_dfree_t(&st->qoff);
  eg_dfree_t(&st->ctr);
  eg_dfree_t(&st->lvl_tab);
  eg_dfree_t(&st->roots);
  eg_dfree_t(&st->ctl);
  if (st->exec) (void)hipGraphExecDestroy(st->exec);
  st->exec = nullptr;
  uint8_t *t = reinterpret_cast<uint8_t *>(st->scan_tmp);
  eg_dfree_t(&t);
  st->scan_tmp = nullptr;
  if (st->h) (void)hipHostFree(st->h);
  st->h = nullptr;
}

template <int NW>
static Lanes<NW> to_lanes(const uint64_t *w) {
  Lanes<NW> r;
  for (int j = 0; j < NW; j++) r.w[j] = w[j];
  return r;
}

int eg_ms_reset(eg_graph *g, MsState *st) {
  const uint32_t n = g->n;
  const uint64_t count = (uint64_t)n * st->nw;
  if (st->nw == 4)
    hipLaunchKernelGGL(k_ms_reset<4>, dim3(stride_grid(count)), dim3(256), 0, st->stream, n,
                       st->seen, st->visit, st->nvisit, st->pend);
  else if (st->nw == 2)
    hipLaunchKernelGGL(k_ms_reset<2>, dim3(stride_grid(count)), dim3(256), 0, st->stream, n,
                       st->seen, st->visit, st->nvisit, st->pend);
  else
    hipLaunchKernelGGL(k_ms_reset<1>, dim3(stride_grid(count)), dim3(256), 0, st->stream, n,
                       st->seen, st->visit, st->nvisit, st->pend);
  EG_LAUNCHED();
  st->qlen = 0;
  st->qdeg = 0;
  return EG_OK;
}

// Long-row segment ORs of a pull level (nw words per segment), allocated
// outside any stream capture.
static int ms_seg_or(MsState *st, const PullPlan &pl) {
  if (st->seg_cap < pl.nseg || !st->seg_or) {
    eg_dfree_t(&st->seg_or);
    EG_TRY(eg_dalloc_t(&st->seg_or, (uint64_t)(pl.nseg ? pl.nseg : 1) * st->nw, "ms segment ors"));
    st->seg_cap = pl.nseg;
  }
  return EG_OK;
}

// Frontier of visit: statistics, and the queue when the next level may push.
template <int NW>
static int frontier_prep(eg_graph *g, const BatchView &bv, MsState *st, bool *push) {
  const uint32_t n = g->n;
  hipStream_t s = st->stream;
  unsigned long long *part = st->ctr + kCtrPart;
  const uint32_t nb = blocks_for(n, kBlock) < kStatBlocks ? blocks_for(n, kBlock) : kStatBlocks;
  hipLaunchKernelGGL(k_ms_stats<NW>, dim3(nb), dim3(kBlock), 0, s, n, st->visit, bv.ptr, part,
                     (const uint32_t *)nullptr, (uint64_t *)nullptr);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_ms_stats_sum, dim3(1), dim3(kBlock), 0, s, part, nb, st->ctr + 1);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(st->h, st->ctr, 24, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  const uint64_t cnt = st->h[1], deg = st->h[2];
  // bfs_simple.c:191: pull once the frontier's edges exceed m/20 (m / ms_push_div)
  *push = bv.sym && (double)(cnt + deg) <= (double)g->m / (double)g->tune.ms_push_div;
  st->qdeg = deg;
  st->qlen = 0;
  if (*push && cnt) {
    EG_HIP(hipMemsetAsync(st->ctr + 3, 0, 8, s));
    hipLaunchKernelGGL(k_ms_compact<NW>, dim3(nb), dim3(kBlock), 0, s, n, st->visit, bv.ptr,
                       st->queue, st->ctr + 3, (const uint32_t *)nullptr);
    EG_LAUNCHED();
    st->qlen = (uint32_t)cnt;
  }
  return EG_OK;
}

int eg_ms_seed(eg_graph *g, const BatchView &bv, MsState *st, uint32_t k, const uint32_t *lanes,
               const uint32_t *roots) {
  if (k == 0) return EG_OK;
  if (k > 64 * st->nw) return eg_fail(EG_EINVAL, "at most %u lanes", 64 * st->nw);
  for (uint32_t i = 0; i < k; i++)
    if (lanes[i] >= 64 * st->nw) return eg_fail(EG_EINVAL, "lane %u out of range", lanes[i]);
  hipStream_t s = st->stream;
  uint32_t h[2 * kMaxLanes];
  memcpy(h, lanes, k * 4);
  memcpy(h + kMaxLanes, roots, k * 4);
  EG_HIP(hipMemcpyAsync(st->roots, h, sizeof(h), hipMemcpyHostToDevice, s));
  if (st->nw == 4)
    hipLaunchKernelGGL(k_ms_seed<4>, dim3(1), dim3(kMaxLanes), 0, s, g->n, k, st->roots,
                       st->roots + kMaxLanes, bv.perm, st->seen, st->visit, st->lvl);
  else if (st->nw == 2)
    hipLaunchKernelGGL(k_ms_seed<2>, dim3(1), dim3(kMaxLanes), 0, s, g->n, k, st->roots,
                       st->roots + kMaxLanes, bv.perm, st->seen, st->visit, st->lvl);
  else
    hipLaunchKernelGGL(k_ms_seed<1>, dim3(1), dim3(kMaxLanes), 0, s, g->n, k, st->roots,
                       st->roots + kMaxLanes, bv.perm, st->seen, st->visit, st->lvl);
  EG_LAUNCHED();
  EG_HIP(hipStreamSynchronize(s));  // h is on the stack
  return EG_OK;
}

template <int NW>
static int ms_level(eg_graph *g, const BatchView &bv, MsState *st, const uint64_t *active_w,
                    const uint16_t *tab, uint64_t *progress) {
  const uint32_t n = g->n;
  hipStream_t s = st->stream;
  const Lanes<NW> active = to_lanes<NW>(active_w);
  bool push = false;
  EG_TRY(frontier_prep<NW>(g, bv, st, &push));
  st->pushed = push;
  EG_HIP(hipMemcpyAsync(st->lvl_tab, tab, 128 * NW, hipMemcpyHostToDevice, s));
  for (int j = 0; j < NW; j++) EG_HIP(hipMemsetAsync(st->ctr + progress_slot(j), 0, 8, s));
  const PullPlan &pl = *bv.pl;
  if (push) {
    if (st->qlen && st->qdeg) {
      hipLaunchKernelGGL(k_ms_qdeg, dim3(blocks_for(st->qlen, 256)), dim3(256), 0, s, st->queue,
                         st->qlen, bv.ptr, st->qoff);
      EG_LAUNCHED();
      EG_TRY(eg_scan_excl(s, st->qoff, st->qoff, st->qlen, st->scan_tmp));
      EgProf prof(g, "msbfs_push", s);
      hipLaunchKernelGGL(k_ms_push<NW>, dim3(blocks_for(st->qdeg, kEdgesPerBlock)), dim3(kBlock),
                         0, s, st->queue, st->qlen, (const unsigned long long *)nullptr, st->qoff,
                         bv.ptr, bv.ms_idx, st->visit, st->seen, st->pend, active,
                         (const uint32_t *)nullptr);
      EG_LAUNCHED();
    }
    const uint32_t nb = blocks_for(n, kBlock) < 4096 ? blocks_for(n, kBlock) : 4096;
    hipLaunchKernelGGL(k_ms_apply<NW>, dim3(nb), dim3(kBlock), 0, s, n, bv.ptr, st->pend,
                       st->seen, st->nvisit, st->lvl, st->lvl_tab, active, st->ctr,
                       (const uint32_t *)nullptr, (unsigned long long *)nullptr, (uint64_t *)nullptr);
    EG_LAUNCHED();
  } else {
    EG_TRY(ms_seg_or(st, pl));
    {
      EgProf prof(g, "msbfs_pull", s);
      if (g->tune.ms_batch == 16)
        hipLaunchKernelGGL((k_ms_pull<16, NW>), dim3(pl.nseg + pl.nchunks), dim3(kBlock), 0, s,
                           n, bv.ptr, bv.ms_idx, pl.chunk_info, pl.seg_info, pl.long_rows,
                           pl.nseg, pl.nseg + pl.nchunks, st->seg_or, st->seen, st->visit, st->nvisit, st->lvl,
                           st->lvl_tab, active, st->ctr, (const uint32_t *)nullptr,
                           (unsigned long long *)nullptr, (uint64_t *)nullptr);
      else
        hipLaunchKernelGGL((k_ms_pull<8, NW>), dim3(pl.nseg + pl.nchunks), dim3(kBlock), 0, s, n,
                           bv.ptr, bv.ms_idx, pl.chunk_info, pl.seg_info, pl.long_rows, pl.nseg,
                           pl.nseg + pl.nchunks, st->seg_or, st->seen, st->visit, st->nvisit, st->lvl, st->lvl_tab,
                           active, st->ctr, (const uint32_t *)nullptr,
                           (unsigned long long *)nullptr, (uint64_t *)nullptr);
      EG_LAUNCHED();
    }
    if (pl.nlong) {
      hipLaunchKernelGGL(k_ms_pull_finish<NW>, dim3(blocks_for(pl.nlong, kBlock)), dim3(kBlock),
                         0, s, n, bv.ptr, pl.long_rows, pl.seg_base, pl.nlong, st->seg_or,
                         st->seen, st->nvisit, st->lvl, st->lvl_tab, active, st->ctr,
                         (const uint32_t *)nullptr, (unsigned long long *)nullptr,
                         (uint64_t *)nullptr);
      EG_LAUNCHED();
    }
  }
  std::swap(st->visit, st->nvisit);
  EG_HIP(hipMemcpyAsync(st->h, st->ctr, 64, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < NW; j++) progress[j] = st->h[progress_slot(j)];
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Device-driven traversal: levels are enqueued in batches with no host wait
// between them; the kernels decide each level's direction and the end of the
// traversal themselves (control words), and two levels (one swap of the
// frontier buffers) are recorded once as a hipGraph of kernel nodes and
// replayed.  The host reads the control words once per batch of levels.
// ---------------------------------------------------------------------------
struct MsRunInfo {
  uint32_t levels = 0, push_levels = 0, syncs = 0;
  bool overflow = false;
};

template <int NW>
static int msd_level(eg_graph *g, const BatchView &bv, MsState *st, const Lanes<NW> &active) {
  const uint32_t n = g->n;
  hipStream_t s = st->stream;
  unsigned long long *fs = st->ctr + kCtrFront;
  const uint32_t nb = std::min(blocks_for(n, kBlock), kStatBlocks);
  const PullPlan &pl = *bv.pl;
  uint32_t *spart = reinterpret_cast<uint32_t *>(st->scan_tmp);
  const unsigned long long *qn = st->ctr + 3;
  const double push_max = (double)g->m / (double)g->tune.ms_push_div;
  hipLaunchKernelGGL(k_msd_decide<NW>, dim3(1), dim3(kBlock), 0, s, fs, push_max,
                     bv.sym ? 1 : 0, st->ctl, st->ctr, st->lvl_tab);
  // push: the frontier's vertices with out-edges, their degree scan, the
  // edge push (a push level's edges are at most push_max) and the apply pass
  const uint32_t nwords = (n + 63) / 64;
  hipLaunchKernelGGL(k_msd_compact_bits, dim3(std::min(blocks_for(nwords, kBlock), kStatBlocks)),
                     dim3(kBlock), 0, s, nwords, st->fbits, st->queue, st->ctr + 3, st->ctl);
  hipLaunchKernelGGL(k_msd_qdeg, dim3(256), dim3(256), 0, s, st->queue, qn, bv.ptr, st->qoff,
                     st->ctl);
  hipLaunchKernelGGL(k_msd_scan_a, dim3(kScanParts), dim3(kBlock), 0, s, st->qoff, qn, spart,
                     st->ctl);
  hipLaunchKernelGGL(k_msd_scan_b, dim3(1), dim3(kScanParts), 0, s, st->qoff, qn, spart, st->ctl);
  hipLaunchKernelGGL(k_msd_scan_c, dim3(kScanParts), dim3(kBlock), 0, s, st->qoff, qn, spart,
                     st->ctl);
  const uint32_t pgrid =
      std::min<uint32_t>(blocks_for((uint64_t)push_max + 1, kEdgesPerBlock), kDevGrid);
  hipLaunchKernelGGL(k_ms_push<NW>, dim3(pgrid), dim3(kBlock), 0, s, st->queue, 0u, qn, st->qoff,
                     bv.ptr, bv.ms_idx, st->visit, st->seen, st->pend, active, st->ctl);
  const uint32_t nba = std::min(blocks_for(n, kBlock), kDevGrid);
  hipLaunchKernelGGL(k_ms_apply<NW>, dim3(nba), dim3(kBlock), 0, s, n, bv.ptr, st->pend,
                     st->seen, st->nvisit, st->lvl, st->lvl_tab, active, st->ctr, st->ctl, fs,
                     st->fbits);
  // pull over the plan
  const uint32_t nitems = pl.nseg + pl.nchunks;
  const uint32_t pull_grid = std::max(1u, std::min(nitems, kDevPullGrid));
  if (g->tune.ms_batch == 16)
    hipLaunchKernelGGL((k_ms_pull<16, NW>), dim3(pull_grid), dim3(kBlock), 0, s, n,
                       bv.ptr, bv.ms_idx, pl.chunk_info, pl.seg_info, pl.long_rows, pl.nseg,
                       nitems, st->seg_or, st->seen, st->visit, st->nvisit, st->lvl, st->lvl_tab, active,
                       st->ctr, st->ctl, fs, st->fbits);
  else
    hipLaunchKernelGGL((k_ms_pull<8, NW>), dim3(pull_grid), dim3(kBlock), 0, s, n,
                       bv.ptr, bv.ms_idx, pl.chunk_info, pl.seg_info, pl.long_rows, pl.nseg,
                       nitems, st->seg_or, st->seen, st->visit, st->nvisit, st->lvl, st->lvl_tab,
                       active, st->ctr, st->ctl, fs, st->fbits);
  if (pl.nlong)
    hipLaunchKernelGGL(k_ms_pull_finish<NW>, dim3(blocks_for(pl.nlong, kBlock)), dim3(kBlock), 0,
                       s, n, bv.ptr, pl.long_rows, pl.seg_base, pl.nlong, st->seg_or, st->seen,
                       st->nvisit, st->lvl, st->lvl_tab, active, st->ctr, st->ctl, fs, st->fbits);
  hipLaunchKernelGGL(k_msd_end<NW>, dim3(1), dim3(1), 0, s, st->ctl, st->ctr);
  EG_LAUNCHED();
  std::swap(st->visit, st->nvisit);
  return EG_OK;
}

// Records two levels into st->exec unless the recording already there was
// made with the same plan, buffers, lanes and shapes.
template <int NW>
static int msd_graph(eg_graph *g, const BatchView &bv, MsState *st, const Lanes<NW> &active) {
  uint64_t key[6] = {reinterpret_cast<uint64_t>(bv.pl), reinterpret_cast<uint64_t>(st->visit),
                     reinterpret_cast<uint64_t>(bv.ms_idx),
                     (uint64_t)g->tune.ms_batch | ((uint64_t)g->tune.ms_push_div << 32), 0, 0};
  for (int j = 0; j < NW; j++) key[4] ^= active.w[j] * (0x9E3779B97F4A7C15ull + 2 * j);
  key[5] = (uint64_t)NW;
  if (st->exec && memcmp(key, st->exec_key, sizeof(key)) == 0) return EG_OK;
  if (st->exec) (void)hipGraphExecDestroy(st->exec);
  st->exec = nullptr;
  hipStream_t s = st->stream;
  hipGraph_t graph = nullptr;
  EG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = msd_level<NW>(g, bv, st, active);
  if (rc == EG_OK) rc = msd_level<NW>(g, bv, st, active);
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (rc == EG_OK && ec != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "stream capture failed: %s", hipGetErrorString(ec));
  if (rc == EG_OK && hipGraphInstantiate(&st->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    st->exec = nullptr;
    rc = eg_fail(EG_EDEVICE, "graph instantiation failed");
  }
  if (graph) (void)hipGraphDestroy(graph);
  if (rc == EG_OK) memcpy(st->exec_key, key, sizeof(key));
  return rc;
}

// One traversal of the seeded lanes to its end (or to the u16 level store's
// limit: info->overflow).
template <int NW>
static int msd_run(eg_graph *g, const BatchView &bv, MsState *st, const uint64_t *active_w,
                   MsRunInfo *info) {
  hipStream_t s = st->stream;
  const Lanes<NW> active = to_lanes<NW>(active_w);
  EG_TRY(ms_seg_or(st, *bv.pl));
  const bool graphs = g->tune.graphs != 0;
  if (graphs) EG_TRY(msd_graph<NW>(g, bv, st, active));
  hipLaunchKernelGGL(k_msd_init, dim3(1), dim3(64), 0, s, st->ctl);
  EG_LAUNCHED();
  // statistics of the seeded frontier; every later level's kernels count the
  // frontier they build
  {
    const uint32_t nb = std::min(blocks_for(g->n, kBlock), kStatBlocks);
    unsigned long long *part = st->ctr + kCtrPart;
    hipLaunchKernelGGL(k_ms_stats<NW>, dim3(nb), dim3(kBlock), 0, s, g->n, st->visit, bv.ptr, part,
                       st->ctl, st->fbits);
    hipLaunchKernelGGL(k_ms_stats_sum, dim3(1), dim3(kBlock), 0, s, part, nb,
                       st->ctr + kCtrFront);
    EG_LAUNCHED();
  }
  // levels per batch: the last traversal's depth first (rounded up to a swap
  // cycle), then 4, 8, 16, 32 at a time (a few idle levels at the end of a
  // deep traversal against one host wait per batch)
  uint32_t batch = std::max<uint32_t>(4u, (st->depth_hint + 1) & ~1u);
  const uint32_t *h = reinterpret_cast<const uint32_t *>(st->h);
  for (;;) {
    for (uint32_t l = 0; l < batch; l += 2) {
      if (graphs) {
        EG_HIP(hipGraphLaunch(st->exec, s));
      } else {
        EG_TRY(msd_level<NW>(g, bv, st, active));
        EG_TRY(msd_level<NW>(g, bv, st, active));
      }
    }
    EG_HIP(hipMemcpyAsync(st->h, st->ctl, kCtlWords * 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    info->syncs++;
    if (h[kCtlDone]) break;
    batch = batch < 4 ? 4 : std::min(2 * batch, 32u);
  }
  info->levels += h[kCtlLevels];
  info->push_levels += h[kCtlPushLevels];
  info->overflow = h[kCtlOverflow] != 0;
  st->depth_hint = h[kCtlLevels];
  return EG_OK;
}

int eg_ms_level(eg_graph *g, const BatchView &bv, MsState *st, const uint64_t *active,
                const uint16_t *tab, uint64_t *progress) {
  if (st->nw == 4) return ms_level<4>(g, bv, st, active, tab, progress);
  return st->nw == 2 ? ms_level<2>(g, bv, st, active, tab, progress)
                     : ms_level<1>(g, bv, st, active, tab, progress);
}

int eg_ms_extract_lane(eg_graph *g, const BatchView &bv, MsState *st, uint32_t lane,
                       uint32_t *out) {
  if (lane >= 64 * st->nw) return eg_fail(EG_EINVAL, "lane %u out of range", lane);
  hipLaunchKernelGGL(k_ms_extract, dim3(blocks_for(g->n, 256)), dim3(256), 0, st->stream, g->n, 1u,
                     lane, st->lvl, st->seen, st->nw, bv.perm, out);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_ms_clear_lanes(eg_graph *g, MsState *st, const uint64_t *lanes) {
  uint64_t any = 0;
  for (uint32_t j = 0; j < st->nw; j++) any |= lanes[j];
  if (!any) return EG_OK;
  if (st->nw == 4)
    hipLaunchKernelGGL(k_ms_clear_lanes<4>, dim3(blocks_for(g->n, 256)), dim3(256), 0, st->stream,
                       g->n, to_lanes<4>(lanes), st->seen, st->visit);
  else if (st->nw == 2)
    hipLaunchKernelGGL(k_ms_clear_lanes<2>, dim3(blocks_for(g->n, 256)), dim3(256), 0, st->stream,
                       g->n, to_lanes<2>(lanes), st->seen, st->visit);
  else
    hipLaunchKernelGGL(k_ms_clear_lanes<1>, dim3(blocks_for(g->n, 256)), dim3(256), 0, st->stream,
                       g->n, to_lanes<1>(lanes), st->seen, st->visit);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_ms_reached(eg_graph *g, MsState *st, uint32_t k, const uint32_t *lanes,
                  uint64_t *reached) {
  if (k == 0) return EG_OK;
  if (k > 64 * st->nw) return eg_fail(EG_EINVAL, "at most %u lanes", 64 * st->nw);
  hipStream_t s = st->stream;
  const uint32_t nb = blocks_for(g->n, kBlock) < 64 ? blocks_for(g->n, kBlock) : 64;
  uint32_t h[kMaxLanes];
  memcpy(h, lanes, k * 4);
  EG_HIP(hipMemcpyAsync(st->roots, h, k * 4, hipMemcpyHostToDevice, s));
  unsigned long long *part = st->ctr + kCtrPart, *sums = st->ctr + kCtrPart + 64 * 64 * st->nw;
  hipLaunchKernelGGL(k_ms_reached, dim3(nb, k), dim3(kBlock), 0, s, g->n, st->seen, st->nw,
                     st->roots, part);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_ms_reached_sum, dim3(k), dim3(64), 0, s, part, nb, sums);
  EG_LAUNCHED();
  EG_HIP(hipMemcpyAsync(st->h + 8, sums, (size_t)k * 8, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < k; i++) reached[i] = st->h[8 + i];
  return EG_OK;
}

int eg_ppr_alloc(eg_graph *g, PprState *st) {
  const uint64_t nn = (uint64_t)g->n * 64;
  EG_TRY(eg_dalloc_t(&st->X, nn, "ppr X"));
  EG_TRY(eg_dalloc_t(&st->Y, nn, "ppr Y"));
  EG_TRY(eg_dalloc_t(&st->C, nn, "ppr C"));
  EG_TRY(eg_dalloc_t(&st->Cn, nn, "ppr Cn"));
  EG_TRY(eg_dalloc_t(&st->mask, g->n, "ppr seed mask"));
  EG_TRY(eg_dalloc_t(&st->w, 128, "ppr seed weights"));
  EG_TRY(eg_dalloc_t(&st->delta, 64, "ppr deltas"));
  if (hipHostMalloc(reinterpret_cast<void **>(&st->h_delta), 64 * sizeof(double),
                    hipHostMallocDefault) != hipSuccess)
    return eg_fail(EG_ENOMEM, "pinned deltas");
  memset(st->hw, 0, sizeof(st->hw));
  return EG_OK;
}

void eg_ppr_free(PprState *st) {
  eg_dfree_t(&st->X);
  eg_dfree_t(&st->Y);
  eg_dfree_t(&st->C);
  eg_dfree_t(&st->Cn);
  eg_dfree_t(&st->mask);
  eg_dfree_t(&st->w);
  eg_dfree_t(&st->seg_part);
  eg_dfree_t(&st->seeds);
  eg_dfree_t(&st->dpart);
  eg_dfree_t(&st->delta);
  if (st->h_delta) (void)hipHostFree(st->h_delta);
  st->h_delta = nullptr;
}

int eg_ppr_reset(eg_graph *g, PprState *st) {
  const uint64_t nn = (uint64_t)g->n * 64;
  hipLaunchKernelGGL(k_ppr_zero, dim3(stride_grid(nn)), dim3(256), 0, st->stream, nn, st->X,
                     st->C, st->Y, st->Cn);
  EG_LAUNCHED();
  EG_HIP(hipMemsetAsync(st->mask, 0, (size_t)g->n * 8, st->stream));
  memset(st->hw, 0, sizeof(st->hw));
  st->full_steps = 0;
  return EG_OK;
}

int eg_ppr_col_init(eg_graph *g, const BatchView &bv, PprState *st, uint32_t q,
                    const uint32_t *seeds, uint32_t ns, double alpha) {
  if (q >= 64 || ns == 0) return eg_fail(EG_EINVAL, "bad PPR column / seeds");
  uint32_t distinct = 0;
  for (uint32_t j = 0; j < ns; j++) {
    bool dup = false;
    for (uint32_t i = 0; i < j; i++) dup |= seeds[i] == seeds[j];
    distinct += !dup;
  }
  st->hw[q] = (float)(1.0 / distinct);
  st->hw[64 + q] = (float)(alpha / distinct);
  st->full_steps = 0;
  hipStream_t s = st->stream;
  if (st->seeds_cap < ns) {
    EG_HIP(hipStreamSynchronize(s));
    eg_dfree_t(&st->seeds);
    EG_TRY(eg_dalloc_t(&st->seeds, ns, "ppr seeds"));
    st->seeds_cap = ns;
  }
  EG_HIP(hipMemcpyAsync(st->seeds, seeds, (size_t)ns * 4, hipMemcpyHostToDevice, s));
  EG_HIP(hipMemcpyAsync(st->w, st->hw, sizeof(st->hw), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_ppr_col_seed, dim3(blocks_for(ns, 64)), dim3(64), 0, s, q, st->seeds, ns,
                     bv.perm, st->w, bv.deg, st->mask, st->X, st->C);
  EG_LAUNCHED();
  EG_HIP(hipStreamSynchronize(s));  // seeds / hw may be reused by the caller
  return EG_OK;
}

// Launches k_ppr_spmm for the narrowest slot width that holds `width`
// columns, with the tuned gather depth.
template <bool DELTA, int D, int W>
static void spmm_go(hipStream_t s, const BatchView &bv, const PprState *st, uint32_t grid,
                    uint32_t nitems, float oma, uint32_t width, uint32_t write_y) {
  const PullPlan &pl = *bv.pl;
  hipLaunchKernelGGL((k_ppr_spmm<DELTA, D, W>), dim3(grid), dim3(kBlock), 0, s, bv.idx,
                     pl.chunk_info, pl.seg_info, pl.nseg, nitems, st->seg_part, bv.ptr, st->C,
                     st->mask, st->w + 64, bv.deg, oma, width, write_y, st->X, st->Y, st->Cn,
                     DELTA ? st->dpart : nullptr);
}

template <bool DELTA, int D>
static void spmm_launch_w(hipStream_t s, const BatchView &bv, const PprState *st, uint32_t grid,
                          uint32_t nitems, float oma, uint32_t width, uint32_t write_y) {
  if (width <= 16)
    spmm_go<DELTA, D, 16>(s, bv, st, grid, nitems, oma, width, write_y);
  else if (width <= 32)
    spmm_go<DELTA, D, 32>(s, bv, st, grid, nitems, oma, width, write_y);
  else
    spmm_go<DELTA, D, 64>(s, bv, st, grid, nitems, oma, width, write_y);
}

static int launch_spmm(eg_graph *g, const BatchView &bv, const PprState *st, bool delta,
                       uint32_t grid, uint32_t nitems, float oma, uint32_t width,
                       uint32_t write_y) {
  const bool d16 = g->tune.spmm_depth == 16;
  hipStream_t s = st->stream;
  if (delta && d16)
    spmm_launch_w<true, 16>(s, bv, st, grid, nitems, oma, width, 1u);
  else if (delta)
    spmm_launch_w<true, 8>(s, bv, st, grid, nitems, oma, width, 1u);
  else if (d16)
    spmm_launch_w<false, 16>(s, bv, st, grid, nitems, oma, width, write_y);
  else
    spmm_launch_w<false, 8>(s, bv, st, grid, nitems, oma, width, write_y);
  EG_LAUNCHED();
  return EG_OK;
}

// Gathered lanes: every active column, rounded up to a multiple of 16.
static uint32_t ppr_width(uint64_t active) {
  if (!active) return 16;
  const uint32_t hi = 63u - (uint32_t)__builtin_clzll(active);
  return (hi / 16 + 1) * 16;
}

// Partial-sum buffers of a step (allocated outside any stream capture).
static int ppr_buffers(const PullPlan &pl, PprState *st, bool want_delta) {
  hipStream_t s = st->stream;
  if (st->seg_cap < (uint64_t)pl.nseg * 64) {
    EG_HIP(hipStreamSynchronize(s));
    eg_dfree_t(&st->seg_part);
    EG_TRY(eg_dalloc_t(&st->seg_part, (uint64_t)(pl.nseg ? pl.nseg : 1) * 64, "ppr partials"));
    st->seg_cap = (uint64_t)(pl.nseg ? pl.nseg : 1) * 64;
  }
  const uint32_t nfin = pl.nlong ? blocks_for((uint64_t)pl.nlong * 64, kBlock) : 0;
  const uint32_t np = pl.nchunks + nfin;
  if (want_delta && st->dpart_cap < (uint64_t)np + kDeltaBlocks) {
    EG_HIP(hipStreamSynchronize(s));
    eg_dfree_t(&st->dpart);
    EG_TRY(eg_dalloc_t(&st->dpart, ((uint64_t)np + kDeltaBlocks) * 64, "ppr delta partials"));
    st->dpart_cap = (uint64_t)np + kDeltaBlocks;
  }
  return EG_OK;
}

int eg_ppr_step(eg_graph *g, const BatchView &bv, PprState *st, double alpha, bool want_delta) {
  const PullPlan &pl = *bv.pl;
  hipStream_t s = st->stream;
  EG_TRY(ppr_buffers(pl, st, want_delta));
  const uint32_t nfin = pl.nlong ? blocks_for((uint64_t)pl.nlong * 64, kBlock) : 0;
  const uint32_t np = pl.nchunks + nfin;
  const float oma = (float)(1.0 - alpha);
  const uint32_t width = ppr_width(st->active);
  // after two full steps since the last column change the static chunks hold
  // their final values in both iterate buffers
  const bool skip = st->full_steps >= 2 && bv.static_chunk0 < pl.nchunks;
  const uint32_t nchunks = skip ? bv.static_chunk0 : pl.nchunks;
  const uint32_t nitems = pl.nseg + nchunks;
  const uint32_t grid = st->max_grid && st->max_grid < nitems ? st->max_grid : nitems;
  // Y rows: always with deltas, and in the first two full steps since the
  // last column change (the static chunks' rows keep those values in both
  // buffers); otherwise only unless the caller skips them (st->skip_y)
  const uint32_t write_y = (want_delta || !st->skip_y || st->full_steps < 2) ? 1u : 0u;
  if (!skip) st->full_steps++;
  if (grid) {
    EgProf prof(st->capturing ? nullptr : g, "ppr_spmm", s);
    EG_TRY(launch_spmm(g, bv, st, want_delta, grid, nitems, oma, width, write_y));
  }
  if (pl.nlong) {
    if (want_delta)
      hipLaunchKernelGGL(k_ppr_finish<true>, dim3(nfin), dim3(kBlock), 0, s, pl.long_rows,
                         pl.seg_base, pl.nlong, st->seg_part, st->mask, st->w + 64, bv.deg, oma,
                         width, 1u, st->X, st->Y, st->Cn, st->dpart, pl.nchunks);
    else
      hipLaunchKernelGGL(k_ppr_finish<false>, dim3(nfin), dim3(kBlock), 0, s, pl.long_rows,
                         pl.seg_base, pl.nlong, st->seg_part, st->mask, st->w + 64, bv.deg, oma,
                         width, write_y, st->X, st->Y, st->Cn, nullptr, 0u);
    EG_LAUNCHED();
  }
  if (want_delta) {
    double *part2 = st->dpart + (uint64_t)np * 64;
    hipLaunchKernelGGL(k_ppr_delta_partial, dim3(kDeltaBlocks), dim3(kBlock), 0, s, st->dpart, np,
                       skip ? nchunks : 0u, skip ? pl.nchunks : 0u, part2);
    EG_LAUNCHED();
    hipLaunchKernelGGL(k_ppr_delta_final, dim3(1), dim3(64), 0, s, part2, st->delta);
    EG_LAUNCHED();
    EG_HIP(hipMemcpyAsync(st->h_delta, st->delta, 64 * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  std::swap(st->X, st->Y);
  std::swap(st->C, st->Cn);
  return EG_OK;
}

int eg_ppr_extract_col(eg_graph *g, const BatchView &bv, PprState *st, uint32_t q, float *out) {
  hipLaunchKernelGGL(k_ppr_extract, dim3(blocks_for(g->n, 256)), dim3(256), 0, st->stream, g->n, 1u,
                     q, st->X, bv.perm, out);
  EG_LAUNCHED();
  return EG_OK;
}

int eg_ppr_clear_cols(eg_graph *g, PprState *st, uint64_t cols) {
  if (!cols) return EG_OK;
  hipLaunchKernelGGL(k_ppr_cols_clear, dim3(stride_grid((uint64_t)g->n * 64)), dim3(256), 0,
                     st->stream, g->n, cols, st->mask, st->X, st->C, st->Y, st->Cn);
  EG_LAUNCHED();
  for (int q = 0; q < 64; q++)
    if ((cols >> q) & 1ull) st->hw[q] = st->hw[64 + q] = 0.f;
  st->full_steps = 0;
  return EG_OK;
}

// ---------------------------------------------------------------------------
// The batched API.
// ---------------------------------------------------------------------------
struct BatchState {
  MsState ms, ms2, ms4;  // one, two or four lane words (64, 128, 256 lanes per traversal)
  PprState ppr;
  bool ms_ok = false, ms2_ok = false, ms4_ok = false, ppr_ok = false;
  uint32_t *ext = nullptr;  // k x n extraction buffer
  uint64_t ext_cap = 0;
  eg_bfs_batch_info last{};  // of the last eg_bfs_batch call
};

void eg_batch_free(eg_graph *g) {
  BatchState *b = g->batch;
  if (!b) return;
  eg_ms_free(&b->ms);
  eg_ms_free(&b->ms2);
  eg_ms_free(&b->ms4);
  eg_ppr_free(&b->ppr);
  eg_dfree_t(&b->ext);
  delete b;
  g->batch = nullptr;
}

static BatchState *batch_state(eg_graph *g) {
  if (!g->batch) g->batch = new BatchState();
  return g->batch;
}

static int ensure_ext(BatchState *b, uint64_t words) {
  if (b->ext_cap < words) {
    eg_dfree_t(&b->ext);
    EG_TRY(eg_dalloc_t(&b->ext, words, "extract"));
    b->ext_cap = words;
  }
  return EG_OK;
}

// The host-driven form of one traversal (the serving engine's level loop):
// two host waits per level.
template <int NW>
static int ms_run_host(eg_graph *g, const BatchView &bv, MsState *st, const uint64_t *active,
                       MsRunInfo *info) {
  uint16_t tab[kMaxLanes];
  for (int level = 2;; level++) {
    if ((uint32_t)level > kMaxLevelValue) {
      info->overflow = true;
      return EG_OK;
    }
    for (uint32_t i = 0; i < kMaxLanes; i++) tab[i] = (uint16_t)level;
    uint64_t progress[4] = {0, 0, 0, 0};
    EG_TRY(ms_level<NW>(g, bv, st, active, tab, progress));
    info->levels++;
    info->push_levels += st->pushed ? 1u : 0u;
    info->syncs += 2;
    if ((progress[0] | progress[1] | progress[2] | progress[3]) == 0) return EG_OK;
  }
}

extern "C" int eg_bfs_batch(eg_graph *g, uint32_t k, const uint32_t *roots, uint32_t *dist,
                            double *device_ms) {
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "batched BFS needs adjacency lists (modes 2/3/4)");
  if (k && !roots) return eg_fail(EG_EINVAL, "null roots");
  for (uint32_t i = 0; i < k; i++)
    if (roots[i] >= g->n) return eg_fail(EG_ERANGE, "root %u >= n", roots[i]);
  // lanes per traversal: the fewest lane words that hold k sources, at most
  // ms_lanes (a 16-byte pull gather fetches two words at once)
  const uint32_t cap = (uint32_t)g->tune.ms_lanes;
  const uint32_t want = k > 128 ? 256u : k > 64 ? 128u : 64u;
  const uint32_t per = want < cap ? want : cap;
  const uint32_t nw = per / 64;
  BatchState *b = batch_state(g);
  MsState *mst = nw == 4 ? &b->ms4 : nw == 2 ? &b->ms2 : &b->ms;
  bool &ok = nw == 4 ? b->ms4_ok : nw == 2 ? b->ms2_ok : b->ms_ok;
  if (!ok) {
    EG_TRY(eg_ms_alloc(g, mst, nw));
    mst->stream = g->stream;
    ok = true;
  }
  BatchView bv;
  EG_TRY(eg_batch_view(g, &bv));
  hipStream_t s = g->stream;
  const uint32_t n = g->n;
  if (dist) EG_TRY(ensure_ext(b, (uint64_t)n * 64));
  eg_bfs_batch_info info{};
  info.lanes = per;
  double total_ms = 0.0;
  uint32_t lanes[kMaxLanes];
  for (uint32_t i = 0; i < kMaxLanes; i++) lanes[i] = i;
  for (uint32_t g0 = 0; g0 < k; g0 += per) {
    const uint32_t kk = k - g0 < per ? k - g0 : per;
    hipEvent_t e0, e1;
    EG_HIP(hipEventCreate(&e0));
    EG_HIP(hipEventCreate(&e1));
    int rc = eg_ms_reset(g, mst);
    if (rc == EG_OK) rc = eg_ms_seed(g, bv, mst, kk, lanes, roots + g0);
    uint64_t active[4];
    for (uint32_t j = 0; j < 4; j++) {
      const uint32_t c = kk > 64 * j ? kk - 64 * j : 0;
      active[j] = c >= 64 ? ~0ull : ((1ull << c) - 1ull);
    }
    MsRunInfo run;
    if (rc == EG_OK) rc = hipEventRecord(e0, s) == hipSuccess ? EG_OK : eg_fail(EG_EDEVICE, "event");
    if (rc == EG_OK) {
      if (g->tune.ms_device)
        rc = nw == 4 ? msd_run<4>(g, bv, mst, active, &run)
             : nw == 2 ? msd_run<2>(g, bv, mst, active, &run)
                       : msd_run<1>(g, bv, mst, active, &run);
      else
        rc = nw == 4 ? ms_run_host<4>(g, bv, mst, active, &run)
             : nw == 2 ? ms_run_host<2>(g, bv, mst, active, &run)
                       : ms_run_host<1>(g, bv, mst, active, &run);
    }
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    total_ms += ms;
    info.traversals++;
    info.levels += run.levels;
    info.push_levels += run.push_levels;
    info.host_syncs += run.syncs;
    if (run.overflow) {
      // deeper than the u16 level store: these sources run one by one on the
      // single-source path, whose levels have no such limit
      info.fallbacks++;
      for (uint32_t i = 0; i < kk; i++) {
        eg_bfs_stats bst;
        EG_TRY(eg_bfs_host(g, roots[g0 + i], 0, 1, dist ? dist + (uint64_t)(g0 + i) * n : nullptr,
                           nullptr, &bst));
        total_ms += bst.device_ms;
      }
      continue;
    }
    // 64 lanes at a time through the n x 64 extraction buffer
    for (uint32_t l0 = 0; dist && l0 < kk; l0 += 64) {
      const uint32_t kl = kk - l0 < 64 ? kk - l0 : 64;
      hipLaunchKernelGGL(k_ms_extract, dim3(stride_grid((uint64_t)n * kl)), dim3(256), 0, s, n,
                         kl, l0, mst->lvl, mst->seen, mst->nw, bv.perm, b->ext);
      EG_LAUNCHED();
      EG_HIP(hipMemcpyAsync(dist + (uint64_t)(g0 + l0) * n, b->ext, (uint64_t)n * kl * 4,
                            hipMemcpyDeviceToHost, s));
      EG_HIP(hipStreamSynchronize(s));
    }
  }
  info.device_ms = total_ms;
  b->last = info;
  if (device_ms) *device_ms = total_ms;
  return EG_OK;
}

extern "C" int eg_bfs_batch_info_get(eg_graph *g, eg_bfs_batch_info *out) {
  if (!g || !out) return eg_fail(EG_EINVAL, "null argument");
  if (!g->batch) return eg_fail(EG_ESTATE, "no batched BFS has run on this graph");
  *out = g->batch->last;
  return EG_OK;
}

// Runs 2 * pairs SpMM iterations by replaying a hipGraph of two iterations
// (one full X/Y, C/Cn swap cycle, so every replay sees the same buffers).
// The captured steps hold only kernel launches.
static int ppr_graph_steps(eg_graph *g, const BatchView &bv, PprState *st, double alpha,
                           int pairs) {
  EG_TRY(ppr_buffers(*bv.pl, st, false));
  hipStream_t s = st->stream;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  EG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  st->capturing = true;
  int rc = eg_ppr_step(g, bv, st, alpha);
  if (rc == EG_OK) rc = eg_ppr_step(g, bv, st, alpha);
  st->capturing = false;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (rc == EG_OK && ec != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "stream capture failed: %s", hipGetErrorString(ec));
  if (rc == EG_OK && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "graph instantiation failed");
  for (int p = 0; p < pairs && rc == EG_OK; p++)
    if (hipGraphLaunch(exec, s) != hipSuccess) rc = eg_fail(EG_EDEVICE, "graph launch failed");
  if (rc == EG_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "graph replay failed");
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return rc;
}

extern "C" int eg_ppr_batch(eg_graph *g, uint32_t k, const uint32_t *seed_ptr,
                            const uint32_t *seeds, double alpha, int32_t iterations, float *out,
                            double *device_ms) {
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "batched PPR needs adjacency lists (modes 2/3/4)");
  if (k && (!seed_ptr || !seeds)) return eg_fail(EG_EINVAL, "null seeds");
  if (!(alpha >= 0.0 && alpha <= 1.0) || iterations < 0)
    return eg_fail(EG_EINVAL, "bad PPR options");
  for (uint32_t q = 0; q < k; q++) {
    if (seed_ptr[q + 1] <= seed_ptr[q]) return eg_fail(EG_EINVAL, "query %u has no seeds", q);
    for (uint32_t j = seed_ptr[q]; j < seed_ptr[q + 1]; j++)
      if (seeds[j] >= g->n) return eg_fail(EG_ERANGE, "seed %u >= n", seeds[j]);
  }
  BatchState *b = batch_state(g);
  if (!b->ppr_ok) {
    EG_TRY(eg_ppr_alloc(g, &b->ppr));
    b->ppr.stream = g->stream;
    b->ppr_ok = true;
  }
  BatchView bv;
  EG_TRY(eg_batch_view(g, &bv));
  hipStream_t s = g->stream;
  const uint32_t n = g->n;
  if (out) EG_TRY(ensure_ext(b, (uint64_t)n * 64));
  double total_ms = 0.0;
  for (uint32_t q0 = 0; q0 < k; q0 += 64) {
    const uint32_t kk = k - q0 < 64 ? k - q0 : 64;
    hipEvent_t e0, e1;
    EG_HIP(hipEventCreate(&e0));
    EG_HIP(hipEventCreate(&e1));
    EG_HIP(hipEventRecord(e0, s));
    int rc = eg_ppr_reset(g, &b->ppr);
    b->ppr.active = kk == 64 ? ~0ull : ((1ull << kk) - 1ull);
    for (uint32_t q = 0; q < kk && rc == EG_OK; q++)
      rc = eg_ppr_col_init(g, bv, &b->ppr, q, seeds + seed_ptr[q0 + q],
                           seed_ptr[q0 + q + 1] - seed_ptr[q0 + q], alpha);
    // Only the last step's Y is read (the output); the steps before it write
    // Cn alone (eg_ppr_step still writes Y in the first two full steps).
    int it = 0;
    if (rc == EG_OK && g->tune.graphs && iterations >= 6) {
      // two eager steps first, so that the recorded steps skip static chunks
      for (; it < 2 && rc == EG_OK; it++) {
        b->ppr.skip_y = true;
        rc = eg_ppr_step(g, bv, &b->ppr, alpha);
      }
      // replayed pairs, leaving at least one eager step to write Y
      const int rest = iterations - it;
      const int pairs = rest % 2 ? rest / 2 : rest / 2 - 1;
      if (rc == EG_OK && pairs > 0) {
        b->ppr.skip_y = true;
        rc = ppr_graph_steps(g, bv, &b->ppr, alpha, pairs);
        it += 2 * pairs;
      }
    }
    for (; it < iterations && rc == EG_OK; it++) {
      b->ppr.skip_y = it + 1 < iterations;
      rc = eg_ppr_step(g, bv, &b->ppr, alpha);
    }
    b->ppr.skip_y = false;
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    total_ms += ms;
    if (out) {
      float *ext = reinterpret_cast<float *>(b->ext);
      hipLaunchKernelGGL(k_ppr_extract, dim3(stride_grid((uint64_t)n * kk)), dim3(256), 0, s,
                         n, kk, 0u, b->ppr.X, bv.perm, ext);
      EG_LAUNCHED();
      EG_HIP(hipMemcpyAsync(out + (uint64_t)q0 * n, ext, (uint64_t)n * kk * 4,
                            hipMemcpyDeviceToHost, s));
      EG_HIP(hipStreamSynchronize(s));
    }
  }
  if (device_ms) *device_ms = total_ms;
  return EG_OK;
}
