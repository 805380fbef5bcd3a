// eg_binned.hip -- binned (two-phase) PageRank over an adjacency orientation.
//
// Reference: the pull iteration of pagerank_simple.c:72-84 / :110-133
// (rank[v] = (1-d)/n + d * sum_{u in in(v)} prev[u]/deg(u)).  This computes
// the same recurrence with another data movement.
//
// The pull kernel (eg_pull.hip) makes one random 4-byte gather of contrib[u]
// per stored entry; on R-MAT graphs that is bound by the device's random
// request rate, not by HBM bandwidth.  The binned form turns every access
// into a stream:
//  * destination rows are cut into bins (at most kBinRows rows, balanced by
//    entries), sources into chunks (at most kChunkSrc sources, balanced by
//    entries);
//  * the entries are stored twice, as the u16 source offset within their
//    chunk in chunk-major order (chunk, bin, row, entry), and as the u16 row
//    offset within their bin in bin-major order (bin, chunk, row, entry); a
//    (chunk, bin) segment is contiguous in both orders;
//  * phase 1 (k_pb_scatter): a workgroup stages its chunk's contributions in
//    LDS, streams the chunk's source offsets and writes every entry's value
//    contrib[u] to its bin-major slot: coalesced writes, LDS reads;
//  * phase 2 (k_pb_accum): a workgroup streams one bin's values and row
//    offsets and adds them into per-row accumulators in LDS, then applies the
//    PageRank update to the bin's rows.  Single rows with more entries than a
//    bin takes are cut into parts summed by separate workgroups.
// Sums are exact: every FP32 term (a contribution, or a product w * x for
// SpMV) is converted to a signed 64-bit fixed-point integer with a scale 2^S
// chosen so that no row sum can overflow, and integer addition is
// associative, so the result does not depend on the order of the entries, on
// the bin/chunk cut or on scheduling.  The row sum is rounded to FP32 once at
// the end.
//
// The entries come from an adjacency orientation (row pointers + sources) or
// from an edge list (destination and source of every entry: the 2-D grid of
// load modes 0/1/6 and the edge array of modes 7/8/9), so the same kernels
// serve pagerank_simple, prgrid (prgrid_cilk.c:62-141), pr_algo_edgecentric
// (pagerank_simple.c:40-61) and spmv (spmv.c:98-136).
//
// Traffic per iteration, stored entries m: 2m (source offsets) + 4m (values
// written) + 4m (values read) + 2m (row offsets) + the per-vertex vectors,
// about 12m + 20n bytes, all streamed.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <cfloat>
#include <cmath>

#include "eg_internal.h"
#include "eg_pr.h"

namespace {

constexpr uint32_t kBinRows = 8192;    // rows per bin (u64 accumulators in LDS: 64 KB)
constexpr uint32_t kChunkSrc = 12288;  // sources per chunk (FP32 window in LDS: 48 KB)
constexpr uint32_t kTileBins = 2048;   // bins one scatter tile may span (table in LDS: 16 KB)
constexpr uint32_t kTileLen = 32768;   // positions per scatter tile
constexpr uint32_t kItemBins = 4;      // entry targets of bins per phase-2 work item
constexpr int kBlock = 1024;           // two workgroups per CU (64 KB LDS each)
constexpr int kWaves = kBlock / 64;
constexpr int kWinPer = (kChunkSrc + kBlock - 1) / kBlock;     // window loads per thread
constexpr int kTabPer = (kTileBins + 1 + kBlock - 1) / kBlock;  // table loads per thread
constexpr int kRowPer = (kBinRows + kBlock - 1) / kBlock;       // rows finished per thread
constexpr int kRowBatch = 4;                                    // of which loaded together
constexpr int kFinBlocks = 128;
// Hot sources of one window, kept in LDS by phase 2 (95 KB): what fits in
// 160 KB beside the accumulators (64 KB + a dummy slot per lane) and the
// workgroup's reduction slots.
constexpr uint32_t kHotMax = kPbHotWindow;
constexpr uint32_t kHotWins = kPbHotWindows;                       // hot windows at most
constexpr int kHotPer4 = (kHotMax / 4 + kBlock - 1) / kBlock;      // hot 16-byte loads per thread
static_assert(kHotMax % 4 == 0, "a hot window is copied in 16-byte pieces");
static_assert(kHotMax < 65536, "a hot entry packs its source offset in 16 bits");
static_assert((kBinRows + 64) * 8 + kHotMax * 4 + kWaves * 24 + 16 <= 160 * 1024,
              "phase 2's LDS must fit one workgroup per CU");

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

// ---------------------------------------------------------------------------
// Fixed point: value * 2^S as an unsigned 64-bit integer.
// ---------------------------------------------------------------------------
// |c| * 2^S rounded to the nearest integer, ties to even, negated for c < 0
// as a two's complement u64 (SIGNED; PageRank terms are never negative).
// Every value that is converted is below the bound B with B * 2^S <= 2^62.
// With FP32 operations only: x = |c| * 2^S is exact (a power-of-two scale);
// h = floor(x / 2^32) < 2^30 is exact; l = x - h * 2^32 (one fused
// multiply-add) is exact, since for x >= 2^32 it keeps the low bits of a
// 24-bit mantissa whose last bit is at 2^9 or above, and for x < 2^32 it is
// x; rint(l) rounds it, and h:rint(l) is the rounded product.  (Checked
// against integer rounding of mantissa * 2^(exponent + S) for every finite
// float below 2^(62 - S) at S = 0, 1, 5, 20, 24, 31, 32, 33, 40, 50, 61, 62.)
template <bool SIGNED>
__device__ __forceinline__ uint64_t to_fixed(float c, int S) {
  const float x = ldexpf(SIGNED ? fabsf(c) : c, S);
  const float h = floorf(x * 0x1p-32f);
  const float l = fmaf(-h, 0x1p32f, x);
  const uint64_t mag = ((uint64_t)(uint32_t)h << 32) | (uint32_t)rintf(l);
  return SIGNED && c < 0.f ? 0ull - mag : mag;
}

__device__ __forceinline__ float from_fixed(uint64_t v, int S) {
  return (float)ldexp((double)(int64_t)v, -S);
}

struct PbTile {
  uint32_t c;       // chunk
  uint32_t p0, p1;  // chunk-major positions [p0, p1)
  uint32_t b0;      // first bin (segment of p0)
  uint32_t nb;      // bins spanned
  uint32_t pad[3];
};

// Phase-2 work item: nb consecutive bins, rows [row0, row0 + nrows) with
// bin-major entries [e0, e1); nrows == 0 marks a part of a long row whose sum
// goes to part_sum[row0].  With s = nwin + 1, bin j of the item has its hot
// entries of window w at [wseg[ws + j s + w], wseg[ws + j s + w + 1]) and its
// scattered entries at [wseg[ws + j s + nwin], wseg[ws + (j + 1) s]) (the
// last bin's end is wseg[ws + nb s] = e1).
struct PbItem {
  uint32_t e0, e1, row0, nrows, ws, nb, pad[2];
};

struct PbDev {
  const PbTile *tiles;
  const uint2 *chunk_src;  // (first source, sources)
  const uint2 *tab;        // [nchunks][nbins + 1]: (chunk-major start, bin-major start)
  const uint16_t *sloc;    // chunk-major source offsets
  const uint16_t *dloc;    // bin-major row offsets
  float *val;              // bin-major values
  const PbItem *items;
  uint32_t nrun;           // phase 2 sums items [0, nrun) (the others keep their statistics)
  const uint4 *long_rows;  // (row, first part, parts, 0)
  unsigned long long *part_sum;
  double2 *item_part;
  double2 *fin_part;
  uint32_t *fin_counter;
  uint32_t nbins, nitems, nlong;
  uint32_t nhot;           // sources below this id are hot
  uint32_t win;            // hot window w holds sources [w * win, (w + 1) * win)
  uint32_t nwin;           // hot windows
  const uint32_t *wseg;    // per item: the nwin + 1 starts of its windows' hot entries
  uint32_t persistent;     // phase 2: workgroups loop over items (else one each)
  uint32_t *work;          // persistent phase 2: items handed out, workgroups done (reset
                           // to 0 by the last workgroup)
  uint32_t xcd_tiles;      // phase 1: interleaved walk of the tile list (else launch order)
  const float *w;          // bin-major weights (SpMV of a weighted graph), or null
};

// ---------------------------------------------------------------------------
// Build kernels.
// ---------------------------------------------------------------------------
// Where the entries come from: rows of an adjacency orientation (ptr, idx)
// or an edge list (row[k], idx[k]); w (may be null): the weight of entry k.
struct PbSrcDev {
  const uint32_t *ptr, *row, *idx;
  const float *w;
};

__device__ __forceinline__ void pb_entry_keys(uint32_t v, uint32_t k, const PbSrcDev &src,
                                              const uint32_t *__restrict__ bin_of,
                                              const uint32_t *__restrict__ chunk_of,
                                              const uint32_t *__restrict__ bin_row0,
                                              const uint32_t *__restrict__ chunk_src0,
                                              uint32_t nbins, uint32_t nchunks, uint32_t nhot,
                                              uint32_t win, uint32_t nwin,
                                              int bin_major, uint32_t *__restrict__ key,
                                              uint32_t *__restrict__ pay,
                                              uint32_t *__restrict__ pay2,
                                              uint32_t *__restrict__ pay3) {
  const uint32_t u = src.idx[k];
  const uint32_t b = bin_of[v];
  const bool hot = u < nhot;
  const uint32_t c = hot ? 0u : chunk_of[u];
  if (bin_major) {
    key[k] = b * (nchunks + nwin) + (hot ? u / win : nwin + c);
    pay[k] = v - bin_row0[b];
    pay2[k] = u;
    if (pay3) pay3[k] = __float_as_uint(src.w[k]);
  } else {
    key[k] = hot ? nchunks * nbins : c * nbins + b;
    pay[k] = hot ? 0u : u - chunk_src0[c];
  }
}

// key / payloads of every entry.  Sources below nhot are hot: they belong to
// no chunk but to the hot window source / win (nwin windows).
// bin_major: key = bin * (nchunks + nwin) + (hot ? window : nwin + chunk),
//   payloads = row - first row of the bin, the source, and the weight bits
//   (pay3);
// else key = hot ? nchunks * nbins : chunk * nbins + bin, payload = source -
//   first source of the chunk (hot entries sort last and are not used).
// Adjacency source: a wave per row, grid-stride; edge list: a thread per entry.
__global__ void k_pb_keys(uint32_t n, uint64_t m, PbSrcDev src,
                          const uint32_t *__restrict__ bin_of,
                          const uint32_t *__restrict__ chunk_of,
                          const uint32_t *__restrict__ bin_row0,
                          const uint32_t *__restrict__ chunk_src0, uint32_t nbins,
                          uint32_t nchunks, uint32_t nhot, uint32_t win, uint32_t nwin,
                          int bin_major, uint32_t *__restrict__ key, uint32_t *__restrict__ pay,
                          uint32_t *__restrict__ pay2, uint32_t *__restrict__ pay3) {
  if (src.row) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride)
      pb_entry_keys(src.row[k], (uint32_t)k, src, bin_of, chunk_of, bin_row0, chunk_src0, nbins,
                    nchunks, nhot, win, nwin, bin_major, key, pay, pay2, pay3);
    return;
  }
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const int lane = __lane_id();
  for (uint32_t v = wave; v < n; v += nwaves) {
    const uint32_t k0 = src.ptr[v], k1 = src.ptr[v + 1];
    for (uint32_t k = k0 + lane; k < k1; k += 64)
      pb_entry_keys(v, k, src, bin_of, chunk_of, bin_row0, chunk_src0, nbins, nchunks, nhot,
                    win, nwin, bin_major, key, pay, pay2, pay3);
  }
}

// Largest |w| as float bits (non-negative floats order like their bits).
__global__ void k_pb_absmax(uint64_t m, const float *__restrict__ w, uint32_t *__restrict__ out) {
  uint32_t mx = 0u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += stride)
    mx = max(mx, __float_as_uint(w[k]) & 0x7FFFFFFFu);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o));
  if (__lane_id() == 0) atomicMax(out, mx);
}

// u16 offsets of the sorted order; for the bin-major order also the packed
// (source within its window | row offset << 16) value slots of the hot
// entries, whose keys modulo key_mod are below nwin.
__global__ void k_pb_narrow(uint64_t m, const uint32_t *__restrict__ in,
                            uint16_t *__restrict__ out, const uint32_t *__restrict__ skey,
                            const uint32_t *__restrict__ src, uint32_t key_mod, uint32_t nwin,
                            uint32_t win, uint32_t *__restrict__ packed) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  out[k] = (uint16_t)in[k];
  if (packed && skey[k] % key_mod < nwin) packed[k] = (src[k] % win) | (in[k] << 16);
}

// ---------------------------------------------------------------------------
// Phase 1: values into bin-major slots.
// ---------------------------------------------------------------------------
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));

// The source offsets are loaded four per lane (8-byte loads; 2-byte loads
// stream at well under half the rate) and passed through a per-wave LDS
// stage, so that the value stores stay lane-contiguous.  Quad j of the step
// at a holds positions a + 256 j + 4 lane .. + 3; quads starting at or after
// w1 are not loaded (a quad that starts below w1 < m ends within the
// allocation, which has 16 spare entries).
template <int Q>
__device__ __forceinline__ void load_quads(const PbDev &d, uint32_t a, uint32_t w1, int lane,
                                           u16x4 (&s)[Q]) {
  const u16x4 *q4 = reinterpret_cast<const u16x4 *>(d.sloc);
#pragma unroll
  for (int j = 0; j < Q; j++) {
    const uint32_t p = a + 256 * j + 4 * lane;
    s[j] = p < w1 ? __builtin_nontemporal_load(q4 + (p >> 2)) : u16x4{0, 0, 0, 0};
  }
}

// Writes the values of positions a + 256 j + 64 q + lane within [w0, w1).
template <int Q>
__device__ __forceinline__ void write_quads(const PbDev &d, uint32_t a, uint32_t w0, uint32_t w1,
                                            int lane, const u16x4 (&s)[Q], uint32_t &bi,
                                            const float *s_win, const uint2 *s_tab,
                                            u16x4 *stage) {
  const uint16_t *st = reinterpret_cast<const uint16_t *>(stage);
#pragma unroll
  for (int j = 0; j < Q; j++) {
    // LDS operations of a wave complete in order: the reads below see this
    // store, and the next store follows them
    stage[lane] = s[j];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t p = a + 256 * j + 64 * q + lane;
      const uint32_t o = st[64 * q + lane];
      if (p >= w0 && p < w1) {
        while (p >= s_tab[bi + 1].x) bi++;
        d.val[s_tab[bi].y + (p - s_tab[bi].x)] = s_win[o];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// B: positions per lane per step (B / 4 quads); the next step's offsets are
// loaded before the current step's values are written.
template <int B>
__global__ __launch_bounds__(kBlock) void k_pb_scatter(PbDev d, const float *__restrict__ contrib) {
  constexpr int Q = B / 4;
  __shared__ float s_win[kChunkSrc];
  __shared__ uint2 s_tab[kTileBins + 1];
  __shared__ u16x4 s_stage[kWaves][64];
  // Workgroups are dealt to the eight XCDs round-robin (workgroup b runs on
  // XCD b % 8), and the build stores the tile list so that launch order gives
  // XCD x the x-th eighth of the chunk-ordered tiles.  With xcd_tiles,
  // workgroup b takes list entry x q + min(x, r) + b / 8 instead (T = 8 q + r
  // tiles): XCD x then walks one short run of tiles in every eighth of the
  // chunk order by turns, so each XCD's work samples the whole source range
  // (64 write regions in flight instead of 8), while the tiles of one chunk
  // still share one XCD and its L2 copy of the chunk's window.
  uint32_t ti = blockIdx.x;
  if (d.xcd_tiles) {
    const uint32_t q = gridDim.x >> 3, r = gridDim.x & 7, x = ti & 7;
    ti = x * q + min(x, r) + (ti >> 3);
  }
  const PbTile t = d.tiles[ti];
  const uint2 cs = d.chunk_src[t.c];
  const uint32_t tid = threadIdx.x;
  // the window and the table: every load issued before the first LDS store
  {
    const uint2 *trow = d.tab + (size_t)t.c * (d.nbins + 1) + t.b0;
    float wv[kWinPer];
    uint2 tv[kTabPer];
#pragma unroll
    for (int j = 0; j < kWinPer; j++) {
      const uint32_t i = tid + j * kBlock;
      wv[j] = i < cs.y ? contrib[cs.x + i] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kTabPer; j++) {
      const uint32_t i = tid + j * kBlock;
      tv[j] = i <= t.nb ? trow[i] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int j = 0; j < kWinPer; j++) {
      const uint32_t i = tid + j * kBlock;
      if (i < cs.y) s_win[i] = wv[j];
    }
#pragma unroll
    for (int j = 0; j < kTabPer; j++) {
      const uint32_t i = tid + j * kBlock;
      if (i <= t.nb) s_tab[i] = tv[j];
    }
  }
  __syncthreads();

  // a contiguous range of positions per wave
  const int lane = __lane_id();
  const uint32_t wave = tid >> 6;
  const uint32_t len = t.p1 - t.p0;
  const uint32_t per = ((len + kWaves - 1) / kWaves + 63) & ~63u;
  const uint32_t w0 = t.p0 + min(len, wave * per);
  const uint32_t w1 = t.p0 + min(len, (wave + 1) * per);
  if (w0 >= w1) return;
  const uint32_t a0 = w0 & ~3u;  // quads are 8-byte aligned
  // segment of this lane's first position: the last bin whose start is <= p
  uint32_t bi = 0;
  {
    const uint32_t p = min(max(a0 + lane, w0), w1 - 1);
    uint32_t lo = 0, hi = t.nb - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (s_tab[mid].x <= p) lo = mid; else hi = mid - 1;
    }
    bi = lo;
  }
  constexpr uint32_t kStep = 256 * Q;
  u16x4 *stage = s_stage[wave];
  u16x4 s0[Q], s1[Q];
  load_quads<Q>(d, a0, w1, lane, s0);
  for (uint32_t a = a0;; a += 2 * kStep) {
    const bool more1 = a + kStep < w1;
    if (more1) load_quads<Q>(d, a + kStep, w1, lane, s1);
    write_quads<Q>(d, a, w0, w1, lane, s0, bi, s_win, s_tab, stage);
    if (!more1) break;
    const bool more2 = a + 2 * kStep < w1;
    if (more2) load_quads<Q>(d, a + 2 * kStep, w1, lane, s0);
    write_quads<Q>(d, a + kStep, w0, w1, lane, s1, bi, s_win, s_tab, stage);
    if (!more2) break;
  }
}

// ---------------------------------------------------------------------------
// Phase 2: per-bin accumulation and the PageRank update.
// ---------------------------------------------------------------------------
// A lane takes groups of 8 consecutive entries, sums runs of equal rows in
// registers and adds each run to its LDS accumulator with one atomic.  Entries
// of a segment are in row order, so a hub row's entries form long runs and
// cost about one atomic per group.  Groups are aligned to 8 entries; entries
// of a group outside the range being summed are masked.
//  * cold entries: one 16-byte load of row offsets, two of values;
//  * hot entries (sources below the hot threshold, first in every bin): their
//    value slot holds (source | row offset << 16) from the build, and the
//    value is read from the LDS copy of the hot sources' contributions, so
//    phase 1 never touches them (two 16-byte loads per group).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Adds the 8 entries of the group that starts at entry e (those outside
// [e0, e1) masked): HOTRANGE, w0 / w1 hold packed (source | row offset << 16)
// words; otherwise w0 holds 8 u16 row offsets and v0 / v1 the values.  part:
// the whole range belongs to one row and is summed into psum.  A run of equal
// row offsets is added with one atomic at its last entry.
template <bool HOTRANGE, bool WEIGHTED, bool SIGNED>
__device__ __forceinline__ void group_accumulate(unsigned long long *acc, const u32x4 w0,
                                                 const u32x4 w1, const f32x4 v0, const f32x4 v1,
                                                 const f32x4 x0, const f32x4 x1,
                                                 uint32_t e, uint32_t e0, uint32_t e1, int S,
                                                 bool part, const float *s_hot,
                                                 unsigned long long &psum) {
  uint32_t key[8];
  float v[8];
  if (HOTRANGE) {
    const uint32_t w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int j = 0; j < 8; j++) {
      key[j] = w[j] >> 16;
      // masked slots may hold other words: keep the read inside the copy
      v[j] = s_hot[min(w[j] & 0xFFFFu, kHotMax - 1)];
    }
  } else {
    const uint32_t ks[8] = {w0.x & 0xFFFFu, w0.x >> 16, w0.y & 0xFFFFu, w0.y >> 16,
                            w0.z & 0xFFFFu, w0.z >> 16, w0.w & 0xFFFFu, w0.w >> 16};
    const float vs[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
    for (int j = 0; j < 8; j++) {
      key[j] = ks[j];
      v[j] = vs[j];
    }
  }
  if (WEIGHTED) {  // SpMV of a weighted graph: the FP32 product w * x
    const float ws8[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] *= ws8[j];
  }
  // This lane's dummy slot behind the accumulators (never read): masked
  // entries and the inner entries of runs add there.
  const uint32_t dummy = kBinRows + (uint32_t)__lane_id();
  if (e < e0 || e + 8 > e1) {  // an edge group of the range
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const bool ok = e + j >= e0 && e + j < e1;
      key[j] = ok ? key[j] : dummy;
      v[j] = ok ? v[j] : 0.f;
    }
  }
  if (part) {
#pragma unroll
    for (int j = 0; j < 8; j++) psum += to_fixed<SIGNED>(v[j], S);
    return;
  }
  // Branch-free: every entry issues one atomic; the last entry of a run adds
  // the run to its row, the others add to the dummy slot, so no lane waits on
  // a branch and dummies never collide.  A masked entry's key is the dummy,
  // which no row offset equals, so it ends the run before it.
  unsigned long long run = 0ull;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    run += to_fixed<SIGNED>(v[j], S);
    const bool end = j == 7 || key[j] != key[j + 1];
    atomicAdd(&acc[end ? key[j] : dummy], run);
    run = end ? 0ull : run;
  }
}

// G groups of 8 entries per lane, as loaded from HBM.
template <int G>
struct PbBatch {
  u32x4 wa[G], wb[G];
  f32x4 va[G], vb[G], xa[G], xb[G];
};

// Issues the loads of the groups base + q * 64 + lane (q < G) below wg1.
template <bool HOTRANGE, bool WEIGHTED, int G>
__device__ __forceinline__ void load_batch(const PbDev &d, uint32_t base, uint32_t wg1, int lane,
                                           PbBatch<G> &b) {
  const u32x4 *dl4 = reinterpret_cast<const u32x4 *>(d.dloc);
  const f32x4 *v4 = reinterpret_cast<const f32x4 *>(d.val);
  const u32x4 *h4 = reinterpret_cast<const u32x4 *>(d.val);
  const f32x4 *w4 = reinterpret_cast<const f32x4 *>(d.w);
  const u32x4 z4 = u32x4{0u, 0u, 0u, 0u};
  const f32x4 zf = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < G; q++) {
    const uint32_t gi = base + q * 64 + lane;
    const bool in = gi < wg1;
    if (HOTRANGE) {
      b.wa[q] = in ? __builtin_nontemporal_load(h4 + 2 * (size_t)gi) : z4;
      b.wb[q] = in ? __builtin_nontemporal_load(h4 + 2 * (size_t)gi + 1) : z4;
      b.va[q] = b.vb[q] = zf;
    } else {
      b.wa[q] = in ? __builtin_nontemporal_load(dl4 + gi) : z4;
      b.wb[q] = z4;
      b.va[q] = in ? __builtin_nontemporal_load(v4 + 2 * (size_t)gi) : zf;
      b.vb[q] = in ? __builtin_nontemporal_load(v4 + 2 * (size_t)gi + 1) : zf;
    }
    if (WEIGHTED) {
      b.xa[q] = in ? __builtin_nontemporal_load(w4 + 2 * (size_t)gi) : zf;
      b.xb[q] = in ? __builtin_nontemporal_load(w4 + 2 * (size_t)gi + 1) : zf;
    } else {
      b.xa[q] = b.xb[q] = zf;
    }
  }
}

template <bool HOTRANGE, bool WEIGHTED, bool SIGNED, int G>
__device__ __forceinline__ void sum_batch(const PbBatch<G> &b, uint32_t base, uint32_t wg1,
                                          int lane, uint32_t e0, uint32_t e1, int S, bool part,
                                          unsigned long long *acc, const float *s_hot,
                                          unsigned long long &psum) {
#pragma unroll
  for (int q = 0; q < G; q++) {
    const uint32_t gi = base + q * 64 + lane;
    if (gi < wg1)
      group_accumulate<HOTRANGE, WEIGHTED, SIGNED>(acc, b.wa[q], b.wb[q], b.va[q], b.vb[q],
                                                   b.xa[q], b.xb[q], gi * 8, e0, e1, S, part,
                                                   s_hot, psum);
  }
}

// The first group of this wave's first block of the range [e0, e1).
template <int G>
__device__ __forceinline__ uint32_t first_block(uint32_t e0) {
  return (e0 >> 3) + (threadIdx.x >> 6) * 64 * G;
}

// Sums the entries [e0, e1) of one item in 8-entry groups, G groups loaded
// per lane before any is summed; the workgroup's waves take consecutive
// 64 G-group blocks, so that together they stream one region of the range.
// This wave starts at the block of group `base`.  HOTRANGE: the entries are
// hot (packed value slots).
template <bool HOTRANGE, bool WEIGHTED, bool SIGNED, int G>
__device__ __forceinline__ void accumulate_from(const PbDev &d, uint32_t base, uint32_t e0,
                                                uint32_t e1, int S, bool part,
                                                unsigned long long *acc, const float *s_hot,
                                                unsigned long long &psum) {
  const int lane = __lane_id();
  const uint32_t g1 = (e1 + 7) >> 3;
  for (; base < g1; base += kWaves * 64 * G) {
    PbBatch<G> b;
    load_batch<HOTRANGE, WEIGHTED, G>(d, base, g1, lane, b);
    sum_batch<HOTRANGE, WEIGHTED, SIGNED, G>(b, base, g1, lane, e0, e1, S, part, acc, s_hot,
                                             psum);
  }
}

template <bool HOTRANGE, bool WEIGHTED, bool SIGNED, int G>
__device__ __forceinline__ void accumulate_range(const PbDev &d, uint32_t e0, uint32_t e1, int S,
                                                 bool part, unsigned long long *acc,
                                                 const float *s_hot,
                                                 unsigned long long &psum) {
  if (e0 >= e1) return;
  accumulate_from<HOTRANGE, WEIGHTED, SIGNED, G>(d, first_block<G>(e0), e0, e1, S, part, acc,
                                                 s_hot, psum);
}

// This thread's 16-byte pieces of hot window w's contributions (sources
// w * win ..., cnt of them): piece j holds sources 4 i .. 4 i + 3 of the
// window, i = threadIdx.x + j * kBlock, zeros past cnt.  The plan keeps nhot
// (so cnt) a multiple of 4 and the launch checks that contrib is 16-byte
// aligned.
__device__ __forceinline__ void hot_load(const PbDev &d, const float *__restrict__ contrib,
                                         uint32_t w, float4 (&hv)[kHotPer4]) {
  const uint32_t cnt = min(d.win, d.nhot - w * d.win);
  const float4 *src = reinterpret_cast<const float4 *>(contrib + (size_t)w * d.win);
#pragma unroll
  for (int j = 0; j < kHotPer4; j++) {
    const uint32_t i4 = threadIdx.x + j * kBlock;
    hv[j] = 4 * i4 < cnt ? src[i4] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Stores the pieces of hot_load that hold sources of the window.
__device__ __forceinline__ void hot_store(const PbDev &d, uint32_t w, const float4 (&hv)[kHotPer4],
                                          float4 *s_hot4) {
  const uint32_t cnt = min(d.win, d.nhot - w * d.win);
#pragma unroll
  for (int j = 0; j < kHotPer4; j++) {
    const uint32_t i4 = threadIdx.x + j * kBlock;
    if (4 * i4 < cnt) s_hot4[i4] = hv[j];
  }
}

// Copies the contributions of hot window w into the LDS copy.  The loads are
// issued before the first barrier, so that they overlap the wait for the
// waves still reading the previous window.
__device__ __forceinline__ void hot_window(const PbDev &d, const float *__restrict__ contrib,
                                           uint32_t w, float4 *s_hot4) {
  float4 hv[kHotPer4];
  hot_load(d, contrib, w, hv);
  __syncthreads();  // every wave is done with the previous window
  hot_store(d, w, hv, s_hot4);
  __syncthreads();
}

// HOT: persistent workgroups (one per CU) that first copy the hot sources'
// contributions into LDS, then take items blockIdx.x, blockIdx.x + gridDim.x,
// ...; otherwise one item per workgroup and no hot range.  With several hot
// windows an item sums each window's entries after loading that window into
// LDS; consecutive items walk the windows in alternating directions, so an
// item starts on the window its predecessor ended with.
// SPMV: rows receive the plain sum (y[r] = sum); otherwise the PageRank
// update.  WEIGHTED: terms are w * x with the bin-major weights.
template <bool HOT, bool SPMV, bool WEIGHTED>
__device__ __forceinline__ void pb_accum(const PbDev &d, int S, float *__restrict__ y,
                                         const PrArgs &a, const float *__restrict__ contrib) {
  __shared__ unsigned long long s_acc[kBinRows + 64];  // + a dummy slot per lane
  __shared__ float4 s_hot4[HOT ? kHotMax / 4 : 1];
  float *s_hot = reinterpret_cast<float *>(s_hot4);
  __shared__ unsigned long long s_ps[kWaves];
  __shared__ double s_d[kWaves], s_l[kWaves];
  __shared__ uint32_t s_next;
  const uint32_t tid = threadIdx.x;
  const int lane = __lane_id();
  const uint32_t wave = tid >> 6;
  if (HOT) {  // window 0; the first barrier of the item loop publishes it
    float4 hv[kHotPer4];
    hot_load(d, contrib, 0, hv);
    hot_store(d, 0, hv, s_hot4);
  }
  uint32_t loaded = 0u, down = 0u;  // window in s_hot; walk direction of the next item
  // One item per workgroup unless the launch is persistent: then a workgroup
  // takes item blockIdx.x first and the next unclaimed item after each one
  // (items run largest first), so that the items balance over the CUs.
  uint32_t item = blockIdx.x;
  for (bool first = true;; first = false) {
    if (!first) {
      if (!d.persistent) break;
      // every thread read s_next before the barriers of the item just done
      if (tid == 0) s_next = gridDim.x + atomicAdd(d.work, 1u);
      __syncthreads();
      item = s_next;
    }
    if (item >= d.nrun) break;  // the same for the whole workgroup
    const PbItem it = d.items[item];
    const bool part = it.nrows == 0;
    if (!part)
      for (uint32_t i = tid; i < it.nrows; i += kBlock) s_acc[i] = 0ull;
    __syncthreads();  // accumulators cleared (and, the first time, the hot copy made)
    unsigned long long psum = 0ull;
    // the hot form runs 16 waves per CU: more loads in flight per lane
    constexpr int G = HOT && !WEIGHTED ? 4 : 2;
    const uint32_t nb = it.nb, stride = d.nwin + 1;
    // start of bin j's entries of hot window w (w = nwin: of its scattered
    // entries), and the end of bin j
    auto wstart = [&](uint32_t j, uint32_t w) { return d.wseg[it.ws + j * stride + w]; };
    auto bin_end = [&](uint32_t j) { return d.wseg[it.ws + (j + 1) * stride]; };
    if (HOT && d.nwin == 1) {
      for (uint32_t j = 0; j < nb; j++)
        accumulate_range<true, WEIGHTED, SPMV, G>(d, wstart(j, 0), wstart(j, 1), S, part, s_acc,
                                                  s_hot, psum);
    } else if (HOT) {
      // groups per lane of a window's first round, loaded ahead of its switch
      constexpr int GS = 2;
      constexpr uint32_t kRound = kWaves * 64 * GS;
      for (uint32_t k = 0; k < d.nwin; k++) {
        const uint32_t w = down ? d.nwin - 1 - k : k;
        // the first bin with entries in w (the same for the whole workgroup)
        uint32_t jf = 0;
        while (jf < nb && wstart(jf, w) == wstart(jf, w + 1)) jf++;
        if (jf == nb) continue;
        uint32_t j = jf;
        if (w != loaded) {
          // The first round of the window's first range is loaded before the
          // switch, so that its latency overlaps the barriers.
          const uint32_t a0 = wstart(jf, w), a1 = wstart(jf, w + 1);
          const uint32_t hb = first_block<GS>(a0), hg1 = (a1 + 7) >> 3;
          PbBatch<GS> bt;
          load_batch<true, WEIGHTED, GS>(d, hb, hg1, lane, bt);
          hot_window(d, contrib, w, s_hot4);
          loaded = w;
          sum_batch<true, WEIGHTED, SPMV, GS>(bt, hb, hg1, lane, a0, a1, S, part, s_acc, s_hot,
                                              psum);
          // the range's entries past its first round (group-aligned start)
          const uint64_t rest = ((uint64_t)(a0 >> 3) + kRound) << 3;
          if (rest < a1)
            accumulate_range<true, WEIGHTED, SPMV, G>(d, (uint32_t)rest, a1, S, part, s_acc, s_hot,
                                                      psum);
          j++;
        }
        for (; j < nb; j++)
          accumulate_range<true, WEIGHTED, SPMV, G>(d, wstart(j, w), wstart(j, w + 1), S, part,
                                                    s_acc, s_hot, psum);
      }
      down ^= 1u;
    }
    for (uint32_t j = 0; j < nb; j++)
      accumulate_range<false, WEIGHTED, SPMV, G>(d, wstart(j, HOT ? d.nwin : 0u), bin_end(j), S,
                                                 part, s_acc, s_hot, psum);
    if (part) {
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) psum += __shfl_xor(psum, o);
      if (lane == 0) s_ps[wave] = psum;
      __syncthreads();
      if (tid == 0) {
        unsigned long long t = 0ull;
        for (int q = 0; q < kWaves; q++) t += s_ps[q];
        d.part_sum[it.row0] = t;
        d.item_part[item] = make_double2(0.0, 0.0);
      }
      __syncthreads();  // s_ps reuse
      continue;
    }
    __syncthreads();
    if (SPMV) {
      for (uint32_t i = tid; i < it.nrows; i += kBlock) y[it.row0 + i] = from_fixed(s_acc[i], S);
      if (tid == 0) d.item_part[item] = make_double2(0.0, 0.0);
      __syncthreads();  // s_acc reuse by the next item
      continue;
    }
    // per-row inputs of this thread's rows, in batches of kRowBatch rows whose
    // loads are all issued before the first use
    const float basep = pr_base(a);
    double dang = 0.0, l1 = 0.0;
    for (int j0 = 0; j0 < kRowPer; j0 += kRowBatch) {
      uint32_t dg[kRowBatch];
      float ro[kRowBatch];
#pragma unroll
      for (int j = 0; j < kRowBatch; j++) {
        const uint32_t i = tid + (j0 + j) * kBlock;
        dg[j] = i < it.nrows ? a.out_deg[it.row0 + i] : 0u;
        ro[j] = i < it.nrows ? a.rank_old[it.row0 + i] : 0.f;
      }
#pragma unroll
      for (int j = 0; j < kRowBatch; j++) {
        const uint32_t i = tid + (j0 + j) * kBlock;
        if (i < it.nrows)
          finish_row<true>(it.row0 + i, from_fixed(s_acc[i], S), y, a, basep, dg[j], ro[j], dang,
                           l1);
      }
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
      d.item_part[item] = make_double2(td, tl);
    }
    __syncthreads();  // s_acc / s_d reuse by the next item
  }
  // the last workgroup out resets the item counters for the next launch
  if (d.persistent && tid == 0) {
    __threadfence();
    if (atomicAdd(d.work + 1, 1u) == gridDim.x - 1) {
      atomicExch(d.work, 0u);
      atomicExch(d.work + 1, 0u);
    }
  }
}

template <bool SPMV, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_pb_accum_hot(PbDev d, int S, float *__restrict__ y,
                                                         PrArgs a,
                                                         const float *__restrict__ contrib) {
  pb_accum<true, SPMV, WEIGHTED>(d, S, y, a, contrib);
}

// Two workgroups per CU: held to 64 VGPRs (8 waves per SIMD).
template <bool SPMV, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_pb_accum_cold(PbDev d, int S, float *__restrict__ y, PrArgs a,
                const float *__restrict__ contrib) {
  pb_accum<false, SPMV, WEIGHTED>(d, S, y, a, contrib);
}

// Long rows (sum of their parts) and the iteration's FP64 statistics; the
// last block to finish writes them and re-arms the counter.
template <bool SPMV>
__global__ __launch_bounds__(256) void k_pb_finish(PbDev d, int S, float *__restrict__ y,
                                                   PrArgs a) {
  __shared__ double s_d[4], s_l[4];
  __shared__ uint32_t s_last;
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const uint32_t gtid = blockIdx.x * 256 + tid;
  const uint32_t gthreads = gridDim.x * 256;
  if (SPMV) {
    for (uint32_t j = gtid; j < d.nlong; j += gthreads) {
      const uint4 lr = d.long_rows[j];
      unsigned long long t = 0ull;
      for (uint32_t q = 0; q < lr.z; q++) t += d.part_sum[lr.y + q];
      y[lr.x] = from_fixed(t, S);
    }
    return;
  }
  const float basep = pr_base(a);
  double dang = 0.0, l1 = 0.0;
  for (uint32_t j = gtid; j < d.nlong; j += gthreads) {
    const uint4 lr = d.long_rows[j];
    unsigned long long t = 0ull;
    for (uint32_t q = 0; q < lr.z; q++) t += d.part_sum[lr.y + q];
    finish_row<true>(lr.x, from_fixed(t, S), y, a, basep, a.out_deg[lr.x], a.rank_old[lr.x],
                     dang, l1);
  }
  for (uint32_t i = gtid; i < d.nitems; i += gthreads) {
    const double2 p = d.item_part[i];
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
    for (int q = 0; q < 4; q++) {
      td += s_d[q];
      tl += s_l[q];
    }
    d.fin_part[blockIdx.x] = make_double2(td, tl);
    __threadfence();
    const uint32_t prev = atomicAdd(d.fin_counter, 1u);
    s_last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return;  // the same for the whole block
  // the last block: every thread sums a strided share of the block partials,
  // then the waves and the block reduce them (a fixed order)
  __threadfence();
  double td = 0.0, tl = 0.0;
  for (uint32_t b = tid; b < gridDim.x; b += 256) {
    const double *pp = reinterpret_cast<const double *>(&d.fin_part[b]);
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
    for (int q = 0; q < 4; q++) {
      sd += s_d[q];
      sl += s_l[q];
    }
    a.stats_out[0] = sd;
    a.stats_out[1] = sl;
    *d.fin_counter = 0u;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct PbPlan {
  const void *for_key = nullptr;  // entry source the plan was built for
  uint32_t nbins = 0, nchunks = 0, ntiles = 0, nitems = 0, nlong = 0, nparts = 0;
  uint32_t nlive = 0;      // items with entries or parts (the rest are last in the list)
  uint32_t tile_len = 0;   // scatter tile length the plan was cut with
  int hot_req = 0;         // hot setting the plan was built for
  uint32_t nhot = 0;       // sources below this id are hot
  uint32_t win = 0;        // sources per hot window
  uint32_t nwin = 0;       // hot windows (ceil(nhot / window))
  uint32_t cus = 0;        // compute units (persistent accumulation grid)
  uint32_t max_rowlen = 0; // most entries of one row
  float wmax = 1.f;        // largest |w| (1 without weights)
  uint64_t nsegs = 0;      // non-empty (chunk, bin) segments
  PbTile *tiles = nullptr;
  uint2 *chunk_src = nullptr;
  uint2 *tab = nullptr;
  uint16_t *sloc = nullptr, *dloc = nullptr;
  float *val = nullptr;
  float *w = nullptr;      // bin-major weights, or null
  PbItem *items = nullptr;
  uint32_t *wseg = nullptr;  // per item: starts of its hot windows' entries
  uint4 *long_rows = nullptr;
  unsigned long long *part_sum = nullptr;
  double2 *item_part = nullptr, *fin_part = nullptr;
  uint32_t *fin_counter = nullptr;
  uint32_t *work = nullptr;  // persistent phase-2 item counters
};

void eg_pb_free(eg_graph *g) {
  PbPlan *p = g->pb;
  if (!p) return;
  eg_dfree_t(&p->tiles);
  eg_dfree_t(&p->chunk_src);
  eg_dfree_t(&p->tab);
  eg_dfree_t(&p->sloc);
  eg_dfree_t(&p->dloc);
  eg_dfree_t(&p->val);
  eg_dfree_t(&p->w);
  eg_dfree_t(&p->items);
  eg_dfree_t(&p->wseg);
  eg_dfree_t(&p->long_rows);
  eg_dfree_t(&p->part_sum);
  eg_dfree_t(&p->item_part);
  eg_dfree_t(&p->fin_part);
  eg_dfree_t(&p->fin_counter);
  eg_dfree_t(&p->work);
  delete p;
  g->pb = nullptr;
}

namespace {

template <typename T>
int upload(eg_graph *g, T **dst, const std::vector<T> &src, const char *what) {
  EG_TRY(eg_dalloc_t(dst, src.empty() ? 1 : src.size(), what));
  if (!src.empty())
    EG_HIP(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                          g->stream));
  return EG_OK;
}

// Sorts the entries into one order and writes their u16 offsets (for the
// bin-major order also the packed hot value slots and the weights); seg
// receives the start of every key's segment (nkeys + 1 entries) on the host.
int pb_order(eg_graph *g, const PbSrcDev &src, const uint32_t *bin_of, const uint32_t *chunk_of,
             const uint32_t *bin_row0, const uint32_t *chunk_src0, uint32_t nbins,
             uint32_t nchunks, uint32_t nhot, uint32_t win, uint32_t nwin, int bin_major,
             uint16_t *out, uint32_t *packed, float *wout, std::vector<uint32_t> &seg) {
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  hipStream_t s = g->stream;
  const uint64_t nkeys =
      bin_major ? (uint64_t)nbins * (nchunks + nwin) : (uint64_t)nchunks * nbins + 1;
  const bool wt = bin_major && wout;
  uint32_t *key = nullptr, *pay = nullptr, *pay2 = nullptr, *pay3 = nullptr, *skey = nullptr,
           *spay = nullptr, *spay2 = nullptr, *segd = nullptr;
  int rc = EG_OK;
  do {
    if ((rc = eg_dalloc_t(&key, m, "binned keys")) || (rc = eg_dalloc_t(&pay, m, "binned payload")) ||
        (rc = eg_dalloc_t(&skey, m, "binned sorted keys")) ||
        (rc = eg_dalloc_t(&spay, m, "binned sorted payload")) ||
        (rc = eg_dalloc_t(&segd, nkeys + 1, "binned segments")))
      break;
    if (bin_major && ((rc = eg_dalloc_t(&pay2, m, "binned sources")) ||
                      (rc = eg_dalloc_t(&spay2, m, "binned sorted sources"))))
      break;
    if (wt && (rc = eg_dalloc_t(&pay3, m, "binned weights")))
      break;
    const uint32_t grid =
        src.row ? std::min<uint64_t>(blocks_for(m, 256), 65536)
                : std::min<uint64_t>(blocks_for((uint64_t)n * 64, 256), 65536);
    hipLaunchKernelGGL(k_pb_keys, dim3(grid), dim3(256), 0, s, n, m, src, bin_of, chunk_of,
                       bin_row0, chunk_src0, nbins, nchunks, nhot, win, nwin, bin_major,
                       key, pay, pay2, wt ? pay3 : nullptr);
    if (hipGetLastError() != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "binned key launch");
      break;
    }
    void *tmp = nullptr;
    EgSortCols c;
    c.np = bin_major ? (wt ? 3 : 2) : 1;
    c.in[0] = pay;
    c.out[0] = spay;
    c.in[1] = pay2;
    c.out[1] = spay2;
    c.in[2] = pay3;
    c.out[2] = reinterpret_cast<uint32_t *>(wout);
    if ((rc = eg_scratch(g, eg_radix_tmp_bytes(m, c.np), &tmp))) break;
    if ((rc = eg_radix_sort(s, m, eg_bits_for(nkeys), key, skey, c, tmp))) break;
    hipLaunchKernelGGL(k_pb_narrow, dim3(blocks_for(m, 256)), dim3(256), 0, s, m, spay, out, skey,
                       spay2, nchunks + nwin, nwin, win,
                       bin_major && nhot ? packed : nullptr);
    if (hipGetLastError() != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "binned narrow launch");
      break;
    }
    if ((rc = eg_ptr_from_sorted(s, skey, m, (uint32_t)nkeys, segd, nullptr))) break;
    seg.resize(nkeys + 1);
    if (hipMemcpyAsync(seg.data(), segd, (nkeys + 1) * 4, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "binned segment download");
      break;
    }
  } while (0);
  (void)hipStreamSynchronize(s);
  eg_dfree_t(&key);
  eg_dfree_t(&pay);
  eg_dfree_t(&pay2);
  eg_dfree_t(&pay3);
  eg_dfree_t(&skey);
  eg_dfree_t(&spay);
  eg_dfree_t(&spay2);
  eg_dfree_t(&segd);
  return rc;
}

// Per-vertex counts of a u32 array of m ids on the host.
int pb_counts(eg_graph *g, const uint32_t *ids, std::vector<uint32_t> &cnt) {
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  cnt.assign(n, 0u);
  uint32_t *d = nullptr;
  EG_TRY(eg_dalloc_t(&d, n, "binned counts"));
  int rc = EG_OK;
  if (hipMemsetAsync(d, 0, (size_t)n * 4, s) != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "binned count clear");
  if (!rc) rc = eg_histogram(s, ids, g->m, d, n);
  if (!rc && (hipMemcpyAsync(cnt.data(), d, (size_t)n * 4, hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = eg_fail(EG_EDEVICE, "binned count download");
  eg_dfree_t(&d);
  return rc;
}

}  // namespace

// Builds (on first use for this entry source) the binned plan of src: the
// entries of row v are the sources u of its in-edges.
int eg_pb_build(eg_graph *g, const PbSource &src) {
  const uint32_t want_tile = g->tune.pr_tile > 0 ? (uint32_t)g->tune.pr_tile : kTileLen;
  const int want_hot = g->tune.pr_hot;
  const void *key = src.row ? (const void *)src.row : (const void *)src.ptr;
  if (g->pb && g->pb->for_key == key && g->pb->tile_len == want_tile && g->pb->hot_req == want_hot)
    return EG_OK;
  eg_pb_free(g);
  const double t0 = eg_wall_ms();
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  hipStream_t s = g->stream;
  if (m == 0 || n == 0) return eg_fail(EG_ENOTSUP, "binned PageRank needs entries");
  int cus = 0;
  EG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
  PbSrcDev sd;
  sd.ptr = src.ptr;
  sd.row = src.row;
  sd.idx = src.idx;
  sd.w = src.w;
  // 1. row lengths and source counts on the host
  std::vector<uint32_t> hptr((size_t)n + 1), scnt;
  if (src.row) {
    std::vector<uint32_t> rcnt;
    EG_TRY(pb_counts(g, src.row, rcnt));
    hptr[0] = 0;
    for (uint32_t v = 0; v < n; v++) hptr[v + 1] = hptr[v] + rcnt[v];
  } else {
    EG_HIP(hipMemcpyAsync(hptr.data(), src.ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
  }
  EG_TRY(pb_counts(g, src.idx, scnt));
  // 2. hot sources, bins of rows and chunks of the other sources, each
  // balanced by entries.  Hot sources form windows of kHotMax, one in LDS at a
  // time.  A hot entry saves about 8 bytes of streaming over a scattered one,
  // and every window after the first costs each work item a switch (about a
  // microsecond); items are about m / 2048 entries, so a window pays from
  // about 2% of the entries on: trailing windows below that are dropped.
  // The hot copy as a whole pays only where the first sources take a fair
  // share of the entries (degree-ordered or R-MAT ids), so it is dropped
  // below 15%.  With EG_PB_GEOMETRY=small in the environment every cap is
  // cut down so that small graphs exercise hot and cold entries in one bin,
  // several hot windows, long-row parts and tile cuts (tests).
  uint32_t rows_cap = kBinRows, src_cap = kChunkSrc, tile_len = want_tile, tile_bins = kTileBins;
  uint32_t hot_cap = want_hot > 0 ? (uint32_t)want_hot : 0u;
  uint32_t win = kHotMax;  // sources per hot window
  uint64_t target = std::max<uint64_t>(m / 2048, 4096);
  double min_hot_share = 0.15, min_win_share = 0.02;
  {
    const char *e = getenv("EG_PB_GEOMETRY");
    if (e && !strcmp(e, "small")) {
      rows_cap = 64;
      src_cap = 96;
      tile_len = 512;
      tile_bins = 16;
      target = 256;
      hot_cap = std::min<uint32_t>(hot_cap, 40);
      win = 16;
      min_hot_share = 0.0;
      min_win_share = 0.0;
    }
  }
  // a multiple of 4: phase 2 copies hot windows in 16-byte pieces
  uint32_t nhot = std::min<uint32_t>(hot_cap, n) & ~3u;
  {
    uint64_t hot_entries = 0, win_entries = 0;
    uint32_t kept = 0;  // sources of the windows kept
    for (uint32_t u = 0; u < nhot; u++) {
      win_entries += scnt[u];
      if (u + 1 == nhot || (u + 1) % win == 0) {  // a window ends
        if (u >= win && (double)win_entries < min_win_share * (double)m) break;
        hot_entries += win_entries;
        win_entries = 0;
        kept = u + 1;
      }
    }
    nhot = kept;
    if ((double)hot_entries < min_hot_share * (double)m) nhot = 0;
  }
  const uint32_t nwin = (nhot + win - 1) / win;
  if (nwin > kHotWins) return eg_fail(EG_EINVAL, "binned plan: %u hot windows", nwin);
  uint32_t max_rowlen = 0;
  std::vector<uint32_t> bin_row0, bin_of(n), chunk_src0, chunk_of(n, 0u);
  {
    uint32_t rows = 0;
    uint64_t ents = 0;
    for (uint32_t v = 0; v < n; v++) {
      const uint32_t dv = hptr[v + 1] - hptr[v];
      max_rowlen = std::max(max_rowlen, dv);
      if (rows == 0 || rows + 1 > rows_cap || ents + dv > target) {
        bin_row0.push_back(v);
        rows = 0;
        ents = 0;
      }
      bin_of[v] = (uint32_t)bin_row0.size() - 1;
      rows++;
      ents += dv;
    }
    bin_row0.push_back(n);
    uint32_t srcs = 0;
    ents = 0;
    for (uint32_t u = nhot; u < n; u++) {
      if (srcs == 0 || srcs + 1 > src_cap || ents + scnt[u] > target) {
        chunk_src0.push_back(u);
        srcs = 0;
        ents = 0;
      }
      chunk_of[u] = (uint32_t)chunk_src0.size() - 1;
      srcs++;
      ents += scnt[u];
    }
    chunk_src0.push_back(n);
  }
  const uint32_t nbins = (uint32_t)bin_row0.size() - 1;
  const uint32_t nchunks = (uint32_t)chunk_src0.size() - 1;
  // Phase-2 work items: runs of consecutive bins with at most rows_cap rows
  // and kItemBins targets of entries together (each switch of hot window then
  // serves them all, while phase 1 keeps the bins' shorter write segments); a
  // bin holding one long row stays alone and is cut into parts.  Row offsets
  // are stored relative to the first row of the bin's item.
  std::vector<uint32_t> grp_bin0, row_base(nbins);
  {
    const uint64_t item_ents = kItemBins * target;
    uint32_t rows = 0;
    uint64_t ents = 0;
    bool alone = false;  // the open item is a long row
    for (uint32_t b = 0; b < nbins; b++) {
      const uint32_t nr = bin_row0[b + 1] - bin_row0[b];
      const uint64_t eb = (uint64_t)hptr[bin_row0[b + 1]] - hptr[bin_row0[b]];
      const bool long_row = nr == 1 && eb > target;
      if (b == 0 || alone || long_row || rows + nr > rows_cap || ents + eb > item_ents) {
        grp_bin0.push_back(b);
        rows = 0;
        ents = 0;
      }
      alone = long_row;
      row_base[b] = bin_row0[grp_bin0.back()];
      rows += nr;
      ents += eb;
    }
    grp_bin0.push_back(nbins);
  }
  if ((uint64_t)nbins * (nchunks + nwin) >= (1ull << 31))
    return eg_fail(EG_ENOTSUP, "binned plan: %u bins x %u chunks", nbins, nchunks);

  PbPlan *p = new PbPlan();
  g->pb = p;
  p->nbins = nbins;
  p->nchunks = nchunks;
  p->nhot = nhot;
  p->win = win;
  p->nwin = nwin;
  p->cus = (uint32_t)cus;
  p->max_rowlen = max_rowlen;
  std::vector<uint32_t> seg_bm, seg_cm;
  {
    uint32_t *d_bin_of = nullptr, *d_chunk_of = nullptr, *d_row0 = nullptr, *d_src0 = nullptr;
    uint32_t *d_wmax = nullptr;
    int rc = EG_OK;
    do {
      if ((rc = upload(g, &d_bin_of, bin_of, "binned bin_of")) ||
          (rc = upload(g, &d_chunk_of, chunk_of, "binned chunk_of")) ||
          (rc = upload(g, &d_row0, row_base, "binned row bases")) ||
          (rc = upload(g, &d_src0, chunk_src0, "binned chunk sources")))
        break;
      if ((rc = eg_dalloc_t(&p->dloc, m + 16, "binned row offsets")) ||
          (rc = eg_dalloc_t(&p->sloc, m + 16, "binned source offsets")) ||
          (rc = eg_dalloc_t(&p->val, m + 16, "binned values")))
        break;
      if (src.w) {
        if ((rc = eg_dalloc_t(&p->w, m + 16, "binned weights")) ||
            (rc = eg_dalloc_t(&d_wmax, 1, "binned weight max")))
          break;
        uint32_t wbits = 0;
        if (hipMemsetAsync(d_wmax, 0, 4, s) != hipSuccess) {
          rc = eg_fail(EG_EDEVICE, "binned weight max clear");
          break;
        }
        hipLaunchKernelGGL(k_pb_absmax, dim3(std::min<uint64_t>(blocks_for(m, 256), 4096)),
                           dim3(256), 0, s, m, src.w, d_wmax);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&wbits, d_wmax, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
          rc = eg_fail(EG_EDEVICE, "binned weight max");
          break;
        }
        memcpy(&p->wmax, &wbits, 4);
      }
      if ((rc = pb_order(g, sd, d_bin_of, d_chunk_of, d_row0, d_src0, nbins, nchunks, nhot,
                         win, nwin, 1, p->dloc, reinterpret_cast<uint32_t *>(p->val), p->w,
                         seg_bm)))
        break;
      rc = pb_order(g, sd, d_bin_of, d_chunk_of, d_row0, d_src0, nbins, nchunks, nhot, win,
                    nwin, 0, p->sloc, nullptr, nullptr, seg_cm);
    } while (0);
    (void)hipStreamSynchronize(s);
    eg_dfree_t(&d_bin_of);
    eg_dfree_t(&d_chunk_of);
    eg_dfree_t(&d_row0);
    eg_dfree_t(&d_src0);
    eg_dfree_t(&d_wmax);
    if (rc) {
      eg_pb_free(g);
      return rc;
    }
  }
  // 3. segment table, scatter tiles, accumulation items.  Bin-major key of
  // (bin b, chunk c) is b * (nchunks + nwin) + nwin + c (the bin's hot
  // entries of window w at + w come first); chunk-major key is c * nbins + b.
  const uint32_t kb = nchunks + nwin;
  std::vector<uint2> tab((size_t)nchunks * (nbins + 1));
  std::vector<uint2> csrc(nchunks);
  std::vector<PbTile> tiles;
  for (uint32_t c = 0; c < nchunks; c++) {
    uint2 *row = &tab[(size_t)c * (nbins + 1)];
    for (uint32_t b = 0; b < nbins; b++)
      row[b] = make_uint2(seg_cm[(size_t)c * nbins + b], seg_bm[(size_t)b * kb + nwin + c]);
    const uint32_t cend = seg_cm[(size_t)(c + 1) * nbins];
    row[nbins] = make_uint2(cend, 0u);
    csrc[c] = make_uint2(chunk_src0[c], chunk_src0[c + 1] - chunk_src0[c]);
    // segment containing position q: last b with row[b].x <= q
    auto seg_of = [&](uint32_t q) {
      uint32_t lo = 0, hi = nbins - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (row[mid].x <= q) lo = mid; else hi = mid - 1;
      }
      return lo;
    };
    for (uint32_t q = row[0].x; q < cend;) {
      PbTile t = {};
      t.c = c;
      t.p0 = q;
      t.b0 = seg_of(q);
      uint32_t q1 = std::min<uint64_t>(cend, (uint64_t)q + tile_len);
      uint32_t b1 = seg_of(q1 - 1);
      if (b1 - t.b0 + 1 > tile_bins) {
        b1 = t.b0 + tile_bins - 1;
        q1 = row[b1 + 1].x;
      }
      t.p1 = q1;
      t.nb = b1 - t.b0 + 1;
      tiles.push_back(t);
      q = q1;
    }
  }
  std::vector<PbItem> items;
  std::vector<uint32_t> wseg;  // per item: nb (nwin + 1) window starts and the end
  std::vector<uint4> long_rows;
  uint32_t nparts = 0;
  // bin b's window starts clamped to [a0, a1), at the end of wseg
  auto push_windows = [&](uint32_t b, uint32_t a0, uint32_t a1) {
    for (uint32_t w = 0; w <= nwin; w++)
      wseg.push_back(std::min(std::max(seg_bm[(size_t)b * kb + w], a0), a1));
  };
  for (size_t gi = 0; gi + 1 < grp_bin0.size(); gi++) {
    const uint32_t b0 = grp_bin0[gi], b1 = grp_bin0[gi + 1];
    const uint32_t e0 = seg_bm[(size_t)b0 * kb], e1 = seg_bm[(size_t)b1 * kb];
    const uint32_t r0 = bin_row0[b0], nr = bin_row0[b1] - r0;
    if (b1 == b0 + 1 && nr == 1 && e1 - e0 > target) {
      const uint32_t np = (uint32_t)((e1 - e0 + target - 1) / target);
      long_rows.push_back(make_uint4(r0, nparts, np, 0u));
      for (uint32_t q = 0; q < np; q++) {
        const uint32_t a0 = e0 + (uint32_t)(q * target);
        const uint32_t a1 = std::min<uint64_t>(e1, (uint64_t)a0 + target);
        PbItem it = {};
        it.e0 = a0;
        it.e1 = a1;
        it.row0 = nparts + q;
        it.nrows = 0;
        it.nb = 1;
        it.ws = (uint32_t)wseg.size();
        push_windows(b0, a0, a1);
        wseg.push_back(a1);
        items.push_back(it);
      }
      nparts += np;
    } else {
      PbItem it = {};
      it.e0 = e0;
      it.e1 = e1;
      it.row0 = r0;
      it.nrows = nr;
      it.nb = b1 - b0;
      it.ws = (uint32_t)wseg.size();
      for (uint32_t b = b0; b < b1; b++) push_windows(b, e0, e1);
      wseg.push_back(e1);
      items.push_back(it);
    }
  }
  // Accumulation items, largest first: workgroups are dispatched in launch
  // order, so the longest items start first and the short ones fill the tail.
  // Items of rows without entries come last, so that phase 2 can leave them
  // out (eg_pb_iterate's skip_static).
  std::stable_sort(items.begin(), items.end(), [](const PbItem &x, const PbItem &y) {
    const bool xs = x.e0 == x.e1, ys = y.e0 == y.e1;
    if (xs != ys) return ys;
    return x.e1 - x.e0 + 4 * x.nrows > y.e1 - y.e0 + 4 * y.nrows;
  });
  uint32_t nlive = 0;
  while (nlive < items.size() && items[nlive].e0 != items[nlive].e1) nlive++;
  // Workgroups are dealt to the 8 XCDs round-robin by launch index.  Launch
  // index j runs tile (j % 8) * ceil(T / 8) + j / 8, so every XCD walks a
  // contiguous eighth of the chunk-ordered tile list and the tiles of one
  // chunk share that XCD's L2 for the chunk's window and table.
  {
    const size_t T = tiles.size(), per = (T + 7) / 8;
    std::vector<PbTile> order;
    order.reserve(T);
    for (size_t j = 0; j < 8 * per; j++) {
      const size_t src = (j % 8) * per + j / 8;
      if (src < T) order.push_back(tiles[src]);
    }
    tiles.swap(order);
  }
  for (size_t k = 0; k + 1 < seg_cm.size(); k++) p->nsegs += seg_cm[k + 1] > seg_cm[k];
  p->tile_len = tile_len;
  p->hot_req = want_hot;
  p->ntiles = (uint32_t)tiles.size();
  p->nitems = (uint32_t)items.size();
  p->nlive = nlive;
  p->nlong = (uint32_t)long_rows.size();
  p->nparts = nparts;
  int rc = EG_OK;
  do {
    if ((rc = upload(g, &p->tab, tab, "binned table")) ||
        (rc = upload(g, &p->chunk_src, csrc, "binned chunks")) ||
        (rc = upload(g, &p->tiles, tiles, "binned tiles")) ||
        (rc = upload(g, &p->items, items, "binned items")) ||
        (rc = upload(g, &p->wseg, wseg, "binned hot windows")) ||
        (rc = upload(g, &p->long_rows, long_rows, "binned long rows")))
      break;
    if ((rc = eg_dalloc_t(&p->part_sum, nparts ? nparts : 1, "binned part sums")) ||
        (rc = eg_dalloc_t(&p->item_part, p->nitems, "binned item partials")) ||
        (rc = eg_dalloc_t(&p->fin_part, kFinBlocks, "binned finish partials")) ||
        (rc = eg_dalloc_t(&p->fin_counter, 1, "binned finish counter")) ||
        (rc = eg_dalloc_t(&p->work, 2, "binned item counters")))
      break;
    if (hipMemsetAsync(p->fin_counter, 0, 4, s) != hipSuccess ||
        hipMemsetAsync(p->work, 0, 8, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = eg_fail(EG_EDEVICE, "binned plan upload");
  } while (0);
  if (rc) {
    eg_pb_free(g);
    return rc;
  }
  p->for_key = key;
  g->prep_plan_ms = eg_wall_ms() - t0;
  // offsets in both orders, value slots (+ bin-major weights), segment table,
  // tiles, items, long rows and partials
  g->prep_plan_bytes = 8ull * (m + 16) + (src.w ? 4ull * (m + 16) : 0ull) +
                       8ull * p->nchunks * ((uint64_t)p->nbins + 1) + sizeof(PbTile) * p->ntiles +
                       (sizeof(PbItem) + 16) * p->nitems + 4ull * wseg.size() +
                       16ull * p->nlong + 8ull * p->nparts;
  return EG_OK;
}

bool eg_pb_selected(const eg_graph *g, int iterations, const void *key) {
  if (g->tune.pr_binned >= 0) return g->tune.pr_binned > 0;
  if (g->m < (1ull << 22)) return false;
  return (g->pb && g->pb->for_key == key) || iterations >= kAutoBinnedIters;
}

// Fixed-point scale for row sums bounded by `bound` (finite; callers check):
// the largest S with bound * 2^S <= 2^62.
int eg_pb_scale(double bound) {
  if (!(bound > 0.0) || !std::isfinite(bound)) bound = 1.0;
  return 62 - (int)ceil(log2(bound));
}

// The binned SpMV sums exactly only when every FP32 product w * x is finite
// and the row-sum bound has a scale: a weight of inf or NaN (a 12-byte record
// may hold one), or |w| |x| at or past FLT_MAX, leaves the sum to the pull
// kernels, which propagate inf and NaN as IEEE arithmetic does.
bool eg_pb_spmv_exact(const eg_graph *g, double xmax) {
  const PbPlan *p = g->pb;
  if (!p || !std::isfinite(p->wmax) || !std::isfinite(xmax)) return false;
  const double prod = (double)p->wmax * xmax;
  if (!(prod < (double)FLT_MAX)) return false;
  return std::isfinite(prod * (double)p->max_rowlen * 1.001);
}

// Launches one pass of the binned kernels from x into y.  PageRank (spmv
// false): x is the input iterate's contributions and a carries the update
// and the statistics of eg_pull.hip.  SpMV: y[v] = sum of w * x[u] (w = 1
// without weights); a is not used.
static int pb_launch(eg_graph *g, const float *x, float *y, const PrArgs &a, int S,
                     hipStream_t stream, bool spmv, bool skip_static) {
  const PbPlan *p = g->pb;
  if (!p) return eg_fail(EG_ESTATE, "binned plan not built");
  hipStream_t s = stream ? stream : g->stream;
  PbDev d;
  d.tiles = p->tiles;
  d.chunk_src = p->chunk_src;
  d.tab = p->tab;
  d.sloc = p->sloc;
  d.dloc = p->dloc;
  d.val = p->val;
  d.items = p->items;
  d.long_rows = p->long_rows;
  d.part_sum = p->part_sum;
  d.item_part = p->item_part;
  d.fin_part = p->fin_part;
  d.fin_counter = p->fin_counter;
  d.work = p->work;
  d.nbins = p->nbins;
  d.nitems = p->nitems;
  d.nrun = skip_static ? p->nlive : p->nitems;
  d.nlong = p->nlong;
  d.nhot = p->nhot;
  d.win = p->win;
  d.nwin = p->nwin;
  d.wseg = p->wseg;
  d.xcd_tiles = g->tune.pr_xcd ? 1u : 0u;
  d.w = spmv ? p->w : nullptr;
  const bool wt = d.w != nullptr;
  if (p->nhot && (reinterpret_cast<uintptr_t>(x) & 15u))
    return eg_fail(EG_EINVAL, "binned pass: contributions not 16-byte aligned");
  {
    EgProf prof(g, spmv ? "spmv_scatter" : "pr_scatter", s);
    const int sb = g->tune.pr_scatter;
    if (p->ntiles == 0) {
      // every source is hot: nothing to scatter
    } else if (sb == 8)
      hipLaunchKernelGGL((k_pb_scatter<8>), dim3(p->ntiles), dim3(kBlock), 0, s, d, x);
    else if (sb == 32)
      hipLaunchKernelGGL((k_pb_scatter<32>), dim3(p->ntiles), dim3(kBlock), 0, s, d, x);
    else
      hipLaunchKernelGGL((k_pb_scatter<16>), dim3(p->ntiles), dim3(kBlock), 0, s, d, x);
    EG_LAUNCHED();
  }
  {
    EgProf prof(g, spmv ? "spmv_accum" : "pr_accum", s);
    d.persistent = p->nhot && g->tune.pr_persist ? 1u : 0u;
    const dim3 grid(d.persistent ? std::min(d.nrun, p->cus) : d.nrun);
    if (d.nrun == 0) {
      // every item keeps its statistics
    } else if (p->nhot) {
      if (!spmv)
        hipLaunchKernelGGL((k_pb_accum_hot<false, false>), grid, dim3(kBlock), 0, s, d, S, y, a, x);
      else if (wt)
        hipLaunchKernelGGL((k_pb_accum_hot<true, true>), grid, dim3(kBlock), 0, s, d, S, y, a, x);
      else
        hipLaunchKernelGGL((k_pb_accum_hot<true, false>), grid, dim3(kBlock), 0, s, d, S, y, a, x);
    } else {
      if (!spmv)
        hipLaunchKernelGGL((k_pb_accum_cold<false, false>), grid, dim3(kBlock), 0, s, d, S, y, a, x);
      else if (wt)
        hipLaunchKernelGGL((k_pb_accum_cold<true, true>), grid, dim3(kBlock), 0, s, d, S, y, a, x);
      else
        hipLaunchKernelGGL((k_pb_accum_cold<true, false>), grid, dim3(kBlock), 0, s, d, S, y, a, x);
    }
    EG_LAUNCHED();
  }
  if (!spmv || p->nlong) {
    EgProf prof(g, spmv ? "spmv_finish" : "pr_finish", s);
    if (spmv)
      hipLaunchKernelGGL((k_pb_finish<true>), dim3(kFinBlocks), dim3(256), 0, s, d, S, y, a);
    else
      hipLaunchKernelGGL((k_pb_finish<false>), dim3(kFinBlocks), dim3(256), 0, s, d, S, y, a);
    EG_LAUNCHED();
  }
  return EG_OK;
}

int eg_pb_iterate(eg_graph *g, const float *contrib, float *y, const PrArgs &a, int S,
                  hipStream_t stream, bool skip_static) {
  return pb_launch(g, contrib, y, a, S, stream, false, skip_static && !a.base_vec);
}

// SpMV pass y = A^T x over the built plan, x and y in device memory; xmax =
// max |x[u]| bounds the terms: a row sum is at most wmax * longest row * xmax.
int eg_pb_spmv(eg_graph *g, const float *x, float *y, double xmax, hipStream_t stream) {
  const PbPlan *p = g->pb;
  if (!p) return eg_fail(EG_ESTATE, "binned plan not built");
  if (!eg_pb_spmv_exact(g, xmax))
    return eg_fail(EG_EINVAL, "binned SpMV needs finite weights and products");
  const double bound = (double)p->wmax * (double)p->max_rowlen * xmax * 1.001;
  PrArgs a = {};
  return pb_launch(g, x, y, a, eg_pb_scale(bound), stream, true, false);
}

extern "C" int eg_pr_binned_geometry(eg_graph *g, uint32_t *out) {
  if (!g || !out) return eg_fail(EG_EINVAL, "null argument");
  const PbPlan *p = g->pb;
  if (!p) return eg_fail(EG_ESTATE, "binned plan not built");
  out[0] = p->nbins;
  out[1] = p->nchunks;
  out[2] = p->ntiles;
  out[3] = p->nitems;
  out[4] = p->nlong;
  out[5] = (uint32_t)std::min<uint64_t>(p->nsegs, 0xFFFFFFFFull);
  out[6] = p->nhot;
  return EG_OK;
}
