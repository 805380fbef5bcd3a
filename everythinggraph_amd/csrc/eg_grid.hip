// eg_grid.hip -- PageRank and SpMV over the 2-D grid (load modes 0/1/6) and
// over the edge array (modes 7/8/9).
//
// Reference: prgrid_algo (prgrid_cilk.c:62-76) streams the cells of one grid
// row after another, cells of a row in parallel, each cell owning a disjoint
// dst range so `rank[dst] += prev[src]/deg[src]` needs no atomics; the apply
// step (:131-135) forms r = (1-d)/n + d*acc.  pr_algo_edgecentric
// (pagerank_simple.c:40-61) and prgrid_algo_col (prgrid_cilk.c:79-88) scatter
// over the edge array with a CAS add.
//
// Grid on the GPU: the parity layout keeps the reference's in-cell order
// (src asc, input order).  For computation a second copy is ordered by
// (cell, dst, src) and indexed by dst tiles of kTile vertices, so one
// workgroup owns a tile of a grid column and walks the cells of that column
// row by row, exactly the reference's row-sequential / column-parallel
// schedule with the column cut into tiles.  The grid rows (source
// partitions) can be split into bands launched one after another: every
// workgroup of a launch then gathers from the same band of x, which stays in
// L2, and the compensated accumulators (s, c) of each dst vertex are carried
// between launches exactly, so the sums are the same for every band width.  Each workgroup stages one
// cell-tile (its entries' dst offsets and values) in LDS per step and every
// lane sums the entries of the vertices it owns with a compensated FP32
// accumulator, in src order (deterministic, and the same order as the CSC
// pull).
//
// Edge array: one thread per entry with an FP32 atomic add into y (the
// reference's CAS add); the sum order is not deterministic.
#include <math.h>

#include <algorithm>
#include <cmath>

#include "eg_internal.h"
#include "eg_pr.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kTile = 1024;            // dst vertices per workgroup
constexpr uint32_t kVPT = kTile / kBlock;   // vertices per lane
constexpr uint32_t kStage = 4096;           // entries staged per LDS round

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

__host__ __device__ inline void part_range(uint32_t n, uint32_t P, uint32_t p, uint32_t &b,
                                           uint32_t &e) {
  // get_partition_range, init_all.c:29-41.
  const uint64_t split = n % P;
  const uint64_t psize = (uint64_t)n / P + 1;
  if (p < split) {
    b = (uint32_t)(p * psize);
    e = (uint32_t)((p + 1) * psize);
  } else {
    const uint64_t sp = split * psize;
    b = (uint32_t)(sp + (p - split) * (psize - 1));
    e = (uint32_t)(sp + (p - split + 1) * (psize - 1));
  }
}

// tile_off[c * (ntiles + 1) + t] = first entry of cell c whose dst offset is
// >= t * kTile.
__global__ void k_grid_tile_offsets(const uint32_t *__restrict__ cell_ptr,
                                    const uint32_t *__restrict__ cdst, uint32_t ncells, uint32_t n,
                                    uint32_t P, uint32_t ntiles, uint32_t *__restrict__ tile_off) {
  const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (uint64_t)ncells * (ntiles + 1)) return;
  const uint32_t c = (uint32_t)(id / (ntiles + 1));
  const uint32_t t = (uint32_t)(id % (ntiles + 1));
  uint32_t b, e;
  part_range(n, P, c % P, b, e);
  const uint64_t target = (uint64_t)b + (uint64_t)t * kTile;
  uint32_t lo = cell_ptr[c], hi = cell_ptr[c + 1];
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)cdst[mid] >= target)
      hi = mid;
    else
      lo = mid + 1;
  }
  tile_off[id] = lo;
}

struct GridArgs {
  const uint32_t *tile_off;
  const uint32_t *csrc;
  const uint32_t *cdst;
  const float *cw;
  uint32_t n, P, ntiles;
  // PageRank
  const float *rank_old;
  float *contrib_new;
  const uint32_t *out_deg;
  const double *dmass_in;
  double2 *part;  // per workgroup (dangling, l1)
  float base, damping;
  double damping_over_n;
  int uniform;
};

// Cells (i, j) for i in [i0, i1): carried accumulators come from / go to
// carry (per dst vertex) unless the band is the first / last.
template <bool PR, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_grid_pull(const float *__restrict__ x,
                                                      float *__restrict__ y, GridArgs a,
                                                      uint32_t i0, uint32_t i1,
                                                      float2 *__restrict__ carry) {
  __shared__ uint32_t s_dst[kStage];
  __shared__ float s_val[kStage];
  __shared__ double s_d[kWaves], s_l[kWaves];
  const uint32_t t = blockIdx.x, j = blockIdx.y;
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  uint32_t pb, pe;
  part_range(a.n, a.P, j, pb, pe);
  const uint32_t vb = pb + t * kTile;
  double dang = 0.0, l1 = 0.0;
  if (vb < pe) {
    const uint32_t ve = vb + kTile < pe ? vb + kTile : pe;
    EgKahan acc[kVPT];
    if (i0 > 0) {
#pragma unroll
      for (uint32_t r = 0; r < kVPT; r++) {
        const uint32_t v = vb + r * kBlock + tid;
        if (v < ve) {
          const float2 c = carry[v];
          acc[r].s = c.x;
          acc[r].c = c.y;
        }
      }
    }
    for (uint32_t i = i0; i < i1; i++) {
      const uint64_t c = (uint64_t)i * a.P + j;
      const uint32_t k0 = a.tile_off[c * (a.ntiles + 1) + t];
      const uint32_t k1 = a.tile_off[c * (a.ntiles + 1) + t + 1];
      for (uint32_t base = k0; base < k1; base += kStage) {
        const uint32_t cnt = k1 - base < kStage ? k1 - base : kStage;
        __syncthreads();
        for (uint32_t q = tid; q < cnt; q += kBlock) {
          const uint32_t k = base + q;
          s_dst[q] = __builtin_nontemporal_load(a.cdst + k) - vb;
          float v = x[__builtin_nontemporal_load(a.csrc + k)];
          if (WEIGHTED) v *= __builtin_nontemporal_load(a.cw + k);
          s_val[q] = v;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < kVPT; r++) {
          const uint32_t lv = r * kBlock + tid;
          // entries of local vertex lv: [lower_bound(lv), lower_bound(lv+1))
          uint32_t lo = 0, hi = cnt;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_dst[mid] < lv) lo = mid + 1; else hi = mid;
          }
          for (uint32_t q = lo; q < cnt && s_dst[q] == lv; q++) acc[r].add(s_val[q]);
        }
      }
    }
    if (i1 < a.P) {
#pragma unroll
      for (uint32_t r = 0; r < kVPT; r++) {
        const uint32_t v = vb + r * kBlock + tid;
        if (v < ve) carry[v] = make_float2(acc[r].s, acc[r].c);
      }
      return;  // not the last band: no results yet
    }
    float basep = 0.f;
    if (PR) {
      double b = (double)a.base;
      if (a.uniform) b += a.damping_over_n * (*a.dmass_in);
      basep = (float)b;
    }
#pragma unroll
    for (uint32_t r = 0; r < kVPT; r++) {
      const uint32_t v = vb + r * kBlock + tid;
      if (v >= ve) continue;
      const float sum = acc[r].value();
      if (PR) {
        const float out = basep + a.damping * sum;
        y[v] = out;
        const uint32_t d = a.out_deg[v];
        a.contrib_new[v] = d ? out / (float)d : 0.f;
        if (d == 0) dang += (double)out;
        l1 += fabs((double)out - (double)a.rank_old[v]);
      } else {
        y[v] = sum;
      }
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
      a.part[(uint64_t)j * a.ntiles + t] = make_double2(td, tl);
    }
  }
}

// Fixed-order FP64 reduction of per-workgroup partials into stats_out; the
// last block re-arms the counter.
__global__ __launch_bounds__(kBlock) void k_stats_reduce(const double2 *__restrict__ part,
                                                         uint32_t count, double2 *fin_part,
                                                         uint32_t *fin_counter,
                                                         double *stats_out) {
  __shared__ double s_d[kWaves], s_l[kWaves];
  __shared__ uint32_t s_last;
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  double d = 0.0, l = 0.0;
  for (uint32_t c = blockIdx.x * kBlock + tid; c < count; c += gridDim.x * kBlock) {
    d += part[c].x;
    l += part[c].y;
  }
  d = eg_wave_sum_f64(d);
  l = eg_wave_sum_f64(l);
  if (lane == 0) {
    s_d[wave] = d;
    s_l[wave] = l;
  }
  __syncthreads();
  if (tid == 0) {
    double td = 0.0, tl = 0.0;
    for (int q = 0; q < kWaves; q++) {
      td += s_d[q];
      tl += s_l[q];
    }
    fin_part[blockIdx.x] = make_double2(td, tl);
    __threadfence();
    s_last = atomicAdd(fin_counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (s_last && tid == 0) {
    __threadfence();
    double td = 0.0, tl = 0.0;
    for (uint32_t b = 0; b < gridDim.x; b++) {
      const double *pp = reinterpret_cast<const double *>(&fin_part[b]);
      const double2 p = make_double2(__builtin_nontemporal_load(pp), __builtin_nontemporal_load(pp + 1));
      td += p.x;
      tl += p.y;
    }
    stats_out[0] = td;
    stats_out[1] = tl;
    *fin_counter = 0u;
  }
}

// ---- edge array ----
// Adds every record's FP32 term (x[src], or the FP32 product w * x[src]) to
// its destination's FP64 accumulator: the terms are added in whatever order
// the atomics land, but in FP64 the sum of up to 2^29 FP32 terms is within a
// few FP64 ulps of exact, far below the FP32 rounding it is then given.
template <bool WEIGHTED>
__global__ void k_edge_push(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            const float *__restrict__ w, uint64_t m, const float *__restrict__ x,
                            double *__restrict__ y) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  float v = x[src[k]];
  if (WEIGHTED) v *= w[k];
  unsafeAtomicAdd(&y[dst[k]], (double)v);
}

// y[v] = the FP32 rounding of the FP64 sum; the accumulator is cleared for
// the next pass.
__global__ void k_edge_round(uint32_t n, double *__restrict__ acc, float *__restrict__ y) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  y[v] = (float)acc[v];
  acc[v] = 0.0;
}

__global__ void k_pr_apply(uint32_t n, float base, float damping, const double *dmass_in,
                           double damping_over_n, int uniform,
                           const uint32_t *__restrict__ out_deg, double *__restrict__ acc,
                           const float *__restrict__ rank_old, float *__restrict__ rank_new,
                           float *__restrict__ contrib_new, double2 *__restrict__ part) {
  __shared__ double s_d[kWaves], s_l[kWaves];
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  double dang = 0.0, l1 = 0.0;
  double b = (double)base;
  if (uniform) b += damping_over_n * (*dmass_in);
  const float basep = (float)b;
  if (v < n) {
    const float out = basep + damping * (float)acc[v];
    acc[v] = 0.0;
    rank_new[v] = out;
    const uint32_t d = out_deg[v];
    contrib_new[v] = d ? out / (float)d : 0.f;
    if (d == 0) dang = out;
    l1 = fabs((double)out - (double)rank_old[v]);
  }
  dang = eg_wave_sum_f64(dang);
  l1 = eg_wave_sum_f64(l1);
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_d[wave] = dang;
    s_l[wave] = l1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double td = 0.0, tl = 0.0;
    for (int q = 0; q < kWaves; q++) {
      td += s_d[q];
      tl += s_l[q];
    }
    part[blockIdx.x] = make_double2(td, tl);
  }
}

}  // namespace

// Per-graph state of the grid compute copy.
struct GridCompute {
  uint32_t *csrc = nullptr, *cdst = nullptr, *tile_off = nullptr;
  float *cw = nullptr;
  uint32_t ntiles = 0;
  double2 *part = nullptr;
  double2 *fin_part = nullptr;
  uint32_t *fin_counter = nullptr;
  double *acc = nullptr;  // edge-array accumulator (FP64, cleared after every pass)
  float2 *carry = nullptr;  // grid: accumulators carried between source bands
  double2 *apart = nullptr;
  uint32_t apart_n = 0;
};

void eg_pr_resolve(const eg_graph *g, const eg_pr_opts *o, double *damping, int *iters,
                   double *init, int *uniform);

static GridCompute *grid_state(eg_graph *g, bool create) {
  if (!g->gridc && create) g->gridc = new GridCompute();
  return g->gridc;
}

void eg_grid_free(eg_graph *g) {
  GridCompute *c = grid_state(g, false);
  if (!c) return;
  eg_dfree_t(&c->csrc);
  eg_dfree_t(&c->cdst);
  eg_dfree_t(&c->cw);
  eg_dfree_t(&c->tile_off);
  eg_dfree_t(&c->part);
  eg_dfree_t(&c->fin_part);
  eg_dfree_t(&c->fin_counter);
  eg_dfree_t(&c->acc);
  eg_dfree_t(&c->carry);
  eg_dfree_t(&c->apart);
  g->gridc = nullptr;
  delete c;
}

static int grid_build_compute(eg_graph *g, GridCompute *c) {
  if (c->tile_off) return EG_OK;
  const uint64_t m = g->m;
  const uint32_t n = g->n, P = g->P;
  const uint32_t ncells = P * P;
  const uint32_t psize_max = n / P + 1;
  const bool w = g->g_w != nullptr;
  hipStream_t s = g->stream;
  c->ntiles = (psize_max + kTile - 1) / kTile;
  // (cell, src, input) order -> stable by dst -> stable by cell gives
  // (cell, dst, src, input).
  uint32_t *ts = nullptr, *td = nullptr, *tk = nullptr;
  float *tw = nullptr;
  EG_TRY(eg_dalloc_t(&ts, m, "grid tmp src"));
  EG_TRY(eg_dalloc_t(&td, m, "grid tmp dst"));
  EG_TRY(eg_dalloc_t(&tk, m, "grid tmp keys"));
  if (w) EG_TRY(eg_dalloc_t(&tw, m, "grid tmp w"));
  EG_TRY(eg_dalloc_t(&c->csrc, m, "grid compute src"));
  EG_TRY(eg_dalloc_t(&c->cdst, m, "grid compute dst"));
  if (w) EG_TRY(eg_dalloc_t(&c->cw, m, "grid compute w"));
  void *tmp = nullptr;
  EG_TRY(eg_scratch(g, eg_radix_tmp_bytes(m, 3), &tmp));
  {
    EgSortCols cols;
    cols.np = w ? 2 : 1;
    cols.in[0] = g->g_src;
    cols.out[0] = ts;
    if (w) {
      cols.in[1] = reinterpret_cast<const uint32_t *>(g->g_w);
      cols.out[1] = reinterpret_cast<uint32_t *>(tw);
    }
    EG_TRY(eg_radix_sort(s, m, eg_bits_for(n), g->g_dst, td, cols, tmp));
  }
  EG_TRY(eg_cell_keys(s, ts, td, m, n, P, tk));
  {
    EgSortCols cols;
    cols.np = w ? 3 : 2;
    cols.in[0] = ts;
    cols.out[0] = c->csrc;
    cols.in[1] = td;
    cols.out[1] = c->cdst;
    if (w) {
      cols.in[2] = reinterpret_cast<const uint32_t *>(tw);
      cols.out[2] = reinterpret_cast<uint32_t *>(c->cw);
    }
    EG_TRY(eg_radix_sort(s, m, eg_bits_for(ncells), tk, nullptr, cols, tmp));
  }
  EG_HIP(hipStreamSynchronize(s));
  eg_dfree_t(&ts);
  eg_dfree_t(&td);
  eg_dfree_t(&tk);
  eg_dfree_t(&tw);
  const uint64_t noff = (uint64_t)ncells * (c->ntiles + 1);
  EG_TRY(eg_dalloc_t(&c->tile_off, noff, "tile offsets"));
  hipLaunchKernelGGL(k_grid_tile_offsets, dim3(blocks_for(noff, 256)), dim3(256), 0, s,
                     g->cell_ptr, c->cdst, ncells, n, P, c->ntiles, c->tile_off);
  EG_LAUNCHED();
  EG_TRY(eg_dalloc_t(&c->part, (uint64_t)P * c->ntiles, "grid partials"));
  EG_TRY(eg_dalloc_t(&c->carry, n, "grid carried sums"));
  EG_TRY(eg_dalloc_t(&c->fin_part, 256, "grid fin partials"));
  EG_TRY(eg_dalloc_t(&c->fin_counter, 1, "grid counter"));
  EG_HIP(hipMemsetAsync(c->fin_counter, 0, 4, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

// One pass over the grid in bands of g->tune.grid_band source partitions
// (0: all partitions in one launch).
template <bool PR, bool WEIGHTED>
static int grid_pass(eg_graph *g, GridCompute *c, const float *x, float *y, const GridArgs &a) {
  const uint32_t P = g->P;
  const uint32_t band = g->tune.grid_band > 0 && (uint32_t)g->tune.grid_band < P
                            ? (uint32_t)g->tune.grid_band
                            : P;
  for (uint32_t i0 = 0; i0 < P; i0 += band) {
    const uint32_t i1 = i0 + band < P ? i0 + band : P;
    hipLaunchKernelGGL((k_grid_pull<PR, WEIGHTED>), dim3(c->ntiles, P), dim3(kBlock), 0,
                       g->stream, x, y, a, i0, i1, c->carry);
    EG_LAUNCHED();
  }
  return EG_OK;
}

int eg_pr_state(eg_graph *g, int iters);
int eg_pr_start(eg_graph *g, double init, const uint32_t *deg, float *rank, float *contrib);
int eg_pr_finish(eg_graph *g, int iters, float *rank, float *spare, float *rank_out, double *dh,
                 double *lh, const uint32_t *unperm);
static int pr_state(eg_graph *g, int iters) { return eg_pr_state(g, iters); }

// The cells' entries as an edge list for the binned kernels: (dst, src)
// pairs of the grid or of the edge array.
static PbSource cells_source(const eg_graph *g, bool grid) {
  PbSource src;
  src.row = grid ? g->g_dst : g->el_dst;
  src.idx = grid ? g->g_src : g->el_src;
  src.w = grid ? g->g_w : g->el_w;
  return src;
}

// PageRank over the cells' entries on the binned kernels (exact row sums;
// the same recurrence and statistics as the adjacency driver).
static int pagerank_binned_cells(eg_graph *g, int iters, double damping, double init,
                                 int uniform, float *rank_out, double *dh, double *lh, bool grid) {
  EG_TRY(pr_state(g, iters));
  const uint32_t n = g->n;
  EG_TRY(eg_pb_build(g, cells_source(g, grid)));
  const int S = eg_pb_scale(std::max(init * (double)n, 1.0) * 1.001);
  float *rank = g->rank_a, *rank_next = g->rank_b;
  float *contrib = g->contrib_a, *contrib_next = g->contrib_b;
  EG_TRY(eg_pr_start(g, init, g->out_deg, rank, contrib));
  PrArgs a;
  a.out_deg = g->out_deg;
  a.base_vec = nullptr;
  a.base = (float)((1.0 - damping) / (double)n);
  a.damping = (float)damping;
  a.damping_over_n = damping / (double)n;
  a.uniform = uniform;
  a.hot = 0xFFFFFFFFu;
  for (int it = 0; it < iters; it++) {
    a.rank_old = rank;
    a.contrib_new = contrib_next;
    a.dmass_in = g->pr_stats + 2 * it;
    a.stats_out = g->pr_stats + 2 * (it + 1);
    EG_TRY(eg_pb_iterate(g, contrib, rank_next, a, S, nullptr, !uniform && it >= 2));
    std::swap(rank, rank_next);
    std::swap(contrib, contrib_next);
  }
  return eg_pr_finish(g, iters, rank, rank_next, rank_out, dh, lh, nullptr);
}

// Shared PageRank driver for the grid and edge-array layouts.
static int pagerank_cells(eg_graph *g, const eg_pr_opts *o, float *rank_out, double *dh,
                          double *lh, bool grid) {
  double damping, init;
  int iters, uniform;
  eg_pr_resolve(g, o, &damping, &iters, &init, &uniform);
  if (iters < 0 || !(damping >= 0.0 && damping <= 1.0))
    return eg_fail(EG_EINVAL, "bad PageRank options");
  if (grid && !g->cell_ptr) return eg_fail(EG_ENOTSUP, "load mode %d has no grid", g->load_mode);
  if (!grid && !g->el_src)
    return eg_fail(EG_ENOTSUP, "load mode %d has no edge array", g->load_mode);
  if (eg_pb_selected(g, iters, cells_source(g, grid).row))
    return pagerank_binned_cells(g, iters, damping, init, uniform, rank_out, dh, lh, grid);
  GridCompute *c = grid_state(g, true);
  if (grid) EG_TRY(grid_build_compute(g, c));
  EG_TRY(pr_state(g, iters));
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  // each allocated on its own: an edge-array SpMV may have made the
  // accumulator alone
  if (!grid && !c->acc) {
    EG_TRY(eg_dalloc_t(&c->acc, n, "edge accumulator"));
    EG_HIP(hipMemsetAsync(c->acc, 0, (size_t)n * 8, s));
  }
  if (!grid && !c->apart) {
    c->apart_n = blocks_for(n, kBlock);
    EG_TRY(eg_dalloc_t(&c->apart, c->apart_n, "apply partials"));
  }
  if (!grid && !c->fin_part) {
    EG_TRY(eg_dalloc_t(&c->fin_part, 256, "fin partials"));
    EG_TRY(eg_dalloc_t(&c->fin_counter, 1, "fin counter"));
    EG_HIP(hipMemsetAsync(c->fin_counter, 0, 4, s));
  }
  float *rank = g->rank_a, *rank_next = g->rank_b;
  float *contrib = g->contrib_a, *contrib_next = g->contrib_b;
  EG_TRY(eg_pr_start(g, init, g->out_deg, rank, contrib));
  const float base = (float)((1.0 - damping) / (double)n);
  for (int it = 0; it < iters; it++) {
    const double *dmass_in = g->pr_stats + 2 * it;
    double *stats_out = g->pr_stats + 2 * (it + 1);
    if (grid) {
      GridArgs a;
      a.tile_off = c->tile_off;
      a.csrc = c->csrc;
      a.cdst = c->cdst;
      a.cw = c->cw;
      a.n = n;
      a.P = g->P;
      a.ntiles = c->ntiles;
      a.rank_old = rank;
      a.contrib_new = contrib_next;
      a.out_deg = g->out_deg;
      a.dmass_in = dmass_in;
      a.part = c->part;
      a.base = base;
      a.damping = (float)damping;
      a.damping_over_n = damping / (double)n;
      a.uniform = uniform;
      {
        EgProf prof(g, "pr_grid");
        EG_TRY((grid_pass<true, false>(g, c, contrib, rank_next, a)));
      }
      hipLaunchKernelGGL(k_stats_reduce, dim3(64), dim3(kBlock), 0, s, c->part,
                         g->P * c->ntiles, c->fin_part, c->fin_counter, stats_out);
      EG_LAUNCHED();
    } else {
      {
        EgProf prof(g, "pr_edges");
        hipLaunchKernelGGL((k_edge_push<false>), dim3(blocks_for(g->m, 256)), dim3(256), 0, s,
                           g->el_src, g->el_dst, (const float *)nullptr, g->m, contrib, c->acc);
        EG_LAUNCHED();
      }
      hipLaunchKernelGGL(k_pr_apply, dim3(c->apart_n), dim3(kBlock), 0, s, n, base,
                         (float)damping, dmass_in, damping / (double)n, uniform, g->out_deg,
                         c->acc, rank, rank_next, contrib_next, c->apart);
      EG_LAUNCHED();
      hipLaunchKernelGGL(k_stats_reduce, dim3(64), dim3(kBlock), 0, s, c->apart, c->apart_n,
                         c->fin_part, c->fin_counter, stats_out);
      EG_LAUNCHED();
    }
    std::swap(rank, rank_next);
    std::swap(contrib, contrib_next);
  }
  return eg_pr_finish(g, iters, rank, rank_next, rank_out, dh, lh, nullptr);
}

int eg_pagerank_grid(eg_graph *g, const eg_pr_opts *o, float *rank_out, double *dh, double *lh) {
  return pagerank_cells(g, o, rank_out, dh, lh, true);
}

int eg_pagerank_edges(eg_graph *g, const eg_pr_opts *o, float *rank_out, double *dh, double *lh) {
  return pagerank_cells(g, o, rank_out, dh, lh, false);
}

int eg_spmv_grid_device(eg_graph *g, const float *x, float *y) {
  GridCompute *c = grid_state(g, true);
  EG_TRY(grid_build_compute(g, c));
  GridArgs a = {};
  a.tile_off = c->tile_off;
  a.csrc = c->csrc;
  a.cdst = c->cdst;
  a.cw = c->cw;
  a.n = g->n;
  a.P = g->P;
  a.ntiles = c->ntiles;
  EgProf prof(g, "spmv_grid");
  if (c->cw) return grid_pass<false, true>(g, c, x, y, a);
  return grid_pass<false, false>(g, c, x, y, a);
}

static int spmv_host(eg_graph *g, const float *x, float *y, bool grid) {
  if (grid && !g->cell_ptr) return eg_fail(EG_ENOTSUP, "load mode %d has no grid", g->load_mode);
  if (!grid && !g->el_src)
    return eg_fail(EG_ENOTSUP, "load mode %d has no edge array", g->load_mode);
  EG_TRY(pr_state(g, 0));
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  EG_HIP(hipMemcpyAsync(g->contrib_a, x, (size_t)n * 4, hipMemcpyHostToDevice, s));
  // Binned kernels sum exactly; their scale needs a finite bound on |x|.
  double xmax = 0.0;
  bool finite = true;
  const bool sel = eg_pb_selected(g, 1, cells_source(g, grid).row);
  if (sel)
    for (uint32_t i = 0; i < n && finite; i++) {
      const double ax = fabs((double)x[i]);
      finite = std::isfinite(ax);
      xmax = std::max(xmax, ax);
    }
  bool binned = sel && finite;
  if (binned) {
    EG_TRY(eg_pb_build(g, cells_source(g, grid)));
    binned = eg_pb_spmv_exact(g, xmax);
  }
  if (binned) {
    EG_TRY(eg_pb_spmv(g, g->contrib_a, g->rank_a, xmax));
  } else if (grid) {
    EG_TRY(eg_spmv_grid_device(g, g->contrib_a, g->rank_a));
  } else {
    GridCompute *c = grid_state(g, true);
    if (!c->acc) {
      EG_TRY(eg_dalloc_t(&c->acc, n, "edge accumulator"));
      EG_HIP(hipMemsetAsync(c->acc, 0, (size_t)n * 8, s));
    }
    EgProf prof(g, "spmv_edges");
    if (g->el_w)
      hipLaunchKernelGGL((k_edge_push<true>), dim3(blocks_for(g->m, 256)), dim3(256), 0, s,
                         g->el_src, g->el_dst, g->el_w, g->m, g->contrib_a, c->acc);
    else
      hipLaunchKernelGGL((k_edge_push<false>), dim3(blocks_for(g->m, 256)), dim3(256), 0, s,
                         g->el_src, g->el_dst, (const float *)nullptr, g->m, g->contrib_a,
                         c->acc);
    EG_LAUNCHED();
    hipLaunchKernelGGL(k_edge_round, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, c->acc,
                       g->rank_a);
    EG_LAUNCHED();
  }
  if (y) EG_HIP(hipMemcpyAsync(y, g->rank_a, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  return EG_OK;
}

int eg_spmv_grid(eg_graph *g, const float *x, float *y) { return spmv_host(g, x, y, true); }
int eg_spmv_edges(eg_graph *g, const float *x, float *y) { return spmv_host(g, x, y, false); }
