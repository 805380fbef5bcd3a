// eg_wcc.hip -- min-label weakly connected components.
//
// Reference: wcc.c.  Labels start at the vertex id (wcc_reset, :128-135) and
// writeMin (:21-27) propagates the minimum to a fixpoint: along out-edges for
// the adjacency modes (wcc_algo, :30-42; with -u this is the component
// minimum), along both directions of every record for the grid and edge-array
// modes (:193-261).
//
// GPU form: every round relaxes all entries with atomicMin on 32-bit labels,
// then shortcuts labels by pointer jumping, L[v] = min(L[v], L[L[v]]).  Both
// steps keep the invariant "L[v] is the id of a vertex that reaches v (is in
// v's component) and L[v] <= v", so the fixpoint is the reference's: for
// out-edge propagation the minimum id among v's ancestors, for two-way
// relaxation the minimum id of v's component.  Two-way relaxation also hooks
// the larger label's vertex, L[L[v]] = min(., L[u]), which is valid because
// both labels lie in the same component.  A round that changes nothing ends
// the loop.
//
// Undirected / symmetric adjacency graphs run on the degree-ordered ids with
// every row hubs first: a round is a pull, each row taking the minimum of its
// neighbours' labels (a lane per row, the whole wave for a row longer than 64
// entries), and a row stops at label 0, which no label undercuts; vertex 0 is
// the highest-degree hub, so after the first round nearly every row of its
// component stops at its first entry.  Labels are then ordered ids (the
// component minimum in that order); a last pass gives every component the
// minimum of its members' caller ids, the reference's label.
#include <algorithm>

#include "eg_internal.h"

namespace {

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

__global__ void k_wcc_init(uint32_t n, uint32_t *__restrict__ L) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) L[v] = v;
}

template <bool TWO_WAY>
__global__ void k_wcc_relax(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst,
                            uint64_t m, uint32_t *__restrict__ L, uint32_t *__restrict__ changed) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (k < m) {
    const uint32_t u = __builtin_nontemporal_load(src + k);
    const uint32_t v = __builtin_nontemporal_load(dst + k);
    const uint32_t lu = __atomic_load_n(&L[u], __ATOMIC_RELAXED);
    const uint32_t lv = __atomic_load_n(&L[v], __ATOMIC_RELAXED);
    if (lu < lv) {
      ch |= atomicMin(&L[v], lu) > lu;
      if (TWO_WAY) ch |= atomicMin(&L[lv], lu) > lu;
    } else if (TWO_WAY && lv < lu) {
      ch |= atomicMin(&L[u], lv) > lv;
      ch |= atomicMin(&L[lu], lv) > lv;
    }
  }
  // one flag per round: only the first waves to see it clear write it
  if (__ballot(ch) && __lane_id() == 0 && __atomic_load_n(changed, __ATOMIC_RELAXED) == 0u)
    atomicOr(changed, 1u);
}

__global__ void k_wcc_shortcut(uint32_t n, uint32_t *__restrict__ L,
                               uint32_t *__restrict__ changed) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (v < n) {
    uint32_t l = __atomic_load_n(&L[v], __ATOMIC_RELAXED);
    uint32_t ll = __atomic_load_n(&L[l], __ATOMIC_RELAXED);
    // follow at most a few hops; later rounds finish the job
    for (int h = 0; h < 8 && ll < l; h++) {
      l = ll;
      ll = __atomic_load_n(&L[l], __ATOMIC_RELAXED);
    }
    ch = atomicMin(&L[v], l) > l;
  }
  // one flag per round: only the first waves to see it clear write it
  if (__ballot(ch) && __lane_id() == 0 && __atomic_load_n(changed, __ATOMIC_RELAXED) == 0u)
    atomicOr(changed, 1u);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}

constexpr uint32_t kPullSerial = 64;  // entries a lane scans before its wave helps

// One pull round over rows [0, n) of (ptr, col): L[v] = min(L[v], L[u] for u
// in row v), 64 rows per wave pass (grid-stride).
__global__ __launch_bounds__(256) void k_wcc_pull(uint32_t n, const uint32_t *__restrict__ ptr,
                                                  const uint32_t *__restrict__ col,
                                                  uint32_t *__restrict__ L,
                                                  uint32_t *__restrict__ changed) {
  const int lane = __lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  bool ch = false;
  for (uint64_t base = (uint64_t)wave * 64; base < n; base += (uint64_t)nwaves * 64) {
    const uint32_t v = (uint32_t)base + (uint32_t)lane;
    uint32_t cur = 0, best = 0, k0 = 0, k1 = 0;
    bool longrow = false;
    if (v < n) {
      cur = best = __atomic_load_n(&L[v], __ATOMIC_RELAXED);
      k0 = ptr[v];
      k1 = ptr[v + 1];
      const uint32_t ke = k1 - k0 > kPullSerial ? k0 + kPullSerial : k1;
      for (uint32_t k = k0; k < ke && best != 0; k += 4) {
        uint32_t u[4];
#pragma unroll
        for (int x = 0; x < 4; x++) u[x] = k + x < ke ? col[k + x] : col[k];
#pragma unroll
        for (int x = 0; x < 4; x++) best = min(best, __atomic_load_n(&L[u[x]], __ATOMIC_RELAXED));
      }
      longrow = best != 0 && ke < k1;
    }
    uint64_t lm = __ballot(longrow);
    while (lm) {
      const int b = __ffsll((unsigned long long)lm) - 1;
      lm &= lm - 1;
      const uint32_t kb0 = __shfl(k0, b) + kPullSerial, kb1 = __shfl(k1, b);
      uint32_t bb = __shfl(best, b);
      for (uint32_t kb = kb0; kb < kb1 && bb != 0; kb += 64) {
        const uint32_t k = kb + (uint32_t)lane;
        const uint32_t x = k < kb1 ? __atomic_load_n(&L[col[k]], __ATOMIC_RELAXED) : 0xFFFFFFFFu;
        bb = min(bb, wave_min_u32(x));
      }
      if (lane == b) best = bb;
    }
    if (v < n && best < cur) ch |= atomicMin(&L[v], best) > best;
  }
  // one flag per round: only the first waves to see it clear write it
  if (__ballot(ch) && __lane_id() == 0 && __atomic_load_n(changed, __ATOMIC_RELAXED) == 0u)
    atomicOr(changed, 1u);
}

// cmin[c] = minimum caller id among the members of component c (c = the
// ordered-id label; cmin starts at all ones).  Label 0 is the component of
// the top hub, usually most vertices: its minimum is kept in a register and
// merged with one atomic per wave (grid-stride), the other labels directly.
__global__ __launch_bounds__(256) void k_wcc_cmin(uint32_t n, const uint32_t *__restrict__ L,
                                                  const uint32_t *__restrict__ id_of,
                                                  uint32_t *__restrict__ cmin) {
  uint32_t m0 = 0xFFFFFFFFu;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t l = L[v], id = id_of[v];
    if (l == 0)
      m0 = min(m0, id);
    else
      atomicMin(&cmin[l], id);
  }
  m0 = wave_min_u32(m0);
  if (__lane_id() == 0 && m0 != 0xFFFFFFFFu) atomicMin(&cmin[0], m0);
}

// out[id_of[v]] = cmin[L[v]]: labels in caller ids.
__global__ void k_wcc_map(uint32_t n, const uint32_t *__restrict__ L,
                          const uint32_t *__restrict__ id_of, const uint32_t *__restrict__ cmin,
                          uint32_t *__restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) out[id_of[v]] = cmin[L[v]];
}

__global__ void k_widen(uint32_t n, const uint32_t *__restrict__ L, int64_t *__restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) out[v] = (int64_t)L[v];
}

}  // namespace

// Undirected / symmetric adjacency graphs on the ordered ids: pull rounds
// and shortcuts on ordered-id labels, then caller-id component minima.
static int wcc_ordered(eg_graph *g, const uint32_t *rows, int32_t *rounds_out) {
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  EgOrdered *o;
  EG_TRY(eg_ordered(g, &o));
  if (!g->wcc_tmp) EG_TRY(eg_dalloc_t(&g->wcc_tmp, 2 * (uint64_t)n, "ordered labels"));
  uint32_t *L = g->wcc_tmp, *cmin = g->wcc_tmp + n;
  hipLaunchKernelGGL(k_wcc_init, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, L);
  EG_LAUNCHED();
  int rounds = 0;
  for (;;) {
    rounds++;
    if ((uint64_t)rounds > (uint64_t)n + 2)
      return eg_fail(EG_EDEVICE, "label propagation did not converge within n rounds");
    EG_HIP(hipMemsetAsync(g->bfs_ctr + 16, 0, 4, s));
    {
      EgProf prof(g, "wcc_relax");
      hipLaunchKernelGGL(k_wcc_pull, dim3(2048), dim3(256), 0, s, n, o->ptr, rows, L,
                         g->bfs_ctr + 16);
      EG_LAUNCHED();
    }
    hipLaunchKernelGGL(k_wcc_shortcut, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, L,
                       g->bfs_ctr + 16);
    EG_LAUNCHED();
    EG_HIP(hipMemcpyAsync(g->h_ctr + 16, g->bfs_ctr + 16, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    if (g->h_ctr[16] == 0) break;
  }
  EG_HIP(hipMemsetAsync(cmin, 0xFF, (size_t)n * 4, s));
  hipLaunchKernelGGL(k_wcc_cmin, dim3(std::min<uint32_t>(blocks_for(n, 256), 2048)), dim3(256), 0,
                     s, n, L, o->old_of_new, cmin);
  EG_LAUNCHED();
  hipLaunchKernelGGL(k_wcc_map, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, L, o->old_of_new,
                     cmin, g->wcc_lab32);
  EG_LAUNCHED();
  if (rounds_out) *rounds_out = rounds;
  return EG_OK;
}

// Runs the label rounds on the device; labels are left in g->wcc_lab32.
int eg_wcc_device(eg_graph *g, int32_t *rounds_out) {
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  if (!g->wcc_lab32) EG_TRY(eg_dalloc_t(&g->wcc_lab32, n, "labels"));
  if (!g->bfs_ctr) EG_TRY(eg_dalloc_t(&g->bfs_ctr, 64, "counters"));
  if (g->out_ptr && n && g->tune.wcc_order) {
    const uint32_t *rows = nullptr;
    EG_TRY(eg_hubs_first_rows(g, &rows));  // null unless undirected / symmetric and ordered
    if (rows) return wcc_ordered(g, rows, rounds_out);
  }
  const uint32_t *src, *dst;
  bool two_way;
  if (g->out_ptr) {
    src = g->out_src;
    dst = g->out_col;
    two_way = false;
  } else if (g->cell_ptr) {
    src = g->g_src;
    dst = g->g_dst;
    two_way = true;
  } else {
    src = g->el_src;
    dst = g->el_dst;
    two_way = true;
  }
  hipLaunchKernelGGL(k_wcc_init, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g->wcc_lab32);
  EG_LAUNCHED();
  int rounds = 0;
  for (;;) {
    rounds++;
    if ((uint64_t)rounds > (uint64_t)n + 2)
      return eg_fail(EG_EDEVICE, "label propagation did not converge within n rounds");
    EG_HIP(hipMemsetAsync(g->bfs_ctr + 16, 0, 4, s));
    if (g->m) {
      EgProf prof(g, "wcc_relax");
      if (two_way)
        hipLaunchKernelGGL(k_wcc_relax<true>, dim3(blocks_for(g->m, 256)), dim3(256), 0, s, src,
                           dst, g->m, g->wcc_lab32, g->bfs_ctr + 16);
      else
        hipLaunchKernelGGL(k_wcc_relax<false>, dim3(blocks_for(g->m, 256)), dim3(256), 0, s, src,
                           dst, g->m, g->wcc_lab32, g->bfs_ctr + 16);
      EG_LAUNCHED();
    }
    hipLaunchKernelGGL(k_wcc_shortcut, dim3(blocks_for(n, 256)), dim3(256), 0, s, n,
                       g->wcc_lab32, g->bfs_ctr + 16);
    EG_LAUNCHED();
    EG_HIP(hipMemcpyAsync(g->h_ctr + 16, g->bfs_ctr + 16, 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    if (g->h_ctr[16] == 0) break;
  }
  if (rounds_out) *rounds_out = rounds;
  return EG_OK;
}

extern "C" int eg_wcc(eg_graph *g, int64_t *labels, int32_t *rounds) {
  EG_TRY(eg_require_loaded(g));
  EG_TRY(eg_wcc_device(g, rounds));
  if (labels) {
    const uint32_t n = g->n;
    if (!g->wcc_label) EG_TRY(eg_dalloc_t(&g->wcc_label, n, "labels64"));
    hipLaunchKernelGGL(k_widen, dim3(blocks_for(n, 256)), dim3(256), 0, g->stream, n,
                       g->wcc_lab32, g->wcc_label);
    EG_LAUNCHED();
    EG_HIP(hipMemcpyAsync(labels, g->wcc_label, (size_t)n * 8, hipMemcpyDeviceToHost, g->stream));
    EG_HIP(hipStreamSynchronize(g->stream));
  }
  return EG_OK;
}
