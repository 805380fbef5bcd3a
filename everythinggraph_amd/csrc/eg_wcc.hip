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

__global__ void k_widen(uint32_t n, const uint32_t *__restrict__ L, int64_t *__restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) out[v] = (int64_t)L[v];
}

}  // namespace

// Runs the label rounds on the device; labels are left in g->wcc_lab32.
int eg_wcc_device(eg_graph *g, int32_t *rounds_out) {
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  if (!g->wcc_lab32) EG_TRY(eg_dalloc_t(&g->wcc_lab32, n, "labels"));
  if (!g->bfs_ctr) EG_TRY(eg_dalloc_t(&g->bfs_ctr, 64, "counters"));
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
