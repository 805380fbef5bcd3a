// eg_order.hip -- degree-ordered compute copy of the pull orientation.
//
// The pull kernels gather x[u] (PageRank contributions, SpMV inputs, PPR
// rows, MS-BFS frontier masks) for every entry u of a row.  On R-MAT graphs a
// few percent of the vertices are the targets of most gathers, but in the
// natural id order they are spread over the whole vector, so most gathers
// miss the 4 MB L2 of an XCD and fetch a full line from the fabric.  This
// copy relabels vertices by descending out-degree (stable by id), which packs
// the hot set into the first few MB of every gathered vector.
//
// Every row keeps its entries in their original order (a stable sort by the
// new row id), so a row's compensated sum sees the same values in the same
// order as in the reference layout: results are bit-identical, only their
// positions are permuted.  EG_PULL_ORDER=none in the environment disables
// the copy (the kernels then run on the layout as built).
#include <stdlib.h>
#include <string.h>

#include "eg_internal.h"

namespace {

inline uint32_t blocks_for(uint64_t count, uint32_t per) {
  return (uint32_t)((count + per - 1) / per);
}

__global__ void k_order_keys(uint32_t n, const uint32_t *__restrict__ deg,
                             uint32_t *__restrict__ key, uint32_t *__restrict__ iota) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  key[v] = 0xFFFFFFFFu - deg[v];
  iota[v] = v;
}

__global__ void k_invert(uint32_t n, const uint32_t *__restrict__ old_of_new,
                         uint32_t *__restrict__ new_of_old) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) new_of_old[old_of_new[i]] = i;
}

// nkey[p] = new id of the row of entry p (a wave per row, grid-stride).
__global__ void k_row_keys(uint32_t n, const uint32_t *__restrict__ ptr,
                           const uint32_t *__restrict__ new_of_old, uint32_t *__restrict__ nkey) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const int lane = __lane_id();
  for (uint32_t v = wave; v < n; v += nwaves) {
    const uint32_t k1 = ptr[v + 1], nv = new_of_old[v];
    for (uint32_t k = ptr[v] + lane; k < k1; k += 64) nkey[k] = nv;
  }
}

__global__ void k_map(uint64_t m, const uint32_t *__restrict__ in,
                      const uint32_t *__restrict__ map, uint32_t *__restrict__ out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) out[k] = map[in[k]];
}

__global__ void k_gather_u32(uint32_t n, const uint32_t *__restrict__ src,
                             const uint32_t *__restrict__ perm, uint32_t *__restrict__ dst) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

}  // namespace

bool eg_reorder_enabled() {
  static int cached = -1;
  if (cached < 0) {
    const char *e = getenv("EG_PULL_ORDER");
    cached = (e && strcmp(e, "none") == 0) ? 0 : 1;
  }
  return cached == 1;
}

bool eg_use_order(const eg_graph *g, int iterations) {
  if (!eg_reorder_enabled()) return false;
  return (g->ord && g->ord->built) || iterations >= kAutoOrderIters;
}

void eg_order_free(eg_graph *g) {
  EgOrdered *o = g->ord;
  if (!o) return;
  eg_dfree_t(&o->old_of_new);
  eg_dfree_t(&o->new_of_old);
  eg_dfree_t(&o->ptr);
  eg_dfree_t(&o->idx);
  eg_dfree_t(&o->w);
  eg_dfree_t(&o->deg);
  eg_plan_free(&o->plan);
  delete o;
  g->ord = nullptr;
}

// dst[i] = src[perm[i]] for float / u32 vectors of length n.
int eg_gather_vec(eg_graph *g, const void *src, const uint32_t *perm, void *dst,
                  hipStream_t stream) {
  hipLaunchKernelGGL(k_gather_u32, dim3(blocks_for(g->n, 256)), dim3(256), 0,
                     stream ? stream : g->stream, g->n,
                     reinterpret_cast<const uint32_t *>(src), perm,
                     reinterpret_cast<uint32_t *>(dst));
  EG_LAUNCHED();
  return EG_OK;
}

int eg_ordered(eg_graph *g, EgOrdered **out) {
  if (g->ord && g->ord->built) {
    *out = g->ord;
    return EG_OK;
  }
  const uint32_t *ptr, *idx;
  const float *w;
  PullPlan *unused;
  // the orientation itself (its own plan is not needed)
  if (g->in_ptr) {
    ptr = g->in_ptr;
    idx = g->in_idx;
    w = g->in_w;
  } else if (g->out_ptr) {
    ptr = g->out_ptr;
    idx = g->out_col;
    w = g->out_w;
  } else {
    return eg_fail(EG_ENOTSUP, "no adjacency layout");
  }
  (void)unused;
  eg_order_free(g);
  const double t0 = eg_wall_ms();
  EgOrdered *o = new EgOrdered();
  g->ord = o;
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  hipStream_t s = g->stream;
  EG_TRY(eg_dalloc_t(&o->old_of_new, n, "order old_of_new"));
  EG_TRY(eg_dalloc_t(&o->new_of_old, n, "order new_of_old"));
  EG_TRY(eg_dalloc_t(&o->ptr, (uint64_t)n + 1, "order ptr"));
  EG_TRY(eg_dalloc_t(&o->idx, m, "order idx"));
  if (w) EG_TRY(eg_dalloc_t(&o->w, m, "order w"));
  EG_TRY(eg_dalloc_t(&o->deg, n, "order deg"));
  uint32_t *key = nullptr, *iota = nullptr, *nkey = nullptr, *nval = nullptr, *skey = nullptr;
  int rc = EG_OK;
  do {
    // 1. vertex order: descending out-degree, stable by id
    if ((rc = eg_dalloc_t(&key, n, "order keys")) || (rc = eg_dalloc_t(&iota, n, "order iota")))
      break;
    hipLaunchKernelGGL(k_order_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g->out_deg,
                       key, iota);
    if ((rc = hipGetLastError() == hipSuccess ? EG_OK : eg_fail(EG_EDEVICE, "order keys launch")))
      break;
    void *tmp = nullptr;
    EgSortCols c;
    c.np = 1;
    c.in[0] = iota;
    c.out[0] = o->old_of_new;
    if ((rc = eg_scratch(g, eg_radix_tmp_bytes(n, 1), &tmp))) break;
    if ((rc = eg_radix_sort(s, n, 32, key, nullptr, c, tmp))) break;
    hipLaunchKernelGGL(k_invert, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, o->old_of_new,
                       o->new_of_old);
    hipLaunchKernelGGL(k_gather_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g->out_deg,
                       o->old_of_new, o->deg);
    // 2. relabelled entries, stably sorted by new row id
    if ((rc = eg_dalloc_t(&nkey, m, "order row keys")) || (rc = eg_dalloc_t(&nval, m, "order vals")) ||
        (rc = eg_dalloc_t(&skey, m, "order sorted keys")))
      break;
    hipLaunchKernelGGL(k_row_keys, dim3(blocks_for((uint64_t)n * 64, 256) < 65536
                                            ? blocks_for((uint64_t)n * 64, 256)
                                            : 65536),
                       dim3(256), 0, s, n, ptr, o->new_of_old, nkey);
    if (m)
      hipLaunchKernelGGL(k_map, dim3(blocks_for(m, 256)), dim3(256), 0, s, m, idx, o->new_of_old,
                         nval);
    if ((rc = hipGetLastError() == hipSuccess ? EG_OK : eg_fail(EG_EDEVICE, "order launch")))
      break;
    EgSortCols c2;
    c2.np = w ? 2 : 1;
    c2.in[0] = nval;
    c2.out[0] = o->idx;
    if (w) {
      c2.in[1] = reinterpret_cast<const uint32_t *>(w);
      c2.out[1] = reinterpret_cast<uint32_t *>(o->w);
    }
    if ((rc = eg_scratch(g, eg_radix_tmp_bytes(m, c2.np), &tmp))) break;
    if ((rc = eg_radix_sort(s, m, eg_bits_for(n), nkey, skey, c2, tmp))) break;
    if ((rc = eg_ptr_from_sorted(s, skey, m, n, o->ptr, nullptr))) break;
  } while (0);
  (void)hipStreamSynchronize(s);
  eg_dfree_t(&key);
  eg_dfree_t(&iota);
  eg_dfree_t(&nkey);
  eg_dfree_t(&nval);
  eg_dfree_t(&skey);
  if (rc) {
    eg_order_free(g);
    return rc;
  }
  EG_TRY(eg_plan_build(g, &o->plan, o->ptr));
  o->built = true;
  g->prep_order_ms = eg_wall_ms() - t0;
  // permutations, row pointers, entries (+ weights), degrees, plan chunks
  g->prep_order_bytes = 4ull * 3 * n + 4ull * (n + 1) + 4ull * m * (w ? 2 : 1) +
                        16ull * o->plan.nchunks + 24ull * o->plan.nseg;
  *out = o;
  return EG_OK;
}
