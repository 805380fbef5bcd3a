// eg_build.hip -- layout builders of libeg (the load modes of init(),
// init_all.c:855-888).
//
// Every layout is a stable sort of the records (input order, reversed copies
// appended with -u) by the layout's key, run on the device with the LSD radix
// sort of eg_prims.hip.  Row / cell offsets and degrees are the run offsets of
// the sorted keys (one binary search per vertex), the contract of the
// reference's boundary detection + scanIBack(min) (init_all.c:675-697): no
// degree is accumulated with atomics, so R-MAT hubs cost nothing extra.
#include <mutex>

#include "eg_internal.h"

static int sort_by(eg_graph *g, uint64_t m, uint64_t nkeys, const uint32_t *kin, uint32_t *kout,
                   EgSortCols cols) {
  void *tmp = nullptr;
  EG_TRY(eg_scratch(g, eg_radix_tmp_bytes(m, cols.np), &tmp));
  return eg_radix_sort(g->stream, m, eg_bits_for(nkeys), kin, kout, cols, tmp);
}

static EgSortCols cols2(const uint32_t *i0, uint32_t *o0, const float *i1, float *o1) {
  EgSortCols c;
  c.np = 1;
  c.in[0] = i0;
  c.out[0] = o0;
  if (i1) {
    c.np = 2;
    c.in[1] = reinterpret_cast<const uint32_t *>(i1);
    c.out[1] = reinterpret_cast<uint32_t *>(o1);
  }
  return c;
}

// deg[v] = number of keys equal to v (keys-only sort, then run lengths).
static int degrees_by_sort(eg_graph *g, const uint32_t *keys, uint32_t *deg) {
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  uint32_t *sorted = nullptr, *ptr = nullptr;
  int rc = EG_OK;
  do {
    if ((rc = eg_dalloc_t(&sorted, m, "sorted keys")) ||
        (rc = eg_dalloc_t(&ptr, (uint64_t)n + 1, "run offsets")))
      break;
    if ((rc = sort_by(g, m, n, keys, sorted, EgSortCols()))) break;
    rc = eg_ptr_from_sorted(g->stream, sorted, m, n, ptr, deg);
  } while (0);
  (void)hipStreamSynchronize(g->stream);
  eg_dfree_t(&sorted);
  eg_dfree_t(&ptr);
  return rc;
}

static int degrees_from_sorted(eg_graph *g, const uint32_t *sorted, uint32_t *deg) {
  uint32_t *ptr = nullptr;
  EG_TRY(eg_dalloc_t(&ptr, (uint64_t)g->n + 1, "run offsets"));
  int rc = eg_ptr_from_sorted(g->stream, sorted, g->m, g->n, ptr, deg);
  (void)hipStreamSynchronize(g->stream);
  eg_dfree_t(&ptr);
  return rc;
}

// init_adj_sort (init_all.c:629-790) and init_adj_count (:444-496): stable sort
// by src with the dst (and weight) as payload; row_ptr from the sorted src;
// CSC = stable sort of the CSR entries by dst (the src is its payload).
static int build_adjacency(eg_graph *g) {
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  const bool w = (g->flags & EG_WEIGHTED) != 0;
  EG_TRY(eg_dalloc_t(&g->out_ptr, (uint64_t)n + 1, "row_ptr"));
  EG_TRY(eg_dalloc_t(&g->out_col, m, "col"));
  EG_TRY(eg_dalloc_t(&g->out_src, m, "row ids"));
  if (w) EG_TRY(eg_dalloc_t(&g->out_w, m, "weights"));
  EG_TRY(sort_by(g, m, n, g->e_src, g->out_src, cols2(g->e_dst, g->out_col, g->e_w, g->out_w)));
  EG_TRY(eg_ptr_from_sorted(g->stream, g->out_src, m, n, g->out_ptr, g->out_deg));
  if (g->flags & EG_SYMMETRIC) {
    // -s declares the graph symmetric: no in-edges are built (init_all.c:714)
    EG_HIP(hipMemcpyAsync(g->in_deg, g->out_deg, (size_t)n * 4, hipMemcpyDeviceToDevice,
                          g->stream));
    return EG_OK;
  }
  EG_TRY(eg_dalloc_t(&g->in_ptr, (uint64_t)n + 1, "col_ptr"));
  EG_TRY(eg_dalloc_t(&g->in_idx, m, "in idx"));
  if (w) EG_TRY(eg_dalloc_t(&g->in_w, m, "in weights"));
  // the record arrays are free now: e_src receives the sorted dst keys
  EG_TRY(sort_by(g, m, n, g->out_col, g->e_src, cols2(g->out_src, g->in_idx, g->out_w, g->in_w)));
  return eg_ptr_from_sorted(g->stream, g->e_src, m, n, g->in_ptr, g->in_deg);
}

// init_grid_sort (init_all.c:252-361, modes 0/6) and init_grid_nosort
// (:363-424, mode 1): records ordered by cell (part(src), part(dst)), with
// (src, input order) inside a cell for the sorted modes and input order for
// mode 1.  cell_ptr is the exact run offset of every cell.
static int build_grid(eg_graph *g) {
  const uint64_t m = g->m;
  const uint32_t n = g->n, P = g->P;
  const bool w = (g->flags & EG_WEIGHTED) != 0;
  const bool presort = g->load_mode != EG_MODE_GRID_NOSORT;
  const uint64_t ncells = (uint64_t)P * P;
  uint32_t *ts = nullptr, *td = nullptr, *tk = nullptr, *tks = nullptr;
  float *tw = nullptr;
  int rc = EG_OK;
  const uint32_t *s_src = g->e_src, *s_dst = g->e_dst;
  const float *s_w = g->e_w;
  do {
    if ((rc = degrees_by_sort(g, g->e_dst, g->in_deg))) break;
    if (presort) {
      if ((rc = eg_dalloc_t(&ts, m, "grid tmp src")) || (rc = eg_dalloc_t(&td, m, "grid tmp dst")))
        break;
      if (w && (rc = eg_dalloc_t(&tw, m, "grid tmp w"))) break;
      if ((rc = sort_by(g, m, n, g->e_src, ts, cols2(g->e_dst, td, g->e_w, tw)))) break;
      if ((rc = degrees_from_sorted(g, ts, g->out_deg))) break;
      s_src = ts;
      s_dst = td;
      s_w = tw;
    } else if ((rc = degrees_by_sort(g, g->e_src, g->out_deg))) {
      break;
    }
    if ((rc = eg_dalloc_t(&tk, m, "cell keys")) || (rc = eg_dalloc_t(&tks, m, "sorted cell keys")))
      break;
    if ((rc = eg_cell_keys(g->stream, s_src, s_dst, m, n, P, tk))) break;
    if ((rc = eg_dalloc_t(&g->g_src, m, "grid src")) || (rc = eg_dalloc_t(&g->g_dst, m, "grid dst")))
      break;
    if (w && (rc = eg_dalloc_t(&g->g_w, m, "grid w"))) break;
    EgSortCols c;
    c.np = w ? 3 : 2;
    c.in[0] = s_src;
    c.out[0] = g->g_src;
    c.in[1] = s_dst;
    c.out[1] = g->g_dst;
    if (w) {
      c.in[2] = reinterpret_cast<const uint32_t *>(s_w);
      c.out[2] = reinterpret_cast<uint32_t *>(g->g_w);
    }
    if ((rc = sort_by(g, m, ncells, tk, tks, c))) break;
    if ((rc = eg_dalloc_t(&g->cell_ptr, ncells + 1, "cell_ptr"))) break;
    rc = eg_ptr_from_sorted(g->stream, tks, m, (uint32_t)ncells, g->cell_ptr, nullptr);
  } while (0);
  (void)hipStreamSynchronize(g->stream);
  eg_dfree_t(&ts);
  eg_dfree_t(&td);
  eg_dfree_t(&tw);
  eg_dfree_t(&tk);
  eg_dfree_t(&tks);
  return rc;
}

// init_edgelist (init_all.c:573-628): mode 7 = stable by src, mode 8 = by src
// then dst (LSD: stable by dst, then stable by src), mode 9 = input order.
static int build_edgelist(eg_graph *g) {
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  const bool w = (g->flags & EG_WEIGHTED) != 0;
  EG_TRY(degrees_by_sort(g, g->e_dst, g->in_deg));
  EG_TRY(eg_dalloc_t(&g->el_src, m, "edge src"));
  EG_TRY(eg_dalloc_t(&g->el_dst, m, "edge dst"));
  if (w) EG_TRY(eg_dalloc_t(&g->el_w, m, "edge w"));
  if (g->load_mode == EG_MODE_EDGES_RAW) {
    EG_HIP(hipMemcpyAsync(g->el_src, g->e_src, m * 4, hipMemcpyDeviceToDevice, g->stream));
    EG_HIP(hipMemcpyAsync(g->el_dst, g->e_dst, m * 4, hipMemcpyDeviceToDevice, g->stream));
    if (w) EG_HIP(hipMemcpyAsync(g->el_w, g->e_w, m * 4, hipMemcpyDeviceToDevice, g->stream));
    return degrees_by_sort(g, g->e_src, g->out_deg);
  }
  if (g->load_mode == EG_MODE_EDGES_SRC) {
    EG_TRY(sort_by(g, m, n, g->e_src, g->el_src, cols2(g->e_dst, g->el_dst, g->e_w, g->el_w)));
    return degrees_from_sorted(g, g->el_src, g->out_deg);
  }
  uint32_t *ts = nullptr, *td = nullptr;
  float *tw = nullptr;
  int rc = EG_OK;
  do {
    if ((rc = eg_dalloc_t(&ts, m, "tmp src")) || (rc = eg_dalloc_t(&td, m, "tmp dst"))) break;
    if (w && (rc = eg_dalloc_t(&tw, m, "tmp w"))) break;
    if ((rc = sort_by(g, m, n, g->e_dst, td, cols2(g->e_src, ts, g->e_w, tw)))) break;
    if ((rc = sort_by(g, m, n, ts, g->el_src, cols2(td, g->el_dst, tw, g->el_w)))) break;
    rc = degrees_from_sorted(g, g->el_src, g->out_deg);
  } while (0);
  (void)hipStreamSynchronize(g->stream);
  eg_dfree_t(&ts);
  eg_dfree_t(&td);
  eg_dfree_t(&tw);
  return rc;
}

// Common tail of every ingest path: records are in g->e_* (input order, with
// reversed copies when undirected); validate ids, build the mode's layout and
// its degrees (counted once: R4 not reproduced), release the records.
int eg_build_layout(eg_graph *g) {
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  // Events and the id flag are released on every exit.
  struct Guard {
    hipEvent_t e[2] = {nullptr, nullptr};
    uint32_t *tmp = nullptr;
    ~Guard() {
      for (hipEvent_t ev : e)
        if (ev) (void)hipEventDestroy(ev);
      eg_dfree_t(&tmp);
    }
  } G;
  EG_HIP(hipEventCreate(&G.e[0]));
  EG_HIP(hipEventCreate(&G.e[1]));
  EG_HIP(hipEventRecord(G.e[0], g->stream));
  int bad = 0;
  EG_TRY(eg_dalloc_t(&G.tmp, 1025, "id flag"));
  int rc = eg_check_ids(g->stream, g->e_src, g->e_dst, m, n, G.tmp, &bad);
  if (rc == EG_OK && bad) rc = eg_fail(EG_ERANGE, "edge file holds a vertex id >= n = %u", n);
  if (rc == EG_OK) rc = eg_dalloc_t(&g->out_deg, n, "out degree");
  if (rc == EG_OK) rc = eg_dalloc_t(&g->in_deg, n, "in degree");
  if (rc == EG_OK) {
    switch (g->load_mode) {
      case EG_MODE_ADJ_SORT:
      case EG_MODE_ADJ_SORT_FULL:
      case EG_MODE_ADJ_COUNT:
        rc = build_adjacency(g);
        break;
      case EG_MODE_GRID_SORT:
      case EG_MODE_GRID_NOSORT:
      case EG_MODE_GRID_SORT_SRC:
        rc = build_grid(g);
        break;
      default:
        rc = build_edgelist(g);
        break;
    }
  }
  if (rc == EG_OK) rc = eg_count_zero(g->stream, g->out_deg, n, G.tmp, &g->ndangling);
  if (rc) {
    (void)hipStreamSynchronize(g->stream);  // nothing in flight reads what the guard frees
    return rc;
  }
  EG_HIP(hipEventRecord(G.e[1], g->stream));
  EG_HIP(hipEventSynchronize(G.e[1]));
  float ms = 0.f;
  EG_HIP(hipEventElapsedTime(&ms, G.e[0], G.e[1]));
  g->build_ms = ms;
  // The records (memblock) are released once the layout exists, as
  // init_adj_sort / init_adj_count do (init_all.c:494,787).
  eg_dfree_t(&g->e_src);
  eg_dfree_t(&g->e_dst);
  eg_dfree_t(&g->e_w);
  g->loaded = true;
  return EG_OK;
}

// ---------------------------------------------------------------------------
// CSR build that allocates nothing (eg_csr_build_noalloc): the stable sort by
// src of init_adj_sort (init_all.c:629-697) over records already in device
// memory, with every temporary (sorted keys, radix ping-pong buffers, digit
// count matrix, scan partials) in one fixed device array.  Each device holds
// its own copy of the array at its own address, found once per device with
// hipGetSymbolAddress; the array's name is never passed to a kernel.  Calls
// that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocScratchBytes = 48ull << 20;
constexpr int kNoAllocMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// Sorted keys first, then the radix sort's workspace.
size_t noalloc_need(uint64_t m) { return round256(m * 4) + eg_radix_tmp_bytes(m, 1); }

// Device address of this device's scratch array (looked up on first use).
int noalloc_scratch(void **out) {
  int dev = 0;
  EG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return eg_fail(EG_EDEVICE, "device %d beyond the %d supported", dev, kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    if (noalloc_need(EG_NOALLOC_MAX_ENTRIES) > kNoAllocScratchBytes)
      return eg_fail(EG_EDEVICE, "scratch array smaller than the workspace of %u entries",
                     EG_NOALLOC_MAX_ENTRIES);
    void *p = nullptr;
    EG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return eg_fail(EG_EDEVICE, "no device address for the scratch array");
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return EG_OK;
}

}  // namespace

extern "C" int eg_csr_build_noalloc(uint32_t n, uint64_t m, const uint32_t *src,
                                    const uint32_t *dst, uint32_t *row_ptr, uint32_t *col,
                                    void *stream) {
  if (m > EG_NOALLOC_MAX_ENTRIES)
    return eg_fail(EG_EINVAL, "%llu entries exceed the %u that the fixed scratch holds",
                   (unsigned long long)m, EG_NOALLOC_MAX_ENTRIES);
  if (n == 0 || !row_ptr || (m && (!src || !dst || !col)))
    return eg_fail(EG_EINVAL, "bad CSR build arguments");
  void *base = nullptr;
  EG_TRY(noalloc_scratch(&base));
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint8_t *t = static_cast<uint8_t *>(base);
  uint32_t *sorted = reinterpret_cast<uint32_t *>(t);
  EgSortCols c;
  c.np = 1;
  c.in[0] = dst;
  c.out[0] = col;
  // every buffer of the sort is written before it is read: nothing to clear
  EG_TRY(eg_radix_sort(s, m, eg_bits_for(n), src, sorted, c, t + round256(m * 4)));
  return eg_ptr_from_sorted(s, sorted, m, n, row_ptr, nullptr);
}
