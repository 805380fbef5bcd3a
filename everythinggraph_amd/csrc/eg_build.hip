// eg_build.hip -- layout builders of libeg (the load modes of init(),
// init_all.c:855-888).
//
// Every layout is a stable sort of the records (input order, reversed copies
// appended with -u) by the layout's key, run on the device with the LSD radix
// sort of eg_prims.hip.  Row / cell offsets and degrees are the run offsets of
// the sorted keys (one binary search per vertex), the contract of the
// reference's boundary detection + scanIBack(min) (init_all.c:675-697): no
// degree is accumulated with atomics, so R-MAT hubs cost nothing extra.
#include <functional>
#include <mutex>

#include "eg_internal.h"

static int sort_by(eg_graph *g, uint64_t m, uint64_t nkeys, const uint32_t *kin, uint32_t *kout,
                   EgSortCols cols, const EgIdCheck *chk = nullptr) {
  void *tmp = nullptr;
  EG_TRY(eg_scratch(g, eg_radix_tmp_bytes(m, cols.np), &tmp));
  return eg_radix_sort(g->stream, m, eg_bits_for(nkeys), kin, kout, cols, tmp, chk);
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
// The sort by src also checks every id against n (chk).
static int build_adjacency(eg_graph *g, const EgIdCheck *chk) {
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  const bool w = (g->flags & EG_WEIGHTED) != 0;
  EG_TRY(eg_dalloc_t(&g->out_ptr, (uint64_t)n + 1, "row_ptr"));
  EG_TRY(eg_dalloc_t(&g->out_col, m, "col"));
  EG_TRY(eg_dalloc_t(&g->out_src, m, "row ids"));
  if (w) EG_TRY(eg_dalloc_t(&g->out_w, m, "weights"));
  EG_TRY(sort_by(g, m, n, g->e_src, g->out_src, cols2(g->e_dst, g->out_col, g->e_w, g->out_w),
                 chk));
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

// ---------------------------------------------------------------------------
// Adjacency with 64-bit row offsets (graphs with 2^32 stored entries or more).
// The records are cut, in input order, into parts of at most 2^31 entries;
// each part is stably sorted by its key with the 32-bit radix sort and gives
// per-row counts (run offsets).  Row offsets are the 64-bit exclusive scan of
// the summed counts, and part p's entries of row v are placed after those of
// the parts before it: the result is one stable sort of all records by key,
// the order of init_adj_sort (init_all.c:629-697) at any size.
// ---------------------------------------------------------------------------
namespace {

// ptr[v] = entries of row v over all parts, deg[v] the same in 32 bits
// (flag set when a row holds 2^32 entries or more).
__global__ void k_wide_counts(uint32_t n, uint32_t nparts, const uint32_t *__restrict__ pp,
                              uint64_t *__restrict__ ptr, uint32_t *__restrict__ deg,
                              uint32_t *__restrict__ flag) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    uint64_t d = 0;
    for (uint32_t p = 0; p < nparts; p++) {
      const uint32_t *q = pp + (uint64_t)p * ((uint64_t)n + 1);
      d += q[v + 1] - q[v];
    }
    ptr[v] = d;
    if (d > 0xFFFFFFFFull) {
      atomicOr(flag, 1u);
      d = 0xFFFFFFFFull;
    }
    deg[v] = (uint32_t)d;
  }
}

// Places one sorted part: its j-th entry (key v) goes to base[v] + j - pp[v],
// base[v] = row offset of v + entries of v in the parts before.
__global__ void k_wide_place(uint64_t len, const uint32_t *__restrict__ sk,
                             const uint32_t *__restrict__ sv, const float *__restrict__ sw,
                             const uint32_t *__restrict__ pp, const uint64_t *__restrict__ base,
                             uint32_t *__restrict__ idx, float *__restrict__ w,
                             uint32_t *__restrict__ rows) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += stride) {
    const uint32_t v = sk[j];
    const uint64_t d = base[v] + (j - pp[v]);
    idx[d] = sv[j];
    if (w) w[d] = sw[j];
    if (rows) rows[d] = v;
  }
}

__global__ void k_wide_advance(uint32_t n, const uint32_t *__restrict__ pp,
                               uint64_t *__restrict__ base) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
    base[v] += pp[v + 1] - pp[v];
}

uint32_t wide_grid(uint64_t count) {
  const uint64_t b = (count + 255) / 256;
  return b < 65536 ? (b ? (uint32_t)b : 1u) : 65536u;
}

}  // namespace

// One orientation: rows keyed by keys[i], entries vals[i] (weights win, may be
// null).  release() is called once the parts are sorted (the inputs are not
// read after it).  Allocates *ptr (n + 1), *idx (m), *w (when win) and, when
// rows is given, *rows (m): the row of every entry.
static int wide_orientation(eg_graph *g, uint64_t part, const uint32_t *keys,
                            const uint32_t *vals, const float *win,
                            const std::function<void()> &release, uint64_t **ptr, uint32_t **idx,
                            float **w, uint32_t *deg, uint32_t **rows) {
  const uint64_t m = g->m;
  const uint32_t n = g->n;
  hipStream_t s = g->stream;
  const uint64_t nparts = m ? (m + part - 1) / part : 0;
  if (nparts > 0xFFFFFFFFull) return eg_fail(EG_EINVAL, "too many sorted parts");
  struct Bufs {
    uint32_t *sk = nullptr, *sv = nullptr, *pp = nullptr, *flag = nullptr;
    float *sw = nullptr;
    uint64_t *base = nullptr, *scan = nullptr;
    ~Bufs() {
      eg_dfree_t(&sk);
      eg_dfree_t(&sv);
      eg_dfree_t(&pp);
      eg_dfree_t(&flag);
      eg_dfree_t(&sw);
      eg_dfree_t(&base);
      eg_dfree_t(&scan);
    }
  } B;
  int rc = EG_OK;
  const uint64_t stride = (uint64_t)n + 1;
  if (m) {
    EG_TRY(eg_dalloc_t(&B.sk, m, "sorted part keys"));
    EG_TRY(eg_dalloc_t(&B.sv, m, "sorted part entries"));
    if (win) EG_TRY(eg_dalloc_t(&B.sw, m, "sorted part weights"));
    EG_TRY(eg_dalloc_t(&B.pp, nparts * stride, "part run offsets"));
    for (uint64_t p = 0; p < nparts && rc == EG_OK; p++) {
      const uint64_t off = p * part, len = m - off < part ? m - off : part;
      rc = sort_by(g, len, n, keys + off, B.sk + off,
                   cols2(vals + off, B.sv + off, win ? win + off : nullptr,
                         win ? B.sw + off : nullptr));
      if (rc == EG_OK)
        rc = eg_ptr_from_sorted(s, B.sk + off, len, n, B.pp + p * stride, nullptr);
    }
    EG_HIP(hipStreamSynchronize(s));  // nothing in flight reads the inputs or the sort buffers
    if (rc) return rc;
    // the sort's workspace (the largest buffer of the build) goes back now
    eg_dfree(&g->scratch);
    g->scratch_bytes = 0;
  }
  release();
  EG_TRY(eg_dalloc_t(ptr, stride, "64-bit row offsets"));
  EG_TRY(eg_dalloc_t(idx, m ? m : 1, "entries"));
  if (win) EG_TRY(eg_dalloc_t(w, m ? m : 1, "entry weights"));
  if (rows) EG_TRY(eg_dalloc_t(rows, m ? m : 1, "entry rows"));
  EG_TRY(eg_dalloc_t(&B.flag, 1, "degree flag"));
  EG_TRY(eg_dalloc(reinterpret_cast<void **>(&B.scan), eg_scan64_tmp_bytes(), "scan partials"));
  EG_HIP(hipMemsetAsync(B.flag, 0, 4, s));
  hipLaunchKernelGGL(k_wide_counts, dim3(wide_grid(n)), dim3(256), 0, s, n, (uint32_t)nparts,
                     B.pp, *ptr, deg, B.flag);
  EG_LAUNCHED();
  EG_TRY(eg_scan_excl_u64(s, *ptr, n, B.scan));
  uint32_t bad = 0;
  EG_HIP(hipMemcpyAsync(&bad, B.flag, 4, hipMemcpyDeviceToHost, s));
  EG_HIP(hipStreamSynchronize(s));
  if (bad) return eg_fail(EG_ENOTSUP, "a vertex holds 2^32 entries or more (32-bit degrees)");
  if (!m) return EG_OK;
  EG_TRY(eg_dalloc_t(&B.base, n, "placement offsets"));
  EG_HIP(hipMemcpyAsync(B.base, *ptr, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  for (uint64_t p = 0; p < nparts; p++) {
    const uint64_t off = p * part, len = m - off < part ? m - off : part;
    const uint32_t *pp = B.pp + p * stride;
    hipLaunchKernelGGL(k_wide_place, dim3(wide_grid(len)), dim3(256), 0, s, len, B.sk + off,
                       B.sv + off, win ? B.sw + off : nullptr, pp,
                       static_cast<const uint64_t *>(B.base), *idx, win ? *w : nullptr,
                       rows ? *rows : nullptr);
    EG_LAUNCHED();
    if (p + 1 < nparts) {
      hipLaunchKernelGGL(k_wide_advance, dim3(wide_grid(n)), dim3(256), 0, s, n, pp, B.base);
      EG_LAUNCHED();
    }
  }
  EG_HIP(hipStreamSynchronize(s));  // before the buffers go
  return EG_OK;
}

// The CSR (and, unless -s, the CSC) of an adjacency mode with 64-bit row
// offsets; the CSC is the stable sort of the CSR entries by column with the
// entry's row as payload (init_all.c:714-767).
static int build_adjacency_wide(eg_graph *g, uint64_t part) {
  const uint32_t n = g->n;
  const bool w = (g->flags & EG_WEIGHTED) != 0;
  g->wide = true;
  const bool csc = !(g->flags & EG_SYMMETRIC);
  uint32_t *rows = nullptr;
  int rc = wide_orientation(
      g, part, g->e_src, g->e_dst, g->e_w,
      [g] {
        eg_dfree_t(&g->e_src);
        eg_dfree_t(&g->e_dst);
        eg_dfree_t(&g->e_w);
      },
      &g->out_ptr64, &g->out_col, w ? &g->out_w : nullptr, g->out_deg, csc ? &rows : nullptr);
  if (rc == EG_OK && !csc)
    rc = hipMemcpyAsync(g->in_deg, g->out_deg, (size_t)n * 4, hipMemcpyDeviceToDevice,
                        g->stream) == hipSuccess
             ? EG_OK
             : eg_fail(EG_EDEVICE, "degree copy failed");
  if (rc == EG_OK && csc)
    rc = wide_orientation(
        g, part, g->out_col, rows, g->out_w, [&rows] { eg_dfree_t(&rows); }, &g->in_ptr64,
        &g->in_idx, w ? &g->in_w : nullptr, g->in_deg, nullptr);
  (void)hipStreamSynchronize(g->stream);
  eg_dfree_t(&rows);
  return rc;
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
  const bool adjacency = g->load_mode == EG_MODE_ADJ_SORT ||
                         g->load_mode == EG_MODE_ADJ_SORT_FULL ||
                         g->load_mode == EG_MODE_ADJ_COUNT;
  const bool wide = adjacency && (g->tune.wide_part > 0 || m >= (1ull << 32));
  // the 32-bit adjacency build checks the ids inside its first sort
  const bool fused_check = adjacency && !wide;
  int rc = EG_OK;
  if (fused_check) {
    EG_HIP(hipMemsetAsync(G.tmp, 0, sizeof(uint32_t), g->stream));
  } else {
    rc = eg_check_ids(g->stream, g->e_src, g->e_dst, m, n, G.tmp, &bad);
    if (rc == EG_OK && bad) rc = eg_fail(EG_ERANGE, "edge file holds a vertex id >= n = %u", n);
  }
  if (rc == EG_OK) rc = eg_dalloc_t(&g->out_deg, n, "out degree");
  if (rc == EG_OK) rc = eg_dalloc_t(&g->in_deg, n, "in degree");
  if (rc == EG_OK) {
    switch (g->load_mode) {
      case EG_MODE_ADJ_SORT:
      case EG_MODE_ADJ_SORT_FULL:
      case EG_MODE_ADJ_COUNT:
        if (wide) {
          rc = build_adjacency_wide(g, 1ull << (g->tune.wide_part > 0 ? g->tune.wide_part : 31));
        } else {
          EgIdCheck chk;
          chk.limit = n;
          chk.vals = g->e_dst;
          chk.flag = G.tmp;
          rc = build_adjacency(g, &chk);
          if (rc == EG_OK) {
            uint32_t h = 0;
            EG_HIP(hipMemcpyAsync(&h, G.tmp, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
            EG_HIP(hipStreamSynchronize(g->stream));
            if (h) rc = eg_fail(EG_ERANGE, "edge file holds a vertex id >= n = %u", n);
          }
        }
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
  // so is the sorts' workspace (later passes take what they need on demand)
  eg_dfree(&g->scratch);
  g->scratch_bytes = 0;
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
