// eg_core.hip -- graph lifecycle, ingest and layout builders of libeg.
//
// Replaces the reference's init() (init_all.c:791-893): records are read from
// the binary edge file (init_all.c:812-825), copied to the device, unpacked
// with the reversed copies of -u appended (preload_graph, init_all.c:81-110),
// and the layout of the load mode is built with device histograms, scans and
// the stable radix sort of eg_prims.hip (init_adj_sort, init_grid_sort,
// init_grid_nosort, init_adj_count, init_edgelist).
#include <errno.h>
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "eg_internal.h"

// ---------------------------------------------------------------------------
// Errors.
// ---------------------------------------------------------------------------
static thread_local char t_err[512] = "";

int eg_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char *eg_last_error(void) { return t_err; }
extern "C" const char *eg_version(void) { return "libeg 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// Memory.
// ---------------------------------------------------------------------------
int eg_dalloc(void **p, size_t bytes, const char *what) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return eg_fail(e == hipErrorOutOfMemory ? EG_ENOMEM : EG_EDEVICE,
                   "hipMalloc(%zu) for %s: %s", bytes, what, hipGetErrorString(e));
  }
  return EG_OK;
}

void eg_dfree(void **p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

int eg_scratch(eg_graph *g, size_t bytes, void **out) {
  if (bytes > g->scratch_bytes) {
    EG_HIP(hipStreamSynchronize(g->stream));
    eg_dfree(&g->scratch);
    g->scratch_bytes = 0;
    EG_TRY(eg_dalloc(&g->scratch, bytes, "scratch"));
    g->scratch_bytes = bytes;
  }
  *out = g->scratch;
  return EG_OK;
}

bool eg_has_in_edges(const eg_graph *g) { return g->in_ptr != nullptr; }

double eg_wall_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int eg_require_loaded_wide_ok(const eg_graph *g) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  if (!g->loaded) return eg_fail(EG_ESTATE, "graph has no edges loaded");
  if (g->server) return eg_fail(EG_ESTATE, "graph is owned by a running server");
  return EG_OK;
}

int eg_require_loaded(const eg_graph *g) {
  EG_TRY(eg_require_loaded_wide_ok(g));
  if (g->wide)
    return eg_fail(EG_ENOTSUP,
                   "the layout has 64-bit row offsets (%llu stored entries): BFS, degrees "
                   "and the CSR / CSC downloads run on it",
                   (unsigned long long)g->m);
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Timing.
// ---------------------------------------------------------------------------
EgProf::EgProf(eg_graph *g_, const char *name, hipStream_t s_) : g(g_) {
  if (!g || !g->prof_on) return;
  stream = s_ ? s_ : g->stream;
  for (auto &s : g->prof)
    if (s.name == name) slot = &s;
  if (!slot) {
    g->prof.emplace_back();
    slot = &g->prof.back();
    slot->name = name;
  }
  if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
    slot = nullptr;
    return;
  }
  (void)hipEventRecord(ev[0], stream);
}

EgProf::~EgProf() {
  if (!slot) return;
  (void)hipEventRecord(ev[1], stream);
  slot->pending.push_back(ev[0]);
  slot->pending.push_back(ev[1]);
}

static void prof_resolve(eg_graph *g) {
  for (auto &s : g->prof) {
    for (size_t i = 0; i + 1 < s.pending.size(); i += 2) {
      float ms = 0.f;
      if (hipEventSynchronize(s.pending[i + 1]) == hipSuccess &&
          hipEventElapsedTime(&ms, s.pending[i], s.pending[i + 1]) == hipSuccess) {
        s.total_ms += ms;
        s.launches++;
      }
      (void)hipEventDestroy(s.pending[i]);
      (void)hipEventDestroy(s.pending[i + 1]);
    }
    s.pending.clear();
  }
}

extern "C" int eg_prof_enable(eg_graph *g, int32_t on) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  g->prof_on = on;
  return EG_OK;
}

extern "C" int eg_prof_reset(eg_graph *g) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  EG_HIP(hipStreamSynchronize(g->stream));
  prof_resolve(g);
  g->prof.clear();
  return EG_OK;
}

extern "C" int eg_prof_read(eg_graph *g, eg_kernel_time *out, int32_t max, int32_t *count) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  EG_HIP(hipStreamSynchronize(g->stream));
  prof_resolve(g);
  int32_t k = 0;
  for (auto &s : g->prof) {
    if (k < max && out) {
      memset(&out[k], 0, sizeof(out[k]));
      strncpy(out[k].name, s.name.c_str(), sizeof(out[k].name) - 1);
      out[k].launches = s.launches;
      out[k].total_ms = s.total_ms;
    }
    k++;
  }
  if (count) *count = k;
  return EG_OK;
}

extern "C" int eg_sync(eg_graph *g) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  EG_HIP(hipStreamSynchronize(g->stream));
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Device.
// ---------------------------------------------------------------------------
extern "C" int eg_device_query(eg_device_info *out) {
  if (!out) return eg_fail(EG_EINVAL, "null output");
  memset(out, 0, sizeof(*out));
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return eg_fail(EG_EDEVICE, "no HIP device: %s", hipGetErrorString(e));
  }
  int dev = 0;
  EG_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  EG_HIP(hipGetDeviceProperties(&prop, dev));
  strncpy(out->name, prop.name, sizeof(out->name) - 1);
  out->compute_units = prop.multiProcessorCount;
  int arch = 0;
  if (sscanf(prop.gcnArchName, "gfx%d", &arch) != 1) arch = 0;
  out->gcn_arch = arch;
  out->hbm_bytes = prop.totalGlobalMem;
  size_t fr = 0, tot = 0;
  EG_HIP(hipMemGetInfo(&fr, &tot));
  out->free_bytes = fr;
  out->peak_hbm_gbs = 8000.0;  // MI355X: 8 TB/s HBM3E (published)
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Graph lifecycle.
// ---------------------------------------------------------------------------
static uint32_t default_P(uint32_t n) {
  // random.c:69-71, in 64-bit arithmetic (the reference's u32 n*8 overflows).
  uint64_t P = ((uint64_t)n * 8 / 1024) / 20;
  if (P > 2000) P = 256;
  if (P == 0) P = 4;
  return (uint32_t)P;
}

extern "C" int eg_graph_create(uint32_t n, uint32_t flags, int32_t load_mode, uint32_t P,
                               eg_graph **out) {
  if (!out) return eg_fail(EG_EINVAL, "null output");
  *out = nullptr;
  if (n == 0) return eg_fail(EG_EINVAL, "n must be > 0");
  if (load_mode < 0 || load_mode > 9)
    return eg_fail(EG_EINVAL, "load mode %d outside 0..9", load_mode);
  if (load_mode == EG_MODE_ADJ_DYNAMIC)
    return eg_fail(EG_ENOTSUP,
                   "load mode 5 (init_adj_dynamic) frees what it builds and leaves no layout");
  if (flags & ~(EG_UNDIRECTED | EG_SYMMETRIC | EG_WEIGHTED))
    return eg_fail(EG_EINVAL, "unknown flags 0x%x", flags);
  if (P == 0) P = default_P(n);
  if ((uint64_t)P * P >= (1ull << 31)) return eg_fail(EG_EINVAL, "grid P=%u too large", P);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return eg_fail(EG_EDEVICE, "no HIP device available: %s", hipGetErrorString(e));
  }
  eg_graph *g = new eg_graph();
  g->n = n;
  g->flags = flags;
  g->load_mode = load_mode;
  g->P = P;
  if (hipGetDevice(&g->device) != hipSuccess ||
      hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    delete g;
    return eg_fail(EG_EDEVICE, "stream creation failed");
  }
  if (hipHostMalloc(reinterpret_cast<void **>(&g->h_ctr), 64 * sizeof(uint32_t),
                    hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(g->stream);
    delete g;
    return eg_fail(EG_ENOMEM, "pinned counters");
  }
  *out = g;
  return EG_OK;
}

static void free_layout(eg_graph *g) {
  eg_dfree_t(&g->e_src);
  eg_dfree_t(&g->e_dst);
  eg_dfree_t(&g->e_w);
  eg_dfree_t(&g->out_ptr);
  eg_dfree_t(&g->out_col);
  eg_dfree_t(&g->out_src);
  eg_dfree_t(&g->bfs_col);
  eg_dfree_t(&g->bfs_ocol);
  eg_dfree_t(&g->out_w);
  eg_dfree_t(&g->in_ptr);
  eg_dfree_t(&g->out_ptr64);
  eg_dfree_t(&g->in_ptr64);
  g->wide = false;
  eg_dfree_t(&g->in_idx);
  eg_dfree_t(&g->in_w);
  eg_dfree_t(&g->out_deg);
  eg_dfree_t(&g->in_deg);
  eg_dfree_t(&g->cell_ptr);
  eg_dfree_t(&g->g_src);
  eg_dfree_t(&g->g_dst);
  eg_dfree_t(&g->g_w);
  eg_dfree_t(&g->el_src);
  eg_dfree_t(&g->el_dst);
  eg_dfree_t(&g->el_w);
  eg_plan_free(&g->plan_in);
  eg_plan_free(&g->plan_out);
  eg_grid_free(g);
  eg_order_free(g);
  eg_pb_free(g);
  eg_batch_free(g);
  g->loaded = false;
  g->ndangling = 0xFFFFFFFFu;
  g->prep_order_ms = g->prep_plan_ms = g->prep_bfs_ms = 0.0;
  g->prep_order_bytes = g->prep_plan_bytes = g->prep_bfs_bytes = 0;
}

static void free_algo_state(eg_graph *g) {
  eg_dfree_t(&g->rank_a);
  eg_dfree_t(&g->rank_b);
  eg_dfree_t(&g->contrib_a);
  eg_dfree_t(&g->contrib_b);
  eg_dfree_t(&g->pr_stats);
  g->pr_stats_cap = 0;
  eg_dfree_t(&g->bfs_dist);
  eg_dfree_t(&g->bfs_parent);
  eg_dfree_t(&g->bfs_visited);
  eg_dfree_t(&g->bfs_front);
  eg_dfree_t(&g->bfs_next);
  eg_dfree_t(&g->bfs_q);
  eg_dfree_t(&g->bfs_nq);
  eg_dfree_t(&g->bfs_f3);
  eg_dfree_t(&g->bfs_qoff);
  eg_dfree_t(&g->bfs_nqoff);
  eg_dfree_t(&g->bfs_part);
  eg_dfree_t(&g->bfs_bmcnt);
  eg_dfree_t(&g->bfs_ctr);
  eg_dfree_t(&g->wcc_label);
  eg_dfree_t(&g->wcc_lab32);
  eg_dfree_t(&g->wcc_tmp);
  eg_dfree_t(&g->ppr_base);
  eg_dfree_t(&g->ppr_seeds);
  g->ppr_seeds_cap = 0;
  eg_dfree_t(&g->sssp_segmin);
  g->sssp_segmin_cap = 0;
  eg_dfree_t(&g->wide_seg);
  eg_dfree_t(&g->wide_ctr);
}

extern "C" int eg_graph_free(eg_graph *g) {
  if (!g) return EG_OK;
  if (g->server) return eg_fail(EG_ESTATE, "graph is owned by a running server; stop it first");
  (void)hipStreamSynchronize(g->stream);
  prof_resolve(g);
  free_layout(g);
  free_algo_state(g);
  eg_dfree(&g->scratch);
  if (g->h_ctr) (void)hipHostFree(g->h_ctr);
  (void)hipStreamDestroy(g->stream);
  delete g;
  return EG_OK;
}

extern "C" int eg_graph_get_info(const eg_graph *g, eg_graph_info *out) {
  if (!g || !out) return eg_fail(EG_EINVAL, "null argument");
  memset(out, 0, sizeof(*out));
  out->n = g->n;
  out->P = g->P;
  out->flags = g->flags;
  out->load_mode = g->load_mode;
  out->m_in = g->m_in;
  out->m = g->m;
  out->has_csr = g->out_ptr != nullptr || g->out_ptr64 != nullptr;
  out->has_csc = g->in_ptr != nullptr || g->in_ptr64 != nullptr;
  out->has_grid = g->cell_ptr != nullptr;
  out->has_edgelist = g->el_src != nullptr;
  out->build_ms = g->build_ms;
  uint64_t b = 0;
  const uint64_t m = g->m, n = g->n;
  if (g->out_ptr) b += 4 * (n + 1) + 8 * m + (g->out_w ? 4 * m : 0);
  if (g->in_ptr) b += 4 * (n + 1) + 4 * m + (g->in_w ? 4 * m : 0);
  if (g->out_ptr64) b += 8 * (n + 1) + 4 * m + (g->out_w ? 4 * m : 0);
  if (g->in_ptr64) b += 8 * (n + 1) + 4 * m + (g->in_w ? 4 * m : 0);
  if (g->cell_ptr) b += 4 * ((uint64_t)g->P * g->P + 1) + 8 * m + (g->g_w ? 4 * m : 0);
  if (g->el_src) b += 8 * m + (g->el_w ? 4 * m : 0);
  b += 8 * n;
  out->device_bytes = b;
  return EG_OK;
}

extern "C" int eg_graph_prep_info(const eg_graph *g, eg_prep_info *out) {
  if (!g || !out) return eg_fail(EG_EINVAL, "null argument");
  eg_graph_info gi;
  EG_TRY(eg_graph_get_info(g, &gi));
  memset(out, 0, sizeof(*out));
  out->layout_ms = g->build_ms;
  out->order_ms = g->prep_order_ms;
  out->binned_plan_ms = g->prep_plan_ms;
  out->bfs_rows_ms = g->prep_bfs_ms;
  out->layout_bytes = gi.device_bytes;
  out->order_bytes = g->prep_order_bytes;
  out->binned_plan_bytes = g->prep_plan_bytes;
  out->bfs_rows_bytes = g->prep_bfs_bytes;
  return EG_OK;
}

int eg_build_layout(eg_graph *g);

static int alloc_records(eg_graph *g, uint64_t m_in) {
  if (g->server) return eg_fail(EG_ESTATE, "graph is owned by a running server");
  free_layout(g);
  const bool undir = (g->flags & EG_UNDIRECTED) != 0;
  const uint64_t m = undir ? 2 * m_in : m_in;
  const bool adjacency = g->load_mode == EG_MODE_ADJ_SORT ||
                         g->load_mode == EG_MODE_ADJ_SORT_FULL ||
                         g->load_mode == EG_MODE_ADJ_COUNT;
  // only the adjacency modes keep 64-bit row offsets
  if (m >= (1ull << 32) && !adjacency)
    return eg_fail(EG_ENOTSUP,
                   "%llu stored entries exceed the 32-bit offsets of load mode %d "
                   "(modes 2/3/4 take 64-bit row offsets)",
                   (unsigned long long)m, g->load_mode);
  g->m_in = m_in;
  g->m = m;
  EG_TRY(eg_dalloc_t(&g->e_src, m, "records src"));
  EG_TRY(eg_dalloc_t(&g->e_dst, m, "records dst"));
  if (g->flags & EG_WEIGHTED) EG_TRY(eg_dalloc_t(&g->e_w, m, "records w"));
  return EG_OK;
}

static int finish_records(eg_graph *g) {
  if (g->flags & EG_UNDIRECTED) EG_TRY(eg_append_reverse(g->stream, g->m_in, g->e_src, g->e_dst, g->e_w));
  return eg_build_layout(g);
}

// ---------------------------------------------------------------------------
// Record ingest: a two-slot pipeline.  Chunk c is copied from its source (the
// caller's memory or the file, by several host threads) into pinned buffer
// c % 2, sent to device staging c % 2 and unpacked into the record arrays;
// while the device works on chunk c the host fills the other slot.  A slot is
// refilled only after the event that follows its unpack has completed.
// ---------------------------------------------------------------------------
namespace {

struct Source {
  const uint8_t *mem = nullptr;  // or
  int fd = -1;
};

// Copies [off, off + len) of the source into dst with up to `threads` threads.
int read_range(const Source &src, uint64_t off, uint64_t len, uint8_t *dst, unsigned threads) {
  const uint64_t per = (len + threads - 1) / threads;
  std::vector<std::thread> pool;
  std::vector<int> err(threads, 0);
  for (unsigned t = 0; t < threads; t++) {
    const uint64_t a = std::min<uint64_t>(len, t * per), b = std::min<uint64_t>(len, a + per);
    if (a >= b) break;
    pool.emplace_back([&, t, a, b] {
      if (src.mem) {
        memcpy(dst + a, src.mem + off + a, b - a);
        return;
      }
      uint64_t done = a;
      while (done < b) {
        const ssize_t r = pread(src.fd, dst + done, b - done, (off_t)(off + done));
        if (r <= 0) {
          err[t] = r < 0 ? errno : EIO;
          return;
        }
        done += (uint64_t)r;
      }
    });
  }
  for (auto &th : pool) th.join();
  for (unsigned t = 0; t < threads; t++)
    if (err[t]) return eg_fail(EG_EIO, "read failed: %s", strerror(err[t]));
  return EG_OK;
}

}  // namespace

static int ingest(eg_graph *g, const Source &src, uint64_t m_in, uint64_t isz) {
  hipStream_t s = g->stream;
  const uint64_t total = m_in * isz;
  const uint64_t want = g->tune.ingest_kib > 0 ? (uint64_t)g->tune.ingest_kib << 10 : 64ull << 20;
  const uint64_t chunk = std::max<uint64_t>(isz, want / isz * isz);
  const uint64_t cap = std::min(chunk, total);
  unsigned threads = std::thread::hardware_concurrency();
  threads = std::max(1u, std::min(threads, 8u));
  uint8_t *pin[2] = {nullptr, nullptr};
  void *dev[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  bool used[2] = {false, false};
  int rc = EG_OK;
  for (int b = 0; b < 2 && rc == EG_OK; b++) {
    if (hipHostMalloc(reinterpret_cast<void **>(&pin[b]), cap, hipHostMallocDefault) != hipSuccess)
      rc = eg_fail(EG_ENOMEM, "pinned ingest buffer");
    if (rc == EG_OK) rc = eg_dalloc(&dev[b], cap, "ingest staging");
    if (rc == EG_OK && hipEventCreateWithFlags(&done[b], hipEventDisableTiming) != hipSuccess)
      rc = eg_fail(EG_EDEVICE, "event creation failed");
  }
  uint64_t off = 0;
  for (uint64_t c = 0; rc == EG_OK && off < total; c++, off += chunk) {
    const int b = (int)(c & 1);
    const uint64_t len = std::min(chunk, total - off);
    if (used[b] && hipEventSynchronize(done[b]) != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "ingest copy failed");
      break;
    }
    rc = read_range(src, off, len, pin[b], threads);
    if (rc) break;
    const uint64_t r0 = off / isz;
    hipError_t e = hipMemcpyAsync(dev[b], pin[b], len, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "upload: %s", hipGetErrorString(e));
      break;
    }
    rc = eg_unpack_records(s, dev[b], len / isz, isz == 12, g->e_src + r0, g->e_dst + r0,
                           g->e_w ? g->e_w + r0 : nullptr);
    if (rc == EG_OK && hipEventRecord(done[b], s) != hipSuccess)
      rc = eg_fail(EG_EDEVICE, "event record failed");
    used[b] = true;
  }
  (void)hipStreamSynchronize(s);
  for (int b = 0; b < 2; b++) {
    if (done[b]) (void)hipEventDestroy(done[b]);
    eg_dfree(&dev[b]);
    if (pin[b]) (void)hipHostFree(pin[b]);
  }
  return rc;
}

extern "C" int eg_graph_load_records(eg_graph *g, const void *records, uint64_t nbytes) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  const uint64_t isz = (g->flags & EG_WEIGHTED) ? 12 : 8;
  const uint64_t m_in = nbytes / isz;  // init_all.c:821 drops a partial record
  if (m_in && !records) return eg_fail(EG_EINVAL, "null records");
  EG_TRY(alloc_records(g, m_in));
  if (m_in) {
    Source src;
    src.mem = static_cast<const uint8_t *>(records);
    EG_TRY(ingest(g, src, m_in, isz));
  }
  return finish_records(g);
}

extern "C" int eg_graph_load_file(eg_graph *g, const char *path) {
  if (!g || !path) return eg_fail(EG_EINVAL, "null argument");
  // init_all.c:812-819, with the open failure reported (R12).
  int fd = open(path, O_RDONLY);
  if (fd < 0) return eg_fail(EG_EIO, "cannot open %s: %s", path, strerror(errno));
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return eg_fail(EG_EIO, "cannot stat %s", path);
  }
  const uint64_t isz = (g->flags & EG_WEIGHTED) ? 12 : 8;
  const uint64_t m_in = (uint64_t)sb.st_size / isz;  // init_all.c:821
  int rc = alloc_records(g, m_in);
  if (rc == EG_OK && m_in) {
    Source src;
    src.fd = fd;
    rc = ingest(g, src, m_in, isz);
  }
  close(fd);
  if (rc) return rc;
  return finish_records(g);
}

extern "C" int eg_graph_load_edges(eg_graph *g, uint64_t m_in, const uint32_t *src,
                                   const uint32_t *dst, const float *w) {
  if (!g) return eg_fail(EG_EINVAL, "null graph");
  if (m_in && (!src || !dst)) return eg_fail(EG_EINVAL, "null edge arrays");
  if ((g->flags & EG_WEIGHTED) && m_in && !w)
    return eg_fail(EG_EINVAL, "weighted graph needs weights");
  EG_TRY(alloc_records(g, m_in));
  if (m_in) {
    EG_HIP(hipMemcpyAsync(g->e_src, src, m_in * 4, hipMemcpyHostToDevice, g->stream));
    EG_HIP(hipMemcpyAsync(g->e_dst, dst, m_in * 4, hipMemcpyHostToDevice, g->stream));
    if (g->e_w) EG_HIP(hipMemcpyAsync(g->e_w, w, m_in * 4, hipMemcpyHostToDevice, g->stream));
  }
  return finish_records(g);
}

static int check_gen(const eg_gen_params *p, uint32_t n);

extern "C" int eg_graph_generate(eg_graph *g, const eg_gen_params *p) {
  if (!g || !p) return eg_fail(EG_EINVAL, "null argument");
  EG_TRY(check_gen(p, g->n));
  EG_TRY(alloc_records(g, p->m_in));
  EG_TRY(eg_gen_device(g->stream, p, g->n, g->e_src, g->e_dst, g->e_w));
  return finish_records(g);
}

static int check_gen(const eg_gen_params *p, uint32_t n) {
  if (p->kind == 0) {
    if (p->scale < 1 || p->scale > 32 || ((uint64_t)1 << p->scale) > (uint64_t)n)
      return eg_fail(EG_EINVAL, "R-MAT scale %d does not fit n = %u", p->scale, n);
    if (p->a < 0 || p->b < 0 || p->c < 0 || p->a + p->b + p->c > 1.0)
      return eg_fail(EG_EINVAL, "R-MAT probabilities out of range");
  } else if (p->kind != 1) {
    return eg_fail(EG_EINVAL, "generator kind %d", p->kind);
  }
  return EG_OK;
}

extern "C" int eg_generate_edges(const eg_gen_params *p, uint32_t n, uint32_t *src,
                                 uint32_t *dst, float *w) {
  if (!p || (p->m_in && (!src || !dst))) return eg_fail(EG_EINVAL, "null argument");
  EG_TRY(check_gen(p, n));
  if (p->m_in == 0) return EG_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return eg_fail(EG_EDEVICE, "no HIP device available");
  }
  uint32_t *ds = nullptr, *dd = nullptr;
  float *dw = nullptr;
  int rc = EG_OK;
  hipStream_t s = nullptr;
  do {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "stream creation failed");
      break;
    }
    if ((rc = eg_dalloc_t(&ds, p->m_in, "gen src")) || (rc = eg_dalloc_t(&dd, p->m_in, "gen dst")))
      break;
    if (w && (rc = eg_dalloc_t(&dw, p->m_in, "gen w"))) break;
    if ((rc = eg_gen_device(s, p, n, ds, dd, dw))) break;
    if (hipMemcpyAsync(src, ds, p->m_in * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(dst, dd, p->m_in * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        (w && hipMemcpyAsync(w, dw, p->m_in * 4, hipMemcpyDeviceToHost, s) != hipSuccess) ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = eg_fail(EG_EDEVICE, "generator download failed");
      break;
    }
  } while (0);
  eg_dfree_t(&ds);
  eg_dfree_t(&dd);
  eg_dfree_t(&dw);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

// ---------------------------------------------------------------------------
// Downloads.
// ---------------------------------------------------------------------------
static int download_ptr(eg_graph *g, const uint32_t *d, uint64_t count, uint64_t *h) {
  std::vector<uint32_t> tmp(count);
  EG_HIP(hipMemcpyAsync(tmp.data(), d, count * 4, hipMemcpyDeviceToHost, g->stream));
  EG_HIP(hipStreamSynchronize(g->stream));
  for (uint64_t i = 0; i < count; i++) h[i] = tmp[i];
  return EG_OK;
}

static int download(eg_graph *g, const void *d, void *h, uint64_t bytes) {
  if (!h || !bytes) return EG_OK;
  if (!d) return eg_fail(EG_ESTATE, "array not present in this layout");
  EG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g->stream));
  EG_HIP(hipStreamSynchronize(g->stream));
  return EG_OK;
}

extern "C" int eg_graph_get_degrees(eg_graph *g, uint32_t *out_deg, uint32_t *in_deg) {
  EG_TRY(eg_require_loaded_wide_ok(g));
  EG_TRY(download(g, g->out_deg, out_deg, (uint64_t)g->n * 4));
  EG_TRY(download(g, g->in_deg, in_deg, (uint64_t)g->n * 4));
  return EG_OK;
}

extern "C" int eg_graph_get_csr(eg_graph *g, uint64_t *row_ptr, uint32_t *col, float *w) {
  EG_TRY(eg_require_loaded_wide_ok(g));
  if (g->out_ptr64) {
    EG_TRY(download(g, g->out_ptr64, row_ptr, ((uint64_t)g->n + 1) * 8));
  } else {
    if (!g->out_ptr) return eg_fail(EG_ESTATE, "load mode %d has no CSR", g->load_mode);
    if (row_ptr) EG_TRY(download_ptr(g, g->out_ptr, (uint64_t)g->n + 1, row_ptr));
  }
  EG_TRY(download(g, g->out_col, col, g->m * 4));
  if (w) EG_TRY(download(g, g->out_w, w, g->m * 4));
  return EG_OK;
}

extern "C" int eg_graph_get_csc(eg_graph *g, uint64_t *col_ptr, uint32_t *row, float *w) {
  EG_TRY(eg_require_loaded_wide_ok(g));
  if (g->in_ptr64) {
    EG_TRY(download(g, g->in_ptr64, col_ptr, ((uint64_t)g->n + 1) * 8));
  } else {
    if (!g->in_ptr) return eg_fail(EG_ESTATE, "no in-edges (load mode %d or -s)", g->load_mode);
    if (col_ptr) EG_TRY(download_ptr(g, g->in_ptr, (uint64_t)g->n + 1, col_ptr));
  }
  EG_TRY(download(g, g->in_idx, row, g->m * 4));
  if (w) EG_TRY(download(g, g->in_w, w, g->m * 4));
  return EG_OK;
}

extern "C" int eg_graph_get_grid(eg_graph *g, uint64_t *cell_ptr, uint32_t *src, uint32_t *dst,
                                 float *w) {
  EG_TRY(eg_require_loaded(g));
  if (!g->cell_ptr) return eg_fail(EG_ESTATE, "load mode %d has no grid", g->load_mode);
  if (cell_ptr) EG_TRY(download_ptr(g, g->cell_ptr, (uint64_t)g->P * g->P + 1, cell_ptr));
  EG_TRY(download(g, g->g_src, src, g->m * 4));
  EG_TRY(download(g, g->g_dst, dst, g->m * 4));
  if (w) EG_TRY(download(g, g->g_w, w, g->m * 4));
  return EG_OK;
}

extern "C" int eg_graph_get_edgelist(eg_graph *g, uint32_t *src, uint32_t *dst, float *w) {
  EG_TRY(eg_require_loaded(g));
  if (!g->el_src) return eg_fail(EG_ESTATE, "load mode %d has no edge array", g->load_mode);
  EG_TRY(download(g, g->el_src, src, g->m * 4));
  EG_TRY(download(g, g->el_dst, dst, g->m * 4));
  if (w) EG_TRY(download(g, g->el_w, w, g->m * 4));
  return EG_OK;
}
