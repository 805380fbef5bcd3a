// eg_tune.hip -- kernel-shape autotuner.  No reference counterpart.
//
// Each tunable kernel has variants that compute bit-identical results (the
// shape changes how many memory requests a lane keeps in flight, never the
// order of a row's additions), so the choice is free to depend on timing:
//  * pull_batch  -- k_pull (PageRank / SpMV / single-vector PPR): entries per
//                   lane per gather pass, 8 or 16;
//  * spmm_depth  -- k_ppr_spmm: row gathers in flight per wave, 8 or 16;
//  * ms_batch    -- k_ms_pull: entries per lane per index pass, 8 or 16;
//  * ms_push_div -- multi-source BFS on symmetric graphs: a level pushes while
//                   the union frontier's vertices plus out-degrees are at most
//                   m / ms_push_div (20 = the single-source rule of
//                   bfs_simple.c:191-197; timed: 5, 20, 40);
//  * pull_persist -- k_pull on the degree-ordered copy: 1 or 2 persistent
//                   workgroups per CU that keep the leading (highest-degree)
//                   entries of the gathered vector in LDS; 0 = off;
//  * pull_hot    -- k_pull on the degree-ordered copy: gathers of sources with
//                   new id >= pull_hot bypass normal caching (non-temporal
//                   loads) so that the hottest sources stay in L2; 0 = off;
//  * ingest_kib  -- record ingest chunk (KiB) of the pinned two-slot pipeline
//                   (0 = 64 MiB); not timed by eg_autotune;
//  * graphs      -- 1: batched PPR iterations are replayed from a hipGraph;
//  * grid_band   -- grid pull: source partitions per launch (0 = all), so that
//                   one launch gathers from a band of x that fits in L2;
//  * pr_binned   -- PageRank by the binned two-phase kernels (1) or the pull
//                   kernel (0), timed against the best pull shape; the two sum
//                   in different arithmetic (exact fixed point vs compensated
//                   FP32), so this is the one choice that changes result bits;
//  * pr_tile     -- binned PageRank: source offsets per scatter workgroup
//                   (0 = 32768); not timed by eg_autotune;
//  * pr_scatter  -- binned PageRank scatter: source offsets loaded per lane
//                   before any value is written, 8, 16 or 32;
//  * pr_hot      -- binned PageRank: sources below this id are read from an
//                   LDS copy in phase 2 instead of being scattered (0..389120,
//                   in windows of 24320 sources loaded one at a time; the plan
//                   keeps the leading windows that hold 2% of the entries;
//                   timed: caps of 1, 4, 8 and 16 windows, each distinct plan
//                   once);
//  * pr_persist  -- binned phase 2 with hot sources: persistent workgroups,
//                   one per CU (1), or one workgroup per work item (0); timed
//                   on the plan of the kept pr_hot;
//  * pr_xcd      -- binned scatter tile walk: 0 = XCD x takes the x-th
//                   eighth of the chunk-ordered tiles (the default); 1 = XCD x
//                   takes a short run of tiles from every eighth by turns
//                   (balances XCDs whose eighths differ in cost); the tiles
//                   of one chunk share an XCD either way;
//  * bfs_order   -- symmetric graphs: BFS runs on the degree-ordered ids with
//                   every row hubs first (2, the default; 1 without the
//                   ordered copy), on the layout's ids with rows hubs first
//                   (1), or on the rows as built (0); levels are the same,
//                   parents may differ;
//  * wcc_order   -- undirected / symmetric graphs: WCC pulls over the
//                   degree-ordered ids, rows hubs first (1, the default), or
//                   relaxes the edges in the layout's ids (0); labels are the
//                   same.
// eg_autotune times every variant on the graph itself through the public
// entry points (device time from HIP events, best of a few repetitions) and
// keeps the fastest.
#include <string.h>

#include <algorithm>
#include <vector>

#include "eg_internal.h"

namespace {

int *param_slot(eg_graph *g, const char *name) {
  if (!strcmp(name, "pull_batch")) return &g->tune.pull_batch;
  if (!strcmp(name, "spmm_depth")) return &g->tune.spmm_depth;
  if (!strcmp(name, "ms_batch")) return &g->tune.ms_batch;
  if (!strcmp(name, "pull_hot")) return &g->tune.pull_hot;
  if (!strcmp(name, "pull_persist")) return &g->tune.pull_persist;
  if (!strcmp(name, "graphs")) return &g->tune.graphs;
  if (!strcmp(name, "grid_band")) return &g->tune.grid_band;
  if (!strcmp(name, "ingest_kib")) return &g->tune.ingest_kib;
  if (!strcmp(name, "pr_binned")) return &g->tune.pr_binned;
  if (!strcmp(name, "pr_tile")) return &g->tune.pr_tile;
  if (!strcmp(name, "pr_scatter")) return &g->tune.pr_scatter;
  if (!strcmp(name, "pr_hot")) return &g->tune.pr_hot;
  if (!strcmp(name, "pr_persist")) return &g->tune.pr_persist;
  if (!strcmp(name, "pr_xcd")) return &g->tune.pr_xcd;
  if (!strcmp(name, "ms_lanes")) return &g->tune.ms_lanes;
  if (!strcmp(name, "ms_device")) return &g->tune.ms_device;
  if (!strcmp(name, "ms_push_div")) return &g->tune.ms_push_div;
  if (!strcmp(name, "bfs_order")) return &g->tune.bfs_order;
  if (!strcmp(name, "wcc_order")) return &g->tune.wcc_order;
  if (!strcmp(name, "wide_part")) return &g->tune.wide_part;
  return nullptr;
}

// Device milliseconds of fn() on the graph's stream, best of reps.
template <typename F>
int time_best(eg_graph *g, int reps, F fn, double *best) {
  hipEvent_t e0, e1;
  EG_HIP(hipEventCreate(&e0));
  EG_HIP(hipEventCreate(&e1));
  *best = 1e300;
  int rc = EG_OK;
  for (int r = 0; r < reps && rc == EG_OK; r++) {
    (void)hipEventRecord(e0, g->stream);
    rc = fn();
    (void)hipEventRecord(e1, g->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < *best) *best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

struct Result {
  const char *param;
  int value;
  double ms;
};

}  // namespace

extern "C" int eg_tune_set(eg_graph *g, const char *param, int32_t value) {
  if (!g || !param) return eg_fail(EG_EINVAL, "null argument");
  int *slot = param_slot(g, param);
  if (!slot) return eg_fail(EG_EINVAL, "unknown tuning parameter '%s'", param);
  if (!strcmp(param, "pull_hot")) {
    if (value < 0) return eg_fail(EG_EINVAL, "pull_hot takes a vertex count >= 0");
  } else if (!strcmp(param, "pull_persist")) {
    if (value < 0 || value > 2) return eg_fail(EG_EINVAL, "pull_persist takes 0, 1 or 2");
  } else if (!strcmp(param, "ingest_kib")) {
    if (value < 0) return eg_fail(EG_EINVAL, "ingest_kib takes a size >= 0");
  } else if (!strcmp(param, "grid_band")) {
    if (value < 0) return eg_fail(EG_EINVAL, "grid_band takes a partition count >= 0");
  } else if (!strcmp(param, "pr_tile")) {
    if (value < 0 || (value > 0 && value < 1024))
      return eg_fail(EG_EINVAL, "pr_tile takes 0 or a length >= 1024");
  } else if (!strcmp(param, "pr_scatter")) {
    if (value != 8 && value != 16 && value != 32)
      return eg_fail(EG_EINVAL, "pr_scatter takes 8, 16 or 32");
  } else if (!strcmp(param, "pr_hot")) {
    if (value < 0 || value > kPbHotWindow * kPbHotWindows)
      return eg_fail(EG_EINVAL, "pr_hot takes 0..%d", kPbHotWindow * kPbHotWindows);
  } else if (!strcmp(param, "bfs_order")) {
    if (value < 0 || value > 2) return eg_fail(EG_EINVAL, "bfs_order takes 0, 1 or 2");
  } else if (!strcmp(param, "pr_persist")) {
    if (value != 0 && value != 1) return eg_fail(EG_EINVAL, "pr_persist takes 0 or 1");
  } else if (!strcmp(param, "pr_xcd")) {
    if (value != 0 && value != 1) return eg_fail(EG_EINVAL, "pr_xcd takes 0 or 1");
  } else if (!strcmp(param, "ms_lanes")) {
    if (value != 64 && value != 128 && value != 256)
      return eg_fail(EG_EINVAL, "ms_lanes takes 64, 128 or 256");
  } else if (!strcmp(param, "ms_device")) {
    if (value != 0 && value != 1) return eg_fail(EG_EINVAL, "ms_device takes 0 or 1");
  } else if (!strcmp(param, "ms_push_div")) {
    if (value < 1 || value > 1000) return eg_fail(EG_EINVAL, "ms_push_div takes 1..1000");
  } else if (!strcmp(param, "wide_part")) {
    if (value != 0 && (value < 10 || value > 31))
      return eg_fail(EG_EINVAL, "wide_part takes 0 or a part size exponent 10..31");
  } else if (!strcmp(param, "wcc_order")) {
    if (value != 0 && value != 1) return eg_fail(EG_EINVAL, "wcc_order takes 0 or 1");
  } else if (!strcmp(param, "pr_binned")) {
    if (value < -1 || value > 1) return eg_fail(EG_EINVAL, "pr_binned takes -1, 0 or 1");
  } else if (!strcmp(param, "graphs")) {
    if (value != 0 && value != 1) return eg_fail(EG_EINVAL, "graphs takes 0 or 1");
  } else if (value != 8 && value != 16) {
    return eg_fail(EG_EINVAL, "%s takes 8 or 16", param);
  }
  *slot = value;
  return EG_OK;
}

extern "C" int eg_tune_get(eg_graph *g, const char *param, int32_t *value) {
  if (!g || !param || !value) return eg_fail(EG_EINVAL, "null argument");
  int *slot = param_slot(g, param);
  if (!slot) return eg_fail(EG_EINVAL, "unknown tuning parameter '%s'", param);
  *value = *slot;
  return EG_OK;
}

extern "C" int eg_autotune(eg_graph *g, uint32_t kernels, eg_tune_entry *out, int32_t cap,
                           int32_t *count) {
  EG_TRY(eg_require_loaded(g));
  if (cap < 0 || (cap > 0 && !out)) return eg_fail(EG_EINVAL, "bad output array");
  const bool adj = g->out_ptr != nullptr;
  std::vector<Result> res;
  if (kernels & EG_TUNE_GRID) {
    if (!g->cell_ptr) return eg_fail(EG_ENOTSUP, "grid tuning needs a grid layout");
    eg_pr_opts o = {0.85, 3, EG_PR_DANGLING_DROP, 0.0, EG_LAYOUT_GRID};
    g->tune.pr_binned = 0;  // the bands are timed on the grid kernel
    EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // builds the compute copy
    const int bands[4] = {1, 2, 4, 0};
    for (int b : bands) {
      if (b >= (int)g->P) continue;
      g->tune.grid_band = b;
      double ms;
      EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); },
                       &ms));
      res.push_back({"grid_band", b, ms / o.iterations});
    }
    double best = 1e300;
    for (const Result &r : res)
      if (!strcmp(r.param, "grid_band") && r.ms < best) {
        best = r.ms;
        g->tune.grid_band = r.value;
      }
    // the binned kernels over the grid's entries against the best band
    for (int pb = 0; pb <= 1; pb++) {
      g->tune.pr_binned = pb;
      double ms;
      EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // warm (binned plan)
      EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); }, &ms));
      res.push_back({"pr_binned", pb, ms / o.iterations});
    }
  }
  const int variants[2] = {8, 16};
  if (kernels & EG_TUNE_PULL) {
    if (!adj) return eg_fail(EG_ENOTSUP, "pull tuning needs adjacency lists");
    eg_pr_opts o = {0.85, 5, EG_PR_DANGLING_DROP, 0.0, EG_LAYOUT_ADJ};
    g->tune.pr_binned = 0;  // the pull shapes are timed on the pull kernel
    // tuning serves long runs, which use the degree-ordered copy: time that
    if (eg_reorder_enabled()) {
      EgOrdered *ord;
      EG_TRY(eg_ordered(g, &ord));
    }
    for (int v : variants) {
      g->tune.pull_batch = v;
      double ms;
      EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // warm (plan, order, state)
      EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); }, &ms));
      res.push_back({"pull_batch", v, ms / o.iterations});
    }
    // keep the faster depth, then try cache limits for the ordered copy
    g->tune.pull_batch = res[0].ms <= res[1].ms ? res[0].value : res[1].value;
    if (eg_reorder_enabled()) {
      const int hots[4] = {0, (int)(g->n / 64), (int)(g->n / 16), (int)(g->n / 4)};
      for (int i = 0; i < 4; i++) {
        const int h = hots[i];
        if (i > 0 && h == hots[i - 1]) continue;
        g->tune.pull_hot = h;
        double ms;
        EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); },
                         &ms));
        res.push_back({"pull_hot", h, ms / o.iterations});
      }
      // keep the best limit, then try persistent workgroups with an LDS hot set
      double best = 1e300;
      for (const Result &r : res)
        if (!strcmp(r.param, "pull_hot") && r.ms < best) {
          best = r.ms;
          g->tune.pull_hot = r.value;
        }
      for (int pp = 0; pp <= 2; pp++) {
        g->tune.pull_persist = pp;
        double ms;
        const int rc = time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); },
                                 &ms);
        if (rc != EG_OK) {
          if (pp == 0 || rc == EG_EDEVICE) return rc;
          continue;  // this device cannot give the variant its LDS: not a candidate
        }
        res.push_back({"pull_persist", pp, ms / o.iterations});
      }
      double best_pp = 1e300;
      for (const Result &r : res)
        if (!strcmp(r.param, "pull_persist") && r.ms < best_pp) {
          best_pp = r.ms;
          g->tune.pull_persist = r.value;
        }
    }
    // the binned scatter's tile order, then the binned two-phase kernels
    // against the best pull shape
    g->tune.pr_binned = 1;
    EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // warm (binned plan)
    double xms[2];
    for (int x = 0; x <= 1; x++) {
      g->tune.pr_xcd = x;
      EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); },
                       &xms[x]));
      res.push_back({"pr_xcd", x, xms[x] / o.iterations});
    }
    g->tune.pr_xcd = xms[1] < xms[0] ? 1 : 0;
    // hot window caps (each plan keeps only the windows that hold 2% of the
    // entries, so caps past the graph's skew give the same plan)
    {
      const int caps[4] = {kPbHotWindow, 4 * kPbHotWindow, 8 * kPbHotWindow,
                           kPbHotWindows * kPbHotWindow};
      double best = 1e300;
      int best_cap = g->tune.pr_hot;
      uint32_t prev_hot = 0xFFFFFFFFu;
      for (int c : caps) {
        g->tune.pr_hot = c;
        EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // warm (binned plan)
        uint32_t geo[7];
        EG_TRY(eg_pr_binned_geometry(g, geo));
        if (geo[6] == prev_hot) continue;  // trimmed to the previous cap's plan
        prev_hot = geo[6];
        double ms;
        EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); },
                         &ms));
        res.push_back({"pr_hot", c, ms / o.iterations});
        if (ms < best) {
          best = ms;
          best_cap = c;
        }
      }
      g->tune.pr_hot = best_cap;
      // persistent phase-2 workgroups against one per work item, on that plan
      EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // warm (binned plan)
      double pms[2];
      for (int pp = 0; pp <= 1; pp++) {
        g->tune.pr_persist = pp;
        EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); },
                         &pms[pp]));
        res.push_back({"pr_persist", pp, pms[pp] / o.iterations});
      }
      g->tune.pr_persist = pms[1] <= pms[0] ? 1 : 0;
    }
    for (int pb = 0; pb <= 1; pb++) {
      g->tune.pr_binned = pb;
      double ms;
      EG_TRY(eg_pagerank(g, &o, nullptr, nullptr, nullptr));  // warm (binned plan)
      EG_TRY(time_best(g, 3, [&] { return eg_pagerank(g, &o, nullptr, nullptr, nullptr); }, &ms));
      res.push_back({"pr_binned", pb, ms / o.iterations});
    }
  }
  std::vector<uint32_t> roots;
  if (kernels & (EG_TUNE_SPMM | EG_TUNE_MSBFS)) {
    if (!adj) return eg_fail(EG_ENOTSUP, "batched tuning needs adjacency lists");
    std::vector<uint32_t> deg(g->n);
    EG_HIP(hipMemcpyAsync(deg.data(), g->out_deg, (size_t)g->n * 4, hipMemcpyDeviceToHost,
                          g->stream));
    EG_HIP(hipStreamSynchronize(g->stream));
    // up to 256 roots drawn uniformly (fixed seed) from the vertices with
    // edges, as queries arrive: the SpMM takes the first 64, the multi-source
    // BFS all of them (one traversal of its widest lane count); ids in
    // order when draws keep missing
    std::vector<uint8_t> taken(g->n, 0);
    uint64_t h = 0x243F6A8885A308D3ull;
    for (uint32_t tries = 0; tries < 64 * 256 && roots.size() < 256; tries++) {
      h += 0x9E3779B97F4A7C15ull;
      uint64_t z = h;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      const uint32_t v = (uint32_t)((z ^ (z >> 31)) % g->n);
      if (deg[v] && !taken[v]) {
        taken[v] = 1;
        roots.push_back(v);
      }
    }
    for (uint32_t v = 0; v < g->n && roots.size() < 256; v++)
      if (deg[v] && !taken[v]) {
        taken[v] = 1;
        roots.push_back(v);
      }
    if (roots.empty()) roots.push_back(0);
  }
  if (kernels & EG_TUNE_SPMM) {
    const uint32_t k = (uint32_t)std::min<size_t>(roots.size(), 64);
    std::vector<uint32_t> sp(k + 1);
    for (uint32_t i = 0; i <= k; i++) sp[i] = i;
    for (int v : variants) {
      g->tune.spmm_depth = v;
      double ms;
      EG_TRY(eg_ppr_batch(g, k, sp.data(), roots.data(), 0.15, 1, nullptr, nullptr));
      EG_TRY(time_best(
          g, 2, [&] { return eg_ppr_batch(g, k, sp.data(), roots.data(), 0.15, 3, nullptr, nullptr); },
          &ms));
      res.push_back({"spmm_depth", v, ms / 3});
    }
  }
  if (kernels & EG_TUNE_MSBFS) {
    const uint32_t k = (uint32_t)roots.size();
    double best = 1e300;
    int best_v = g->tune.ms_batch;
    for (int v : variants) {
      g->tune.ms_batch = v;
      double ms;
      EG_TRY(eg_bfs_batch(g, k, roots.data(), nullptr, nullptr));
      EG_TRY(time_best(g, 2, [&] { return eg_bfs_batch(g, k, roots.data(), nullptr, nullptr); },
                       &ms));
      res.push_back({"ms_batch", v, ms});
      if (ms < best) {
        best = ms;
        best_v = v;
      }
    }
    g->tune.ms_batch = best_v;
    // the push / pull threshold of the union frontier (symmetric graphs only
    // push); levels are the same whichever direction a level takes
    if ((g->flags & (EG_UNDIRECTED | EG_SYMMETRIC)) || !g->in_ptr) {
      for (int div : {5, 20, 40}) {
        g->tune.ms_push_div = div;
        double ms;
        EG_TRY(eg_bfs_batch(g, k, roots.data(), nullptr, nullptr));
        EG_TRY(time_best(g, 2, [&] { return eg_bfs_batch(g, k, roots.data(), nullptr, nullptr); },
                         &ms));
        res.push_back({"ms_push_div", div, ms});
      }
    }
  }
  // keep the fastest variant of each parameter (of equal timings, the first)
  for (size_t i = 0; i < res.size(); i++) {
    bool best = true;
    for (size_t j = 0; j < res.size(); j++)
      if (!strcmp(res[j].param, res[i].param) &&
          (res[j].ms < res[i].ms || (res[j].ms == res[i].ms && j < i)))
        best = false;
    if (best) *param_slot(g, res[i].param) = res[i].value;
  }
  int32_t nout = 0;
  for (const Result &r : res) {
    if (nout < cap) {
      memset(&out[nout], 0, sizeof(out[nout]));
      strncpy(out[nout].param, r.param, sizeof(out[nout].param) - 1);
      out[nout].value = r.value;
      out[nout].ms = r.ms;
      out[nout].chosen = *param_slot(g, r.param) == r.value;
    }
    nout++;
  }
  if (count) *count = nout;
  return EG_OK;
}
