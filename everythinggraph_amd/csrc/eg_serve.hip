// eg_serve.hip -- serving form: continuous batching of BFS and personalized
// PageRank queries over one resident graph.  No reference counterpart; each
// query's result is defined as eg_bfs's levels (bfs_simple.c:237-281, root =
// 1) or as the PPR recurrence of eg_ppr_batch run until the query's own L1
// change reaches the tolerance.
//
// One engine thread owns every device step.  BFS queries occupy lanes of
// 64-lane groups (eg_batch.h MsState): a query seeded at a level boundary is
// advanced by every later level of its group, with its own level count, and
// retires at the first level at which it reaches nothing.  PPR queries occupy
// columns of one 64-column block: a query initialised at an iteration
// boundary retires when its column's L1 change falls to the tolerance.  BFS
// levels run on one stream and PPR iterations on another, so a level's
// kernels fill the device while an SpMM iteration is in flight.  Per-query
// results are independent of the other lanes and columns: levels are exact,
// and a PPR column's arithmetic (summation order included) involves only
// that column.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "eg_batch.h"

namespace {

using Clock = std::chrono::steady_clock;

double now_ms() {
  return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count();
}

struct Query {
  uint64_t ticket = 0;
  int32_t kind = EG_REQ_BFS;
  uint32_t root = 0;
  std::vector<uint32_t> seeds;
  void *out = nullptr;
  double t_submit = 0.0, t_start = 0.0, t_done = 0.0;
  uint32_t level = 0;  // BFS: largest level so far (root = 1); PPR: iterations
  bool done = false;
  bool claimed = false;  // a thread waits on this ticket (one waiter per ticket)
  eg_serve_result res{};
  std::string err;
};

struct LaneGroup {
  MsState st;
  uint32_t cap = 64;  // usable lanes
  Query *lane[64] = {};
  uint64_t active = 0;
};

constexpr uint16_t kMaxLevel = 65534;

}  // namespace

struct eg_server {
  eg_graph *g = nullptr;
  eg_serve_opts o{};
  BatchView bv{};
  hipStream_t s_bfs = nullptr, s_ppr = nullptr;
  hipEvent_t ppr_ev = nullptr, ppr_t0 = nullptr;  // end / start of the SpMM step in flight
  bool ppr_busy = false;
  int ppr_wclass = 0;        // slot width of that step: 0, 1, 2 = 16, 32, 64 columns
  // Measured device ms of an SpMM step per width class and of a
  // single-vector step (moving averages; 0 = not measured yet).
  double spmm_ms[3] = {0.0, 0.0, 0.0};
  double solo_ms = 0.0;
  double ppr_at = 0.0;  // host clock (ms) at the launch of the SpMM step in flight
  std::vector<LaneGroup *> groups;
  uint32_t max_groups = 1;
  bool bfs_nomem = false;  // no lane group could be allocated
  PprState ppr;
  bool ppr_alloc = false, ppr_nomem = false;
  Query *col[64] = {};
  uint64_t cols = 0;
  // single-vector PPR query stepped on the PPR stream (light load).  Up to
  // two iterations are in flight, indexed by the parity of the iteration
  // they compute, so the device keeps a step queued while the engine thread
  // waits on a BFS level.
  Ppr1 solo;
  Query *solo_q = nullptr;
  int solo_fly = 0;                        // iterations in flight (0..2)
  hipEvent_t solo_ev[2] = {}, solo_t0[2] = {};
  double solo_at[2] = {0.0, 0.0};          // expected host-clock start of each (ms)
  double solo_last = -1.0;                 // last known L1 change (< 0: none yet)
  double *h_solo = nullptr;  // pinned, L1 change per iteration parity
  uint32_t *stage = nullptr;    // n words, BFS stream
  float *stage_ppr = nullptr;   // n floats, PPR stream
  uint64_t device_bytes = 0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Query *> pend_bfs, pend_ppr;
  std::unordered_map<uint64_t, Query *> all;
  uint64_t next_ticket = 1;
  bool stopping = false;
  int fatal = EG_OK;
  std::string fatal_msg;
  eg_serve_stats stats{};
  std::thread th;
};

namespace {

uint64_t ms_group_bytes(uint64_t n) {
  return 32 * n + 128 * n + 8 * n + 600 + eg_scan_tmp_bytes(n + 64) + 8 * (8 + 2 * 1024 + 4160);
}

uint64_t ppr_block_bytes(const eg_graph *g, const PullPlan &pl) {
  const uint64_t n = g->n;
  const uint64_t nfin = ((uint64_t)pl.nlong * 64 + 255) / 256;
  return 4 * 256 * n + 8 * n + 512 * (uint64_t)(pl.nseg ? pl.nseg : 1) +
         512 * ((uint64_t)pl.nchunks + nfin + 512) + 4096;
}

bool memory_allows(eg_server *S, uint64_t bytes) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  return (uint64_t)free_b >= bytes + S->o.hbm_reserve;
}

// Marks q finished (status rc) and wakes its waiter.
void complete(eg_server *S, Query *q, int rc, const char *msg) {
  std::lock_guard<std::mutex> lk(S->mu);
  q->t_done = now_ms();
  q->res.kind = q->kind;
  q->res.status = rc;
  q->res.queue_ms = (q->t_start > 0.0 ? q->t_start : q->t_done) - q->t_submit;
  q->res.latency_ms = q->t_done - q->t_submit;
  if (rc != EG_OK && msg) q->err = msg;
  q->done = true;
  if (rc == EG_OK)
    S->stats.completed++;
  else
    S->stats.failed++;
  if (q->t_start > 0.0) {
    if (q->kind == EG_REQ_BFS)
      S->stats.bfs_inflight--;
    else
      S->stats.ppr_inflight--;
  }
  S->cv_done.notify_all();
}

// Takes up to k queries from a pending queue.
std::vector<Query *> take(eg_server *S, std::deque<Query *> &dq, uint32_t k) {
  std::vector<Query *> v;
  std::lock_guard<std::mutex> lk(S->mu);
  while (k-- && !dq.empty()) {
    v.push_back(dq.front());
    dq.pop_front();
  }
  S->stats.pending = (uint32_t)(S->pend_bfs.size() + S->pend_ppr.size());
  return v;
}

void fail_pending(eg_server *S, std::deque<Query *> &dq, int rc, const char *msg) {
  for (Query *q : take(S, dq, 0xFFFFFFFFu)) complete(S, q, rc, msg);
}

void started(eg_server *S, Query *q) {
  std::lock_guard<std::mutex> lk(S->mu);
  q->t_start = now_ms();
  if (q->kind == EG_REQ_BFS)
    S->stats.bfs_inflight++;
  else
    S->stats.ppr_inflight++;
}

int ensure_stage(eg_server *S) {
  if (!S->stage) {
    EG_TRY(eg_dalloc_t(&S->stage, S->g->n, "serve staging"));
    EG_TRY(eg_dalloc_t(&S->stage_ppr, S->g->n, "serve staging"));
    S->device_bytes += 8ull * S->g->n;
  }
  return EG_OK;
}

// ---- PPR ----
int ppr_prepare(eg_server *S) {
  if (S->ppr_alloc || S->ppr_nomem) return EG_OK;
  if (!memory_allows(S, ppr_block_bytes(S->g, *S->bv.pl))) {
    S->ppr_nomem = true;
    return EG_OK;
  }
  EG_TRY(eg_ppr_alloc(S->g, &S->ppr));
  S->ppr.stream = S->s_ppr;
  S->ppr.max_grid = S->o.ppr_grid;
  EG_TRY(eg_ppr_reset(S->g, &S->ppr));
  S->ppr_alloc = true;
  S->device_bytes += ppr_block_bytes(S->g, *S->bv.pl);
  std::lock_guard<std::mutex> lk(S->mu);
  S->stats.ppr_columns = S->o.ppr_columns;
  return EG_OK;
}

int ppr_admit(eg_server *S) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->pend_ppr.empty()) return EG_OK;
  }
  EG_TRY(ppr_prepare(S));
  if (S->ppr_nomem) {
    fail_pending(S, S->pend_ppr, EG_ENOMEM, "not enough free device memory for the PPR block");
    return EG_OK;
  }
  const uint32_t cap = S->o.ppr_columns;
  uint32_t nfree = 0;
  for (uint32_t q = 0; q < cap; q++) nfree += !((S->cols >> q) & 1ull);
  for (Query *qr : take(S, S->pend_ppr, nfree)) {
    uint32_t q = 0;
    while ((S->cols >> q) & 1ull) q++;
    EG_TRY(eg_ppr_col_init(S->g, S->bv, &S->ppr, q, qr->seeds.data(), (uint32_t)qr->seeds.size(),
                           S->o.ppr_alpha));
    S->col[q] = qr;
    S->cols |= 1ull << q;
    qr->level = 0;
    started(S, qr);
  }
  return EG_OK;
}

// After a completed step: count it for every column, retire converged ones.
int ppr_after(eg_server *S) {
  uint64_t retire = 0;
  for (uint32_t q = 0; q < 64; q++) {
    if (!((S->cols >> q) & 1ull)) continue;
    Query *qr = S->col[q];
    qr->level++;
    qr->res.delta = S->ppr.h_delta[q];
    if (qr->res.delta <= S->o.ppr_tol || (int32_t)qr->level >= S->o.ppr_max_iter)
      retire |= 1ull << q;
  }
  if (!retire) return EG_OK;
  for (uint32_t q = 0; q < 64; q++) {
    if (!((retire >> q) & 1ull)) continue;
    Query *qr = S->col[q];
    if (qr->out) {
      EG_TRY(ensure_stage(S));
      EG_TRY(eg_ppr_extract_col(S->g, S->bv, &S->ppr, q, S->stage_ppr));
      EG_HIP(hipMemcpyAsync(qr->out, S->stage_ppr, (size_t)S->g->n * 4, hipMemcpyDeviceToHost,
                            S->s_ppr));
      EG_HIP(hipStreamSynchronize(S->s_ppr));
    }
    qr->res.levels = qr->level;
    qr->res.reached = 0;
  }
  EG_TRY(eg_ppr_clear_cols(S->g, &S->ppr, retire));
  for (uint32_t q = 0; q < 64; q++) {
    if (!((retire >> q) & 1ull)) continue;
    Query *qr = S->col[q];
    S->col[q] = nullptr;
    S->cols &= ~(1ull << q);
    complete(S, qr, EG_OK, nullptr);
  }
  return EG_OK;
}

// Slot-width class of an SpMM step over the columns `cols` (as ppr_width).
int width_class(uint64_t cols) {
  if (!cols) return 0;
  const int hi = 63 - __builtin_clzll(cols);
  return hi < 16 ? 0 : hi < 32 ? 1 : 2;
}

void ewma(double &avg, double ms) { avg = avg > 0.0 ? 0.75 * avg + 0.25 * ms : ms; }

int ppr_launch(eg_server *S) {
  S->ppr.active = S->cols;
  S->ppr_wclass = width_class(S->cols);
  EG_HIP(hipEventRecord(S->ppr_t0, S->s_ppr));
  EG_TRY(eg_ppr_step(S->g, S->bv, &S->ppr, S->o.ppr_alpha, true));
  EG_HIP(hipEventRecord(S->ppr_ev, S->s_ppr));
  S->ppr_busy = true;
  S->ppr_at = now_ms();
  std::lock_guard<std::mutex> lk(S->mu);
  S->stats.ppr_steps++;
  return EG_OK;
}

// ---- BFS ----
int bfs_add_group(eg_server *S, bool *ok) {
  *ok = false;
  if (S->groups.size() >= S->max_groups) return EG_OK;
  const uint64_t bytes = ms_group_bytes(S->g->n);
  if (!memory_allows(S, bytes)) return EG_OK;
  LaneGroup *G = new LaneGroup();
  const uint32_t i = (uint32_t)S->groups.size();
  G->cap = std::min<uint32_t>(64, S->o.bfs_lanes - 64 * i);
  int rc = eg_ms_alloc(S->g, &G->st, 1);
  if (rc == EG_OK) {
    G->st.stream = S->s_bfs;
    rc = eg_ms_reset(S->g, &G->st);
  }
  if (rc != EG_OK) {
    eg_ms_free(&G->st);
    delete G;
    return rc;
  }
  S->groups.push_back(G);
  S->device_bytes += bytes;
  *ok = true;
  std::lock_guard<std::mutex> lk(S->mu);
  S->stats.bfs_lanes += G->cap;
  return EG_OK;
}

int bfs_admit(eg_server *S) {
  uint32_t want;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    want = (uint32_t)S->pend_bfs.size();
  }
  if (!want) return EG_OK;
  uint32_t nfree = 0;
  for (LaneGroup *G : S->groups) nfree += G->cap - (uint32_t)__builtin_popcountll(G->active);
  while (nfree < want) {
    bool ok = false;
    EG_TRY(bfs_add_group(S, &ok));
    if (!ok) break;
    nfree += S->groups.back()->cap;
  }
  if (S->groups.empty()) {
    S->bfs_nomem = true;
    fail_pending(S, S->pend_bfs, EG_ENOMEM, "not enough free device memory for a BFS lane group");
    return EG_OK;
  }
  std::vector<Query *> got = take(S, S->pend_bfs, nfree);
  size_t next = 0;
  for (LaneGroup *G : S->groups) {
    uint32_t lanes[64], roots[64], k = 0;
    for (uint32_t l = 0; l < G->cap && next < got.size(); l++) {
      if ((G->active >> l) & 1ull) continue;
      Query *qr = got[next++];
      G->lane[l] = qr;
      G->active |= 1ull << l;
      qr->level = 1;
      lanes[k] = l;
      roots[k] = qr->root;
      k++;
      started(S, qr);
    }
    if (k) EG_TRY(eg_ms_seed(S->g, S->bv, &G->st, k, lanes, roots));
  }
  return EG_OK;
}

int bfs_level(eg_server *S, LaneGroup *G) {
  uint16_t tab[64] = {};
  uint64_t deep = 0;
  for (uint32_t l = 0; l < 64; l++) {
    if (!((G->active >> l) & 1ull)) continue;
    if (G->lane[l]->level >= kMaxLevel)
      deep |= 1ull << l;
    else
      tab[l] = (uint16_t)(G->lane[l]->level + 1);
  }
  uint64_t progress = 0;
  const uint64_t run = G->active & ~deep;
  if (run) EG_TRY(eg_ms_level(S->g, S->bv, &G->st, &run, tab, &progress));
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.bfs_levels++;
  }
  const uint64_t retire = (run & ~progress) | deep;
  for (uint32_t l = 0; l < 64; l++)
    if ((progress >> l) & 1ull) G->lane[l]->level++;
  if (!retire) return EG_OK;
  uint32_t lanes[64], k = 0;
  for (uint32_t l = 0; l < 64; l++)
    if ((retire >> l) & 1ull) lanes[k++] = l;
  uint64_t reached[64];
  EG_TRY(eg_ms_reached(S->g, &G->st, k, lanes, reached));
  for (uint32_t i = 0; i < k; i++) {
    Query *qr = G->lane[lanes[i]];
    qr->res.levels = qr->level;
    qr->res.reached = reached[i];
    if (qr->out && !((deep >> lanes[i]) & 1ull)) {
      EG_TRY(ensure_stage(S));
      EG_TRY(eg_ms_extract_lane(S->g, S->bv, &G->st, lanes[i], S->stage));
      EG_HIP(hipMemcpyAsync(qr->out, S->stage, (size_t)S->g->n * 4, hipMemcpyDeviceToHost,
                            S->s_bfs));
      EG_HIP(hipStreamSynchronize(S->s_bfs));
    }
  }
  EG_TRY(eg_ms_clear_lanes(S->g, &G->st, &retire));
  for (uint32_t i = 0; i < k; i++) {
    Query *qr = G->lane[lanes[i]];
    G->lane[lanes[i]] = nullptr;
    G->active &= ~(1ull << lanes[i]);
    if ((deep >> lanes[i]) & 1ull) {
      // deeper than the 16-bit level store: the query runs again on the
      // single-source path, whose levels have no such limit, so its result
      // does not depend on whether it was batched
      eg_bfs_stats st{};
      const int rc =
          eg_bfs_host(S->g, qr->root, 0, 1, static_cast<uint32_t *>(qr->out), nullptr, &st);
      if (rc == EG_EDEVICE) return rc;
      qr->res.levels = (uint32_t)st.levels;
      qr->res.reached = st.reached;
      complete(S, qr, rc, rc ? eg_last_error() : nullptr);
    } else {
      complete(S, qr, EG_OK, nullptr);
    }
  }
  return EG_OK;
}

bool bfs_active(eg_server *S) {
  for (LaneGroup *G : S->groups)
    if (G->active) return true;
  return false;
}

int plain_round(eg_server *S);
int poll(eg_server *S, hipEvent_t ev, bool *done);

size_t waiting(eg_server *S, const std::deque<Query *> &dq) {
  std::lock_guard<std::mutex> lk(S->mu);
  return dq.size();
}

// Light-load PPR: the oldest waiting query runs on the single-vector kernel,
// one iteration per round on the PPR stream, overlapping BFS work.
int solo_start(eg_server *S) {
  std::vector<Query *> q = take(S, S->pend_ppr, 1);
  if (q.empty()) return EG_OK;
  Query *qr = q[0];
  started(S, qr);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.served_plain++;
  }
  qr->level = 0;
  const int rc = eg_ppr1_begin(S->g, &S->solo, qr->seeds.data(), (uint32_t)qr->seeds.size(),
                               S->o.ppr_alpha, S->o.ppr_max_iter, S->s_ppr, S->h_solo);
  if (rc != EG_OK) {
    if (rc == EG_EDEVICE) return rc;
    complete(S, qr, rc, eg_last_error());
    return EG_OK;
  }
  S->solo_q = qr;
  S->solo_fly = 0;
  S->solo_last = -1.0;
  return EG_OK;
}

// Launches the solo query's next iteration; its L1 change lands in the
// pinned slot of its parity.
int solo_launch(eg_server *S) {
  const int p = (S->solo.it + 1) & 1;
  S->solo.h_delta = S->h_solo + p;
  EG_HIP(hipEventRecord(S->solo_t0[p], S->s_ppr));
  EG_TRY(eg_ppr1_step(S->g, &S->solo));
  EG_HIP(hipEventRecord(S->solo_ev[p], S->s_ppr));
  const double t = now_ms();
  S->solo_at[p] = S->solo_fly ? std::max(t, S->solo_at[p ^ 1] + S->solo_ms) : t;
  S->solo_fly++;
  std::lock_guard<std::mutex> lk(S->mu);
  S->stats.ppr_steps++;
  return EG_OK;
}

// Keeps the solo pipeline full: one iteration always, a second while the
// last known L1 change is above 2 tol (the change shrinks by about 1 - alpha
// per iteration, so the iteration in flight will not end the query).
int solo_fill(eg_server *S) {
  while (S->solo_q && S->solo_fly < 2 && S->solo.it < S->solo.max_iter) {
    if (S->solo_fly && S->solo_last >= 0.0 && S->solo_last <= 2.0 * S->o.ppr_tol) break;
    EG_TRY(solo_launch(S));
  }
  return EG_OK;
}

// Takes the solo iterations that have ended, oldest first.  The query ends
// at the first whose L1 change is <= tol (or at the iteration cap); an
// iteration launched after that one is taken back, so the output is the
// ending iteration's iterate, as on the one-step-at-a-time path.
int solo_progress(eg_server *S) {
  while (S->solo_fly) {
    const int j = S->solo.it - S->solo_fly + 1, p = j & 1;  // oldest in flight
    bool done = false;
    EG_TRY(poll(S, S->solo_ev[p], &done));
    if (!done) return EG_OK;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, S->solo_t0[p], S->solo_ev[p]) == hipSuccess) ewma(S->solo_ms, ms);
    S->solo_fly--;
    const double delta = S->h_solo[p];
    S->solo_last = delta;
    if (delta <= S->o.ppr_tol || j >= S->o.ppr_max_iter) {
      while (S->solo_fly) {
        EG_TRY(eg_ppr1_undo(S->g, &S->solo));
        S->solo_fly--;
      }
      Query *qr = S->solo_q;
      qr->level = (uint32_t)j;
      qr->res.delta = delta;
      EG_TRY(eg_ppr1_output(S->g, &S->solo, static_cast<float *>(qr->out)));
      qr->res.levels = qr->level;
      qr->res.reached = 0;
      S->solo_q = nullptr;
      complete(S, qr, EG_OK, nullptr);
      return EG_OK;
    }
  }
  return EG_OK;
}

// Light-load BFS: the oldest waiting query on the direction-optimising
// single-source path (synchronous, on the graph's stream).
int plain_bfs(eg_server *S) {
  std::vector<Query *> q = take(S, S->pend_bfs, 1);
  if (q.empty()) return EG_OK;
  Query *qr = q[0];
  started(S, qr);
  eg_bfs_stats st{};
  const int rc = eg_bfs_host(S->g, qr->root, 0, 1, static_cast<uint32_t *>(qr->out), nullptr, &st);
  if (rc == EG_EDEVICE) return rc;
  qr->res.levels = (uint32_t)st.levels;
  qr->res.reached = st.reached;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.bfs_levels += (uint64_t)st.levels;
    S->stats.served_plain++;
  }
  complete(S, qr, rc, rc ? eg_last_error() : nullptr);
  return EG_OK;
}

int poll(eg_server *S, hipEvent_t ev, bool *done) {
  const hipError_t e = hipEventQuery(ev);
  *done = e == hipSuccess;
  if (e != hipSuccess && e != hipErrorNotReady) EG_HIP(e);
  return EG_OK;
}

// Largest PPR backlog served one query at a time on the single-vector path.
// An SpMM step over w columns costs spmm_ms of the width class that holds w;
// it replaces w single-vector steps, so the block pays once w exceeds the
// measured ratio of the two.  Until both are measured: twice the BFS
// threshold.
size_t ppr_light_limit(const eg_server *S, size_t light, size_t waiting_now) {
  const size_t base = 2 * light;
  if (!light) return base;
  const size_t w = waiting_now < 64 ? waiting_now : 64;
  const double blk = S->spmm_ms[width_class(w ? (1ull << (w - 1)) : 1ull)];
  // Until both steps are timed, a block of w columns is taken to cost about
  // 4 + w/8 single-vector steps (the binned kernels on RMAT-24).
  const double ratio =
      blk > 0.0 && S->solo_ms > 0.0 ? blk / S->solo_ms : 4.0 + (double)w / 8.0;
  const size_t lim = (size_t)ceil(ratio);
  return lim < base ? base : (lim > 63 ? 63 : lim);
}

// Only PPR work is in flight: wait until a step ends or a BFS query arrives.
// The engine sleeps on the queue until shortly before the step's expected end
// (its moving-average duration from the host-side launch time), then polls
// the step events, so a step's end is seen within microseconds rather than a
// timer quantum (a sleep of 200 us cost about 10% of a 1.4 ms iteration).
int wait_ppr(eg_server *S) {
  double end = 0.0;  // earliest expected end (host ms); 0 = unknown
  const int p0 = (S->solo.it - S->solo_fly + 1) & 1;  // oldest solo iteration in flight
  if (S->solo_fly && S->solo_ms > 0.0) end = S->solo_at[p0] + S->solo_ms;
  if (S->ppr_busy && S->spmm_ms[S->ppr_wclass] > 0.0) {
    const double e = S->ppr_at + S->spmm_ms[S->ppr_wclass];
    end = end > 0.0 ? std::min(end, e) : e;
  }
  const double margin = 0.15;  // ms: timer slack of a sleep plus jitter of the step
  const double t = now_ms();
  if (end > 0.0 && end - margin > t) {
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->cv_work.wait_for(lk, std::chrono::duration<double, std::milli>(end - margin - t),
                            [&] { return !S->pend_bfs.empty(); }))
      return EG_OK;
  }
  // poll for at most 20 ms before returning to the round loop
  const double stop = now_ms() + 20.0;
  for (;;) {
    bool done = false;
    if (S->solo_fly) {
      EG_TRY(poll(S, S->solo_ev[p0], &done));
      if (done) return EG_OK;
    }
    if (S->ppr_busy) {
      EG_TRY(poll(S, S->ppr_ev, &done));
      if (done) return EG_OK;
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (!S->pend_bfs.empty()) return EG_OK;
    }
    if (now_ms() > stop) return EG_OK;
    std::this_thread::yield();
  }
}

int batched_round(eg_server *S) {
  const size_t light = S->o.idle_plain > 0 ? (size_t)S->o.idle_plain : 0;
  // ---- PPR side (PPR stream) ----
  bool done = false;
  if (S->ppr_busy) {
    EG_TRY(poll(S, S->ppr_ev, &done));
    if (done) {
      S->ppr_busy = false;
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, S->ppr_t0, S->ppr_ev) == hipSuccess)
        ewma(S->spmm_ms[S->ppr_wclass], ms);
      EG_TRY(ppr_after(S));
    }
  }
  EG_TRY(solo_progress(S));
  const size_t light_ppr = ppr_light_limit(S, light, waiting(S, S->pend_ppr));
  if (!S->solo_q && !S->cols && light_ppr) {
    const size_t w = waiting(S, S->pend_ppr);
    if (w > 0 && w <= light_ppr) EG_TRY(solo_start(S));
  }
  EG_TRY(solo_fill(S));
  if (!S->ppr_busy) {
    // A short queue waits for the single-vector path; a longer one (or one
    // that arrives while columns are busy) joins the SpMM block.
    if (S->cols || !light_ppr || waiting(S, S->pend_ppr) > light_ppr) EG_TRY(ppr_admit(S));
    if (S->cols) EG_TRY(ppr_launch(S));
  }
  // ---- BFS side ----
  bool ran = false;
  if (light && !bfs_active(S)) {
    const size_t w = waiting(S, S->pend_bfs);
    if (w > 0 && w <= light) {
      EG_TRY(plain_bfs(S));
      ran = true;
    }
  }
  if (!ran) {
    EG_TRY(bfs_admit(S));
    for (LaneGroup *G : S->groups) {
      if (!G->active) continue;
      EG_TRY(bfs_level(S, G));
      ran = true;
    }
  }
  if (!ran && (S->ppr_busy || S->solo_fly)) EG_TRY(wait_ppr(S));
  return EG_OK;
}

// Plain mode: the oldest query alone, through the single-query paths (the
// direction-optimising BFS and the single-vector PPR pull).
int plain_round(eg_server *S) {
  Query *qr = nullptr;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    const bool b = !S->pend_bfs.empty(), p = !S->pend_ppr.empty();
    if (b && (!p || S->pend_bfs.front()->ticket < S->pend_ppr.front()->ticket)) {
      qr = S->pend_bfs.front();
      S->pend_bfs.pop_front();
    } else if (p) {
      qr = S->pend_ppr.front();
      S->pend_ppr.pop_front();
    }
    S->stats.pending = (uint32_t)(S->pend_bfs.size() + S->pend_ppr.size());
  }
  if (!qr) return EG_OK;
  started(S, qr);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.served_plain++;
  }
  if (qr->kind == EG_REQ_BFS) {
    eg_bfs_stats st{};
    const int rc = eg_bfs_host(S->g, qr->root, 0, 1, static_cast<uint32_t *>(qr->out), nullptr, &st);
    if (rc == EG_EDEVICE) return rc;
    qr->res.levels = (uint32_t)st.levels;
    qr->res.reached = st.reached;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stats.bfs_levels += (uint64_t)st.levels;
    }
    complete(S, qr, rc, rc ? eg_last_error() : nullptr);
    return EG_OK;
  }
  int iters = 0;
  double delta = 0.0;
  const int rc = eg_ppr_single(S->g, qr->seeds.data(), (uint32_t)qr->seeds.size(), S->o.ppr_alpha,
                               S->o.ppr_tol, S->o.ppr_max_iter, static_cast<float *>(qr->out),
                               &iters, &delta);
  if (rc == EG_EDEVICE) return rc;
  qr->res.levels = (uint32_t)iters;
  qr->res.delta = delta;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stats.ppr_steps += (uint64_t)iters;
  }
  complete(S, qr, rc, rc ? eg_last_error() : nullptr);
  return EG_OK;
}

void fail_everything(eg_server *S) {
  std::string msg = S->fatal_msg;
  fail_pending(S, S->pend_bfs, S->fatal, msg.c_str());
  fail_pending(S, S->pend_ppr, S->fatal, msg.c_str());
  for (LaneGroup *G : S->groups)
    for (uint32_t l = 0; l < 64; l++)
      if (G->lane[l]) {
        complete(S, G->lane[l], S->fatal, msg.c_str());
        G->lane[l] = nullptr;
      }
  for (uint32_t q = 0; q < 64; q++)
    if (S->col[q]) {
      complete(S, S->col[q], S->fatal, msg.c_str());
      S->col[q] = nullptr;
    }
  S->cols = 0;
  if (S->solo_q) {
    complete(S, S->solo_q, S->fatal, msg.c_str());
    S->solo_q = nullptr;
  }
}

void engine(eg_server *S) {
  (void)hipSetDevice(S->g->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        const bool work = !S->pend_bfs.empty() || !S->pend_ppr.empty() || S->cols ||
                          S->ppr_busy || S->solo_q || bfs_active(S);
        if (work) break;
        if (S->stopping) return;
        S->cv_work.wait(lk);
      }
    }
    const int rc = S->o.plain ? plain_round(S) : batched_round(S);
    if (rc != EG_OK) {
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->fatal = rc;
        S->fatal_msg = eg_last_error();
      }
      fail_everything(S);
      return;
    }
  }
}

int submit(eg_server *S, Query *q, uint64_t *ticket) {
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->fatal) {
    const int rc = S->fatal;
    const std::string msg = S->fatal_msg;
    lk.unlock();
    delete q;
    return eg_fail(rc, "server failed: %s", msg.c_str());
  }
  if (S->stopping) {
    lk.unlock();
    delete q;
    return eg_fail(EG_ESTATE, "server is stopping");
  }
  if (S->pend_bfs.size() + S->pend_ppr.size() >= S->o.max_pending) {
    lk.unlock();
    delete q;
    return eg_fail(EG_EAGAIN, "%u queries already pending", S->o.max_pending);
  }
  q->ticket = S->next_ticket++;
  q->t_submit = now_ms();
  S->all[q->ticket] = q;
  (q->kind == EG_REQ_BFS ? S->pend_bfs : S->pend_ppr).push_back(q);
  S->stats.submitted++;
  S->stats.pending = (uint32_t)(S->pend_bfs.size() + S->pend_ppr.size());
  *ticket = q->ticket;
  lk.unlock();
  S->cv_work.notify_one();
  return EG_OK;
}

void destroy(eg_server *S) {
  if (S->s_bfs) (void)hipStreamSynchronize(S->s_bfs);
  if (S->s_ppr) (void)hipStreamSynchronize(S->s_ppr);
  for (LaneGroup *G : S->groups) {
    eg_ms_free(&G->st);
    delete G;
  }
  S->groups.clear();
  if (S->ppr_alloc) eg_ppr_free(&S->ppr);
  eg_dfree_t(&S->stage);
  eg_dfree_t(&S->stage_ppr);
  for (auto &kv : S->all) delete kv.second;
  S->all.clear();
  if (S->ppr_ev) (void)hipEventDestroy(S->ppr_ev);
  if (S->ppr_t0) (void)hipEventDestroy(S->ppr_t0);
  for (int p = 0; p < 2; p++) {
    if (S->solo_ev[p]) (void)hipEventDestroy(S->solo_ev[p]);
    if (S->solo_t0[p]) (void)hipEventDestroy(S->solo_t0[p]);
  }
  if (S->h_solo) (void)hipHostFree(S->h_solo);
  if (S->s_bfs) (void)hipStreamDestroy(S->s_bfs);
  if (S->s_ppr) (void)hipStreamDestroy(S->s_ppr);
  if (S->g && S->g->server == S) S->g->server = nullptr;
  delete S;
}

}  // namespace

extern "C" int eg_serve_start(eg_graph *g, const eg_serve_opts *opts, eg_server **out) {
  if (!out) return eg_fail(EG_EINVAL, "null output");
  *out = nullptr;
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "serving needs adjacency lists (modes 2/3/4)");
  eg_serve_opts o{};
  if (opts) o = *opts;
  if (o.bfs_lanes == 0) o.bfs_lanes = 64;
  if (o.ppr_columns == 0) o.ppr_columns = 64;
  if (o.hbm_reserve == 0) o.hbm_reserve = 1ull << 30;
  if (o.ppr_alpha == 0.0) o.ppr_alpha = 0.15;
  if (o.ppr_tol == 0.0) o.ppr_tol = 1e-6;
  if (o.ppr_max_iter == 0) o.ppr_max_iter = 200;
  if (o.max_pending == 0) o.max_pending = 1u << 20;
  if (o.idle_plain == 0) o.idle_plain = 2;
  if (o.ppr_grid == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) !=
            hipSuccess ||
        cus <= 0)
      cus = 256;
    o.ppr_grid = 8u * (uint32_t)cus;
  }
  if (o.ppr_columns > 64 || !(o.ppr_alpha > 0.0 && o.ppr_alpha <= 1.0) || !(o.ppr_tol > 0.0) ||
      o.ppr_max_iter < 0 || o.bfs_lanes > 64 * 64)
    return eg_fail(EG_EINVAL, "bad serving options");
  if (o.plain) {
    o.bfs_lanes = 1;
    o.ppr_columns = 1;
  }
  EG_HIP(hipSetDevice(g->device));
  EG_HIP(hipStreamSynchronize(g->stream));
  eg_server *S = new eg_server();
  S->g = g;
  S->o = o;
  S->max_groups = (o.bfs_lanes + 63) / 64;
  int rc = eg_batch_view(g, &S->bv);  // builds the pull plan / ordered copy now
  // and the single-vector PPR and BFS layouts, which would otherwise stall the
  // engine thread on the first query that needs them
  if (rc == EG_OK) rc = eg_ppr1_prepare(g, o.ppr_max_iter);
  if (rc == EG_OK) rc = eg_bfs_prepare(g);
  if (rc == EG_OK && hipStreamCreateWithFlags(&S->s_bfs, hipStreamNonBlocking) != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "stream creation failed");
  if (rc == EG_OK && hipStreamCreateWithFlags(&S->s_ppr, hipStreamNonBlocking) != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "stream creation failed");
  // timed events: the engine measures its SpMM and single-vector steps
  if (rc == EG_OK && (hipEventCreate(&S->ppr_ev) != hipSuccess ||
                      hipEventCreate(&S->ppr_t0) != hipSuccess ||
                      hipEventCreate(&S->solo_ev[0]) != hipSuccess ||
                      hipEventCreate(&S->solo_t0[0]) != hipSuccess ||
                      hipEventCreate(&S->solo_ev[1]) != hipSuccess ||
                      hipEventCreate(&S->solo_t0[1]) != hipSuccess))
    rc = eg_fail(EG_EDEVICE, "event creation failed");
  if (rc == EG_OK && hipHostMalloc(reinterpret_cast<void **>(&S->h_solo), 2 * sizeof(double),
                                   hipHostMallocDefault) != hipSuccess)
    rc = eg_fail(EG_ENOMEM, "pinned solo delta");
  if (rc == EG_OK && hipStreamSynchronize(g->stream) != hipSuccess)
    rc = eg_fail(EG_EDEVICE, "graph stream failed");
  if (rc != EG_OK) {
    destroy(S);
    return rc;
  }
  g->server = S;
  S->th = std::thread(engine, S);
  *out = S;
  return EG_OK;
}

extern "C" int eg_serve_submit_bfs(eg_server *S, uint32_t root, uint32_t *out, uint64_t *ticket) {
  if (!S || !ticket) return eg_fail(EG_EINVAL, "null argument");
  if (root >= S->g->n) return eg_fail(EG_ERANGE, "root %u >= n = %u", root, S->g->n);
  Query *q = new Query();
  q->kind = EG_REQ_BFS;
  q->root = root;
  q->out = out;
  return submit(S, q, ticket);
}

extern "C" int eg_serve_submit_ppr(eg_server *S, const uint32_t *seeds, uint32_t nseeds,
                                   float *out, uint64_t *ticket) {
  if (!S || !ticket || (nseeds && !seeds)) return eg_fail(EG_EINVAL, "null argument");
  if (nseeds == 0) return eg_fail(EG_EINVAL, "a PPR query needs at least one seed");
  for (uint32_t i = 0; i < nseeds; i++)
    if (seeds[i] >= S->g->n) return eg_fail(EG_ERANGE, "seed %u >= n = %u", seeds[i], S->g->n);
  Query *q = new Query();
  q->kind = EG_REQ_PPR;
  q->seeds.assign(seeds, seeds + nseeds);
  q->out = out;
  return submit(S, q, ticket);
}

extern "C" int eg_serve_wait(eg_server *S, uint64_t ticket, double timeout_ms,
                             eg_serve_result *res) {
  if (!S) return eg_fail(EG_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->all.find(ticket);
  if (it == S->all.end()) return eg_fail(EG_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
  Query *q = it->second;
  // The waiter that sees the query done frees it: a second concurrent waiter
  // on the same ticket would read freed memory, so it is refused.
  if (q->claimed)
    return eg_fail(EG_EINVAL, "ticket %llu already has a waiter", (unsigned long long)ticket);
  q->claimed = true;
  if (timeout_ms < 0.0) {
    S->cv_done.wait(lk, [&] { return q->done; });
  } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms),
                                  [&] { return q->done; })) {
    q->claimed = false;
    return eg_fail(EG_ETIMEDOUT, "query %llu still running", (unsigned long long)ticket);
  }
  const eg_serve_result r = q->res;
  const std::string err = q->err;
  // by key: the wait released the lock, and a submit that rehashed the table
  // in the meantime invalidated `it`
  S->all.erase(ticket);
  lk.unlock();
  delete q;
  if (res) *res = r;
  if (r.status != EG_OK) eg_fail(r.status, "%s", err.c_str());
  return EG_OK;
}

extern "C" int eg_serve_stats_get(eg_server *S, eg_serve_stats *st) {
  if (!S || !st) return eg_fail(EG_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *st = S->stats;
  st->device_bytes = S->device_bytes;
  return EG_OK;
}

extern "C" int eg_serve_stop(eg_server *S) {
  if (!S) return EG_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  const int rc = S->fatal;
  const std::string msg = S->fatal_msg;
  destroy(S);
  if (rc != EG_OK) return eg_fail(rc, "server failed: %s", msg.c_str());
  return EG_OK;
}

// ---------------------------------------------------------------------------
// Poisson replay.
// ---------------------------------------------------------------------------
namespace {
uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double unit_open(uint64_t &s) { return ((splitmix(s) >> 11) + 1) * 0x1.0p-53; }  // (0, 1]

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)ceil(p * (double)v.size());
  if (k < 1) k = 1;
  return v[k - 1];
}
}  // namespace

extern "C" int eg_serve_replay(eg_graph *g, const eg_serve_opts *opts, const eg_replay_params *p,
                               eg_replay_report *r) {
  if (!p || !r) return eg_fail(EG_EINVAL, "null argument");
  EG_TRY(eg_require_loaded(g));
  if (!g->out_ptr) return eg_fail(EG_ENOTSUP, "serving needs adjacency lists (modes 2/3/4)");
  if (!(p->qps > 0.0) || !(p->bfs_fraction >= 0.0 && p->bfs_fraction <= 1.0))
    return eg_fail(EG_EINVAL, "bad replay parameters");
  const uint32_t max_seeds = p->ppr_max_seeds ? p->ppr_max_seeds : 8;
  std::vector<uint32_t> deg(g->n);
  EG_HIP(hipMemcpyAsync(deg.data(), g->out_deg, (size_t)g->n * 4, hipMemcpyDeviceToHost,
                        g->stream));
  EG_HIP(hipStreamSynchronize(g->stream));
  std::vector<uint32_t> cand;
  for (uint32_t v = 0; v < g->n; v++)
    if (deg[v]) cand.push_back(v);
  if (cand.empty()) return eg_fail(EG_EINVAL, "graph has no vertex with out-edges");
  // The stream: arrival offsets (ms), kinds, roots / seed lists.
  const uint32_t N = p->requests;
  std::vector<double> at(N);
  std::vector<int> kind(N);
  std::vector<std::vector<uint32_t>> who(N);
  uint64_t rs = p->seed;
  double t = 0.0;
  for (uint32_t i = 0; i < N; i++) {
    t += -log(unit_open(rs)) / p->qps * 1e3;
    at[i] = t;
    kind[i] = unit_open(rs) <= p->bfs_fraction ? EG_REQ_BFS : EG_REQ_PPR;
    const uint32_t cnt = kind[i] == EG_REQ_BFS ? 1 : 1 + (uint32_t)(splitmix(rs) % max_seeds);
    for (uint32_t j = 0; j < cnt; j++) who[i].push_back(cand[splitmix(rs) % cand.size()]);
  }
  eg_server *S = nullptr;
  EG_TRY(eg_serve_start(g, opts, &S));
  std::vector<uint64_t> tickets(N, 0);
  std::vector<double> sub(N, 0.0);
  const auto t0 = Clock::now();
  const double t0ms = now_ms();
  int rc = EG_OK;
  for (uint32_t i = 0; i < N && rc == EG_OK; i++) {
    std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double, std::milli>(at[i])));
    sub[i] = now_ms();
    rc = kind[i] == EG_REQ_BFS
             ? eg_serve_submit_bfs(S, who[i][0], nullptr, &tickets[i])
             : eg_serve_submit_ppr(S, who[i].data(), (uint32_t)who[i].size(), nullptr, &tickets[i]);
  }
  std::vector<double> lat, lat_b, lat_p;
  memset(r, 0, sizeof(*r));
  double last = t0ms;
  for (uint32_t i = 0; i < N; i++) {
    if (!tickets[i]) continue;
    eg_serve_result res{};
    const int w = eg_serve_wait(S, tickets[i], -1.0, &res);
    if (w != EG_OK || res.status != EG_OK) {
      r->failed++;
      continue;
    }
    r->completed++;
    lat.push_back(res.latency_ms);
    if (kind[i] == EG_REQ_BFS) {
      r->bfs++;
      lat_b.push_back(res.latency_ms);
    } else {
      r->ppr++;
      lat_p.push_back(res.latency_ms);
    }
    last = std::max(last, sub[i] + res.latency_ms);
  }
  eg_serve_stats_get(S, &r->stats);
  const int stop_rc = eg_serve_stop(S);
  if (rc != EG_OK) return rc;
  if (stop_rc != EG_OK) return stop_rc;
  r->wall_s = (last - t0ms) * 1e-3;
  r->qps = r->wall_s > 0.0 ? r->completed / r->wall_s : 0.0;
  r->p50_ms = percentile(lat, 0.50);
  r->p99_ms = percentile(lat, 0.99);
  double sum = 0.0, mx = 0.0;
  for (double x : lat) {
    sum += x;
    mx = std::max(mx, x);
  }
  r->mean_ms = lat.empty() ? 0.0 : sum / (double)lat.size();
  r->max_ms = mx;
  r->bfs_p50_ms = percentile(lat_b, 0.50);
  r->bfs_p99_ms = percentile(lat_b, 0.99);
  r->ppr_p50_ms = percentile(lat_p, 0.50);
  r->ppr_p99_ms = percentile(lat_p, 0.99);
  return EG_OK;
}
