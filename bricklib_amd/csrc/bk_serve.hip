// bk_serve.hip -- the serving loop behind bk_server_* (host code).
//
// Stencil solves of differing grid size, stencil, dtype and sweep count
// arrive over time.  The loop admits every arrived request whose two
// fields fit the HBM budget (fields from a field arena, resident until the
// request completes), runs all admitted requests together for a chunk of
// ticks through batched persistent kernels (bk_batch_run: one kernel per
// (stencil, dtype, schedule) group per tick over the concatenated bricks of
// every request in it, hipGraph-replayed), retires the completed ones
// (their memory returns to the budget at once) and admits new arrivals
// between chunks: a request that arrives while others are on the GPU joins
// the next chunk, so their work overlaps on the device.
//
// The admitted requests are split over one or two lanes, each a batch on
// its own stream, joined once per chunk: lane 0 holds the FP64 cubes and
// the other dtypes' stars, lane 1 the other cubes and the FP64 stars, so
// one lane's compute-bound cube ticks run beside the other lane's
// memory-bound star ticks.  With pipelining the next chunk is queued
// before the host waits for the current one.
//
// Schedules: with tuning on, each request sweeps by the schedule
// bk_tune_sweep found fastest (among those the batched kernels run) for
// its stencil, dtype and size class -- brick counts binned by powers of
// two, each class timed once on a cubic grid of that size -- otherwise by
// the fixed defaults (2x2 column quads of 4 levels for the cube; for the
// stars the configured star_plan).
//
// The reference has no serving loop; its closest precedent is the strong
// scaling driver's one launch over many subdomains (strong/main.cu:85-99).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "bk_api.h"
#include "bk_internal.h"

namespace {

int sfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return bk::plan_fail(code, buf);
}

#define BK_TRY(call)              \
  do {                            \
    int rc_ = (call);             \
    if (rc_ != BK_OK) return rc_; \
  } while (0)

using Clock = std::chrono::steady_clock;

// Tables shared by every request of one brick-grid shape: adjacency,
// interior list, ghost pairs and the plans requests of this shape sweep by.
struct Shape {
  long nb[3] = {0, 0, 0};
  int ghost = 0;
  uint32_t nbricks = 0, nlist = 0, npairs = 0;
  std::vector<uint32_t> adj, list;
  void *d_adj = nullptr, *d_list = nullptr, *d_pairs = nullptr;
  std::map<std::pair<int, uint32_t>, bk_plan> plans;  // (shape, kcol)
  ~Shape() {
    for (auto &p : plans) (void)bk_plan_destroy(p.second);
    for (void *p : {d_adj, d_list, d_pairs})
      if (p) (void)bk_free(p);
  }
  int build(const long n[3], int g) {
    for (int a = 0; a < 3; ++a) nb[a] = n[a];
    ghost = g;
    long padded[3];
    for (int a = 0; a < 3; ++a) padded[a] = nb[a] + 2 * ghost;
    nbricks = (uint32_t)(padded[0] * padded[1] * padded[2]);
    std::vector<uint32_t> grid(nbricks);
    adj.resize((size_t)nbricks * 27);
    BK_TRY(bk_init_grid(padded, ghost ? 0 : 1, grid.data(), adj.data()));
    if (ghost) {
      BK_TRY(bk_interior_list(nb, nullptr, &nlist));
      list.resize(nlist);
      BK_TRY(bk_interior_list(nb, list.data(), &nlist));
    } else {
      list = grid;
      nlist = nbricks;
    }
    BK_TRY(bk_alloc(adj.size() * 4, &d_adj));
    BK_TRY(bk_memcpy_h2d(d_adj, adj.data(), adj.size() * 4, nullptr));
    BK_TRY(bk_alloc((size_t)nlist * 4, &d_list));
    BK_TRY(bk_memcpy_h2d(d_list, list.data(), (size_t)nlist * 4, nullptr));
    if (ghost) {
      BK_TRY(bk_ghost_pairs(nb, nullptr, &npairs));
      std::vector<uint32_t> pairs(2 * (size_t)npairs);
      BK_TRY(bk_ghost_pairs(nb, pairs.data(), &npairs));
      if (npairs) {
        BK_TRY(bk_alloc(pairs.size() * 4, &d_pairs));
        BK_TRY(bk_memcpy_h2d(d_pairs, pairs.data(), pairs.size() * 4, nullptr));
      }
    }
    return BK_OK;
  }
  int plan(int shape, uint32_t kcol, bk_plan *out) {
    auto it = plans.find({shape, kcol});
    if (it != plans.end()) {
      *out = it->second;
      return BK_OK;
    }
    bk_plan p = nullptr;
    BK_TRY(bk_plan_create(adj.data(), nbricks, list.data(), nlist, shape, kcol, &p));
    plans[{shape, kcol}] = p;
    *out = p;
    return BK_OK;
  }
};

enum ReqState { kPending = 0, kActive = 1, kDone = 2, kReleased = 3 };

struct Req {
  uint32_t id = 0;
  bk_serve_req spec;
  std::vector<double> coeff;
  double arrival = 0.0, admitted = 0.0, finished = 0.0;
  int state = kPending;
  Shape *shape = nullptr;
  bk_plan plan = nullptr;        // the schedule chosen at admission
  bk_plan plan_light = nullptr;  // default quad schedules of several levels: half as many
  bk_plan plan_used = nullptr;   // the plan of its last chunk
  void *f0 = nullptr, *f1 = nullptr;
  size_t field_bytes = 0;
  uint32_t parity = 0, done = 0;
  int lane = 0;
  uint32_t remaining() const { return spec.steps - done; }
  uint64_t cells() const { return (uint64_t)spec.nb[0] * spec.nb[1] * spec.nb[2] * 512; }
};

struct Inflight {
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<Req *> done;
  bool queued = false;
};

}  // namespace

struct bk_server_s {
  bk_serve_config cfg;
  int device = 0;
  bk_arena arena = nullptr;
  bk_stream streams[2] = {nullptr, nullptr};
  bk_batch batches[2] = {nullptr, nullptr};
  std::map<std::tuple<long, long, long, int>, std::unique_ptr<Shape>> shapes;
  std::map<std::tuple<int, int, int>, bk_tune_cand> tuned;  // (class, stencil, dtype)
  std::vector<std::unique_ptr<Req>> reqs;  // by id
  std::vector<Req *> pending;              // submitted, not admitted (arrival order)
  std::vector<Req *> active;
  Inflight slots[2];
  int cur_slot = 0;
  double used = 0.0;
  Clock::time_point t0;
  bool realtime = true;
  // statistics since the last bk_server_reset_stats
  std::vector<double> latencies;
  double work = 0.0, peak = 0.0;
  uint64_t chunks = 0;
  double stat_t0 = 0.0, stat_t1 = 0.0;
  // threading
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool running = false, stop = false;
  int worker_rc = BK_OK;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  int lane_of(int stencil, int dtype) const {
    if (cfg.lanes == 1) return 0;
    return (stencil == BK_CUBE125) == (dtype == BK_F64) ? 0 : 1;
  }
};

namespace {

// size class of a grid: log2 of its brick count, rounded
int size_class(const long nb[3]) {
  const double n = (double)nb[0] * nb[1] * nb[2];
  return (int)std::lround(std::log2(std::max(1.0, n)));
}

int get_shape(bk_server_s *s, const long nb[3], Shape **out) {
  const auto key = std::make_tuple(nb[0], nb[1], nb[2], s->cfg.ghost);
  auto it = s->shapes.find(key);
  if (it != s->shapes.end()) {
    *out = it->second.get();
    return BK_OK;
  }
  std::unique_ptr<Shape> sh(new Shape());
  BK_TRY(sh->build(nb, s->cfg.ghost));
  *out = sh.get();
  s->shapes[key] = std::move(sh);
  return BK_OK;
}

// the schedule of a request: tuned per (size class, stencil, dtype) or the
// defaults
int schedule_of(bk_server_s *s, const Req &r, int *shape, uint32_t *kcol) {
  const int st = r.spec.stencil;
  if (s->cfg.tune) {
    const int c = size_class(r.spec.nb);
    const auto key = std::make_tuple(c, st, r.spec.dtype);
    bk_tune_cand pick;
    bool have = false;
    {
      // the map is read by bk_server_stats on other threads
      std::lock_guard<std::mutex> lock(s->mu);
      auto it = s->tuned.find(key);
      if (it != s->tuned.end()) {
        pick = it->second;
        have = true;
      }
    }
    if (!have) {
      // a cubic grid of about 2^c bricks, timed without holding the lock
      const long e = std::max(1L, std::lround(std::cbrt(std::ldexp(1.0, c))));
      const long cells[3] = {8 * e, 8 * e, 8 * e};
      bk_tune_cand best[16];
      uint32_t n = 0;
      BK_TRY(bk_tune_sweep(st, r.spec.dtype, r.coeff.data(), r.spec.theta, r.spec.cond_mode,
                           cells, s->cfg.ghost, 3, 2, 0.02,
                           BK_TUNE_LAYOUT_8x8x8 | BK_TUNE_BATCH_SCHEDULES, best, 16, &n,
                           s->streams[0]));
      if (n == 0) return sfail(BK_ERR_HIP, "bk_server: tuning returned no candidate");
      pick = best[0];
      std::lock_guard<std::mutex> lock(s->mu);
      s->tuned.emplace(key, pick);
    }
    *shape = pick.plan_shape;
    *kcol = pick.kcol;
    return BK_OK;
  }
  if (st == BK_CUBE125) {
    // quads of two levels (one while the load is light: quad_levels)
    *shape = BK_PLAN_QUAD;
    *kcol = 2;
  } else if (s->cfg.star_plan == BK_PLAN_QUAD) {
    // levels per unit as the single-grid defaults: FP64 radius 1-2 one
    // level, the 25-point two; FP32/BF16 two, the 25-point four
    *shape = BK_PLAN_QUAD;
    const bool f64 = r.spec.dtype == BK_F64;
    *kcol = st == BK_STAR25 ? (f64 ? 2 : 4) : (f64 ? 1 : 2);
  } else if (s->cfg.star_plan == BK_PLAN_COLUMN) {
    *shape = BK_PLAN_COLUMN;
    *kcol = 8;
  } else {
    *shape = 0;
    *kcol = 0;
  }
  return BK_OK;
}

int admit(bk_server_s *s, Req &r) {
  BK_TRY(get_shape(s, r.spec.nb, &r.shape));
  int shape = 0;
  uint32_t kcol = 0;
  BK_TRY(schedule_of(s, r, &shape, &kcol));
  r.plan = nullptr;
  if (shape) BK_TRY(r.shape->plan(shape, kcol, &r.plan));
  r.plan_light = nullptr;
  if (shape == BK_PLAN_QUAD && kcol > 1 && !s->cfg.tune)
    BK_TRY(r.shape->plan(BK_PLAN_QUAD, kcol / 2, &r.plan_light));
  r.lane = s->lane_of(r.spec.stencil, r.spec.dtype);
  BK_TRY(bk_arena_alloc(s->arena, r.field_bytes, &r.f0));
  int rc = bk_arena_alloc(s->arena, r.field_bytes, &r.f1);
  if (rc != BK_OK) {
    (void)bk_arena_free(s->arena, r.f0);
    r.f0 = nullptr;
    return rc;
  }
  // filled on its lane's stream, which runs all its ticks
  BK_TRY(bk_fill_uniform_grid(r.spec.dtype, r.f0, 512, r.spec.nb, s->cfg.ghost, r.spec.seed,
                              -1.0, 1.0, s->streams[r.lane]));
  return BK_OK;  // done and parity start at 0 (a request is admitted once)
}

void release_fields(bk_server_s *s, Req &r) {
  if (r.f0) (void)bk_arena_free(s->arena, r.f0);
  if (r.f1) (void)bk_arena_free(s->arena, r.f1);
  r.f0 = r.f1 = nullptr;
}

void retire(bk_server_s *s, std::vector<Req *> &done) {
  const double t = s->now();
  std::lock_guard<std::mutex> lock(s->mu);
  for (Req *r : done) {
    r->finished = t;
    r->state = kDone;
    if (!r->spec.keep) {
      release_fields(s, *r);
      r->state = kReleased;
    }
    s->used -= 2.0 * (double)r->field_bytes;
    s->latencies.push_back(r->finished - r->arrival);
    s->work += (double)r->cells() * r->spec.steps;
    s->stat_t1 = t;
  }
  done.clear();
  s->cv.notify_all();
}

// Requests with a default quad schedule of several levels run it while
// their lane's units of one (stencil, dtype) group fill that group's
// persistent kernel four times over, and units of half the levels below
// that: a tick of a light load lasts about one unit, so halving the unit
// halves the tick, while a heavy load spends less on halo planes with
// longer units.  Cubes of 2 and 1 levels: Poisson p99 29.6 -> 21.7 ms with
// one level throughout, saturating 76.3 % of the single-grid rates with two
// throughout against 74.9 % with one.
using GroupKey = std::pair<int, int>;  // (stencil, dtype)
void quad_levels(bk_server_s *s, int lane, std::map<GroupKey, bool> &light) {
  std::map<GroupKey, uint64_t> units;
  for (Req *r : s->active)
    if (r->lane == lane && r->plan_light)
      units[{r->spec.stencil, r->spec.dtype}] += bk::plan_view(r->plan).nunits;
  for (auto &u : units) {
    const int st = u.first.first, dt = u.first.second;
    const uint32_t cap = st == BK_CUBE125 ? bk::cube_batched_capacity(dt)
                                          : bk::star_quads_batched_capacity(st, dt);
    light[u.first] = u.second < 4ull * cap;
  }
}

// one chunk of ticks over the active requests, lane by lane
int run_chunk(bk_server_s *s, std::vector<Req *> &done) {
  uint32_t ticks = 0;
  for (Req *r : s->active) ticks = std::max(ticks, r->remaining());
  ticks = std::min(ticks, s->cfg.chunk);
  for (int l = 0; l < s->cfg.lanes; ++l) {
    std::map<GroupKey, bool> light;
    quad_levels(s, l, light);
    std::vector<bk_batch_req> recs;
    std::vector<Req *> mine;
    for (Req *r : s->active) {
      if (r->lane != l) continue;
      bk_batch_req q;
      std::memset(&q, 0, sizeof(q));
      q.stencil = r->spec.stencil;
      q.dtype = r->spec.dtype;
      q.d_adj = static_cast<const uint32_t *>(r->shape->d_adj);
      q.d_list = static_cast<const uint32_t *>(r->shape->d_list);
      q.nlist = r->shape->nlist;
      q.plan = r->plan_light && light[{r->spec.stencil, r->spec.dtype}] ? r->plan_light : r->plan;
      q.d_pairs = static_cast<const uint32_t *>(r->shape->d_pairs);
      q.npairs = r->shape->npairs;
      q.d_f0 = r->f0;
      q.d_f1 = r->f1;
      q.step = 512;
      q.h_coeff = r->coeff.data();
      q.theta = r->spec.theta;
      q.cond_mode = r->spec.cond_mode;
      q.parity = r->parity;
      q.ticks = std::min(ticks, r->remaining());
      recs.push_back(q);
      mine.push_back(r);
    }
    if (recs.empty()) continue;
    BK_TRY(bk_batch_set(s->batches[l], recs.data(), (uint32_t)recs.size(), s->streams[l]));
    BK_TRY(bk_batch_run(s->batches[l], ticks, s->cfg.use_graph, s->streams[l]));
    // read by bk_server_info on other threads
    std::lock_guard<std::mutex> lock(s->mu);
    for (size_t q = 0; q < mine.size(); ++q) {
      mine[q]->done += recs[q].ticks;
      mine[q]->parity ^= recs[q].ticks & 1u;
      mine[q]->plan_used = recs[q].plan;
    }
  }
  s->chunks += 1;
  std::vector<Req *> still;
  for (Req *r : s->active) (r->remaining() == 0 ? done : still).push_back(r);
  s->active.swap(still);
  return BK_OK;
}

int sync_slot(bk_server_s *s, Inflight &f) {
  for (int l = 0; l < s->cfg.lanes; ++l) {
    hipError_t e = hipEventSynchronize(f.ev[l]);
    if (e != hipSuccess) return sfail(BK_ERR_HIP, "bk_server: %s", hipGetErrorString(e));
  }
  f.queued = false;
  retire(s, f.done);
  return BK_OK;
}

// One pass of the loop: admit, queue a chunk, retire the chunk before it.
// *idle: nothing is active or in flight (the caller may wait for arrivals).
int step(bk_server_s *s, bool *idle) {
  *idle = false;
  std::vector<Req *> admit_now;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    const double t = s->now();
    size_t q = 0;
    double used = s->used;
    size_t nact = s->active.size();
    while (q < s->pending.size()) {
      Req *r = s->pending[q];
      if ((s->realtime ? r->arrival : 0.0) > t) break;
      if (nact >= s->cfg.max_active || used + 2.0 * (double)r->field_bytes > s->cfg.hbm_budget)
        break;
      used += 2.0 * (double)r->field_bytes;
      ++nact;
      admit_now.push_back(r);
      ++q;
    }
    s->pending.erase(s->pending.begin(), s->pending.begin() + (long)q);
    s->used = used;
    s->peak = std::max(s->peak, used);
  }
  for (Req *r : admit_now) {
    BK_TRY(admit(s, *r));  // fields, plan and fill, outside the lock
    std::lock_guard<std::mutex> lock(s->mu);
    r->admitted = s->now();
    r->state = kActive;  // publishes what admit wrote to bk_server_info's readers
    s->active.push_back(r);
  }
  Inflight &prev = s->slots[s->cur_slot ^ 1];
  if (s->active.empty() && !prev.queued) {
    *idle = true;
    return BK_OK;
  }
  Inflight &cur = s->slots[s->cur_slot];
  if (!s->active.empty()) {
    BK_TRY(run_chunk(s, cur.done));
    if (s->cfg.pipeline) {
      for (int l = 0; l < s->cfg.lanes; ++l) {
        hipError_t e = hipEventRecord(cur.ev[l], static_cast<hipStream_t>(s->streams[l]));
        if (e != hipSuccess) return sfail(BK_ERR_HIP, "bk_server: %s", hipGetErrorString(e));
      }
      cur.queued = true;
    } else {
      for (int l = 0; l < s->cfg.lanes; ++l) BK_TRY(bk_stream_sync(s->streams[l]));
      retire(s, cur.done);
    }
  }
  // the previous chunk, queued before this one: its requests are complete
  // once its events are
  if (prev.queued) BK_TRY(sync_slot(s, prev));
  s->cur_slot ^= 1;
  return BK_OK;
}

bool all_done(bk_server_s *s) {
  std::lock_guard<std::mutex> lock(s->mu);
  return s->pending.empty() && s->active.empty() && !s->slots[0].queued && !s->slots[1].queued;
}

// wait until the next pending arrival (or a submission)
void wait_arrival(bk_server_s *s) {
  std::unique_lock<std::mutex> lock(s->mu);
  if (s->stop) return;
  if (s->pending.empty()) {
    s->cv.wait_for(lock, std::chrono::milliseconds(50));
    return;
  }
  const double dt = (s->realtime ? s->pending.front()->arrival : 0.0) - s->now();
  if (dt > 0) s->cv.wait_for(lock, std::chrono::duration<double>(std::min(dt, 0.05)));
}

void worker_main(bk_server_s *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    {
      std::lock_guard<std::mutex> lock(s->mu);
      if (s->stop && s->pending.empty() && s->active.empty() && !s->slots[0].queued &&
          !s->slots[1].queued)
        break;
    }
    bool idle = false;
    int rc = step(s, &idle);
    if (rc != BK_OK) {
      std::lock_guard<std::mutex> lock(s->mu);
      s->worker_rc = rc;
      s->cv.notify_all();
      break;
    }
    if (idle) wait_arrival(s);
  }
  std::lock_guard<std::mutex> lock(s->mu);
  s->running = false;
  s->cv.notify_all();
}

int check_spec(const bk_serve_req *q) {
  if (!q) return sfail(BK_ERR_ARG, "bk_server_submit: null request");
  if (bk_stencil_ncoeff(q->stencil) == 0)
    return sfail(BK_ERR_ARG, "bk_server_submit: unknown stencil %d", q->stencil);
  if (bk_dtype_size(q->dtype) == 0)
    return sfail(BK_ERR_ARG, "bk_server_submit: unknown dtype %d", q->dtype);
  if (!q->h_coeff) return sfail(BK_ERR_ARG, "bk_server_submit: null coefficients");
  for (int a = 0; a < 3; ++a)
    if (q->nb[a] <= 0 || q->nb[a] > 4096)
      return sfail(BK_ERR_ARG, "bk_server_submit: nb must be in [1, 4096]");
  if (q->steps == 0) return sfail(BK_ERR_ARG, "bk_server_submit: zero sweeps");
  if (q->stencil == BK_COND7 && q->cond_mode != BK_COND_REFERENCE &&
      q->cond_mode != BK_COND_THRESHOLD)
    return sfail(BK_ERR_ARG, "bk_server_submit: unknown cond mode %d", q->cond_mode);
  return BK_OK;
}

}  // namespace

extern "C" {

int bk_server_create(const bk_serve_config *cfg, bk_server *out) {
  if (!cfg || !out) return sfail(BK_ERR_ARG, "bk_server_create: null pointer");
  if (cfg->lanes != 1 && cfg->lanes != 2)
    return sfail(BK_ERR_ARG, "bk_server_create: lanes must be 1 or 2");
  if (cfg->chunk == 0) return sfail(BK_ERR_ARG, "bk_server_create: chunk must be positive");
  if (cfg->max_active == 0 || cfg->max_active > 2048)
    return sfail(BK_ERR_ARG, "bk_server_create: max_active must be in [1, 2048]");
  if (cfg->ghost != 0 && cfg->ghost != 1)
    return sfail(BK_ERR_ARG, "bk_server_create: ghost must be 0 or 1");
  if (!(cfg->hbm_budget > 0.0)) return sfail(BK_ERR_ARG, "bk_server_create: budget must be > 0");
  if (cfg->star_plan != 0 && cfg->star_plan != BK_PLAN_QUAD && cfg->star_plan != BK_PLAN_COLUMN)
    return sfail(BK_ERR_ARG, "bk_server_create: star_plan must be 0, BK_PLAN_QUAD or "
                 "BK_PLAN_COLUMN");
  std::unique_ptr<bk_server_s> s(new bk_server_s());
  s->cfg = *cfg;
  if (hipGetDevice(&s->device) != hipSuccess)
    return sfail(BK_ERR_HIP, "bk_server_create: no current device");
  BK_TRY(bk_arena_create(cfg->slab_bytes ? cfg->slab_bytes : ((size_t)1 << 30), 0, &s->arena));
  for (int l = 0; l < cfg->lanes; ++l) {
    BK_TRY(bk_stream_create(&s->streams[l]));
    BK_TRY(bk_batch_create(cfg->max_active, &s->batches[l]));
  }
  for (auto &f : s->slots)
    for (int l = 0; l < cfg->lanes; ++l)
      if (hipEventCreateWithFlags(&f.ev[l], hipEventDisableTiming) != hipSuccess)
        return sfail(BK_ERR_HIP, "bk_server_create: event creation failed");
  s->t0 = Clock::now();
  *out = s.release();
  return BK_OK;
}

int bk_server_destroy(bk_server s) {
  if (!s) return BK_OK;
  (void)bk_server_stop(s);
  for (int l = 0; l < s->cfg.lanes; ++l)
    if (s->streams[l]) (void)bk_stream_sync(s->streams[l]);
  for (auto &r : s->reqs) release_fields(s, *r);
  for (int l = 0; l < 2; ++l) {
    if (s->batches[l]) (void)bk_batch_destroy(s->batches[l]);
    if (s->streams[l]) (void)bk_stream_destroy(s->streams[l]);
  }
  for (auto &f : s->slots)
    for (hipEvent_t e : f.ev)
      if (e) (void)hipEventDestroy(e);
  s->shapes.clear();
  if (s->arena) (void)bk_arena_destroy(s->arena);
  delete s;
  return BK_OK;
}

int bk_server_submit(bk_server s, const bk_serve_req *q, uint32_t *id) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_submit: null server");
  BK_TRY(check_spec(q));
  std::unique_ptr<Req> r(new Req());
  r->spec = *q;
  r->coeff.assign(q->h_coeff, q->h_coeff + bk_stencil_ncoeff(q->stencil));
  r->spec.h_coeff = nullptr;
  const long g = s->cfg.ghost;
  r->field_bytes = (size_t)(q->nb[0] + 2 * g) * (q->nb[1] + 2 * g) * (q->nb[2] + 2 * g) * 512 *
                   (size_t)bk_dtype_size(q->dtype);
  if (2.0 * (double)r->field_bytes > s->cfg.hbm_budget)
    return sfail(BK_ERR_ARG, "bk_server_submit: a request of %zu bytes exceeds the budget",
                 2 * r->field_bytes);
  std::lock_guard<std::mutex> lock(s->mu);
  r->id = (uint32_t)s->reqs.size();
  r->arrival = q->arrival >= 0.0 ? q->arrival : s->now();
  // pending stays in arrival order
  auto pos = std::upper_bound(s->pending.begin(), s->pending.end(), r->arrival,
                              [](double t, const Req *p) { return t < p->arrival; });
  s->pending.insert(pos, r.get());
  if (id) *id = r->id;
  s->reqs.push_back(std::move(r));
  s->cv.notify_all();
  return BK_OK;
}

int bk_server_prewarm(bk_server s, const bk_serve_req *qs, uint32_t n) {
  if (!s || (n && !qs)) return sfail(BK_ERR_ARG, "bk_server_prewarm: null pointer");
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->running) return sfail(BK_ERR_ARG, "bk_server_prewarm: the server is running");
  }
  // shapes and plans are the loop's own tables: no thread serves meanwhile
  // (the caller's contract), so only the tuned map takes the lock
  for (uint32_t q = 0; q < n; ++q) {
    BK_TRY(check_spec(&qs[q]));
    Req r;
    r.spec = qs[q];
    r.coeff.assign(qs[q].h_coeff, qs[q].h_coeff + bk_stencil_ncoeff(qs[q].stencil));
    BK_TRY(get_shape(s, r.spec.nb, &r.shape));
    int shape = 0;
    uint32_t kcol = 0;
    BK_TRY(schedule_of(s, r, &shape, &kcol));
    bk_plan p = nullptr;
    if (shape) BK_TRY(r.shape->plan(shape, kcol, &p));
    if (shape == BK_PLAN_QUAD && kcol > 1 && !s->cfg.tune)
      BK_TRY(r.shape->plan(BK_PLAN_QUAD, kcol / 2, &p));
  }
  return bk_device_sync();
}

int bk_server_run(bk_server s, int realtime) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_run: null server");
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->running) return sfail(BK_ERR_ARG, "bk_server_run: the server's thread is running");
    s->realtime = realtime != 0;
    s->t0 = Clock::now();
    s->stat_t0 = 0.0;
    s->stat_t1 = 0.0;
    // a saturating run: every queued request is present from the start
    if (!s->realtime)
      for (Req *r : s->pending) r->arrival = 0.0;
  }
  while (!all_done(s)) {
    bool idle = false;
    BK_TRY(step(s, &idle));
    if (idle) wait_arrival(s);
  }
  return BK_OK;
}

int bk_server_start(bk_server s) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_start: null server");
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->running) return BK_OK;
  s->realtime = true;
  s->stop = false;
  s->worker_rc = BK_OK;
  s->running = true;
  s->worker = std::thread(worker_main, s);
  return BK_OK;
}

int bk_server_stop(bk_server s) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_stop: null server");
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  std::lock_guard<std::mutex> lock(s->mu);
  s->stop = false;
  return s->worker_rc;
}

int bk_server_wait(bk_server s, uint32_t id, double timeout_s) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_wait: null server");
  std::unique_lock<std::mutex> lock(s->mu);
  if (id >= s->reqs.size()) return sfail(BK_ERR_ARG, "bk_server_wait: unknown request %u", id);
  const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(timeout_s));
  while (s->reqs[id]->state < kDone) {
    if (s->worker_rc != BK_OK) return s->worker_rc;
    if (!s->running) return sfail(BK_ERR_ARG, "bk_server_wait: the server is not running");
    if (s->cv.wait_until(lock, deadline) == std::cv_status::timeout &&
        s->reqs[id]->state < kDone)
      return sfail(BK_ERR_ARG, "bk_server_wait: request %u not done within %.3f s", id,
                   timeout_s);
  }
  return BK_OK;
}

int bk_server_info(bk_server s, uint32_t id, bk_serve_info *out) {
  if (!s || !out) return sfail(BK_ERR_ARG, "bk_server_info: null pointer");
  std::lock_guard<std::mutex> lock(s->mu);
  if (id >= s->reqs.size()) return sfail(BK_ERR_ARG, "bk_server_info: unknown request %u", id);
  const Req &r = *s->reqs[id];
  std::memset(out, 0, sizeof(*out));
  out->state = r.state;
  out->arrival = r.arrival;
  out->admitted = r.admitted;
  out->finished = r.finished;
  out->sweeps_done = r.done;
  out->field_bytes = r.field_bytes;
  out->d_result = r.state == kDone ? (r.parity ? r.f1 : r.f0) : nullptr;
  // the plan is written by admission before the state leaves kPending
  const bk_plan p = r.plan_used ? r.plan_used : r.plan;
  const bool admitted = r.state >= kActive && p;
  out->plan_shape = admitted ? bk::plan_view(p).shape : 0;
  out->kcol = admitted ? bk::plan_view(p).kcol : 0;
  return BK_OK;
}

int bk_server_result(bk_server s, uint32_t id, void *h_out, size_t bytes) {
  if (!s || !h_out) return sfail(BK_ERR_ARG, "bk_server_result: null pointer");
  const void *src = nullptr;
  size_t fb = 0;
  bk_stream st = nullptr;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (id >= s->reqs.size()) return sfail(BK_ERR_ARG, "bk_server_result: unknown request %u", id);
    const Req &r = *s->reqs[id];
    if (r.state != kDone)
      return sfail(BK_ERR_ARG, "bk_server_result: request %u is not complete with its fields "
                   "kept", id);
    src = r.parity ? r.f1 : r.f0;
    fb = r.field_bytes;
    st = s->streams[r.lane];
  }
  if (bytes != fb)
    return sfail(BK_ERR_ARG, "bk_server_result: the field holds %zu bytes, not %zu", fb, bytes);
  return bk_memcpy_d2h(h_out, src, fb, st);
}

int bk_server_release(bk_server s, uint32_t id) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_release: null server");
  std::lock_guard<std::mutex> lock(s->mu);
  if (id >= s->reqs.size()) return sfail(BK_ERR_ARG, "bk_server_release: unknown request %u", id);
  Req &r = *s->reqs[id];
  if (r.state == kDone) {
    for (int l = 0; l < s->cfg.lanes; ++l) BK_TRY(bk_stream_sync(s->streams[l]));
    release_fields(s, r);
    r.state = kReleased;
  }
  return BK_OK;
}

int bk_server_stats(bk_server s, bk_serve_stats *out) {
  if (!s || !out) return sfail(BK_ERR_ARG, "bk_server_stats: null pointer");
  std::lock_guard<std::mutex> lock(s->mu);
  std::memset(out, 0, sizeof(*out));
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) -> double {
    if (l.empty()) return 0.0;
    // linear interpolation between closest ranks (numpy's default)
    const double x = p / 100.0 * (double)(l.size() - 1);
    const size_t lo = (size_t)std::floor(x), hi = std::min(lo + 1, l.size() - 1);
    return l[lo] + (l[hi] - l[lo]) * (x - (double)lo);
  };
  out->requests = (uint32_t)l.size();
  out->p50_s = pct(50.0);
  out->p99_s = pct(99.0);
  double sum = 0.0;
  for (double v : l) sum += v;
  out->mean_s = l.empty() ? 0.0 : sum / (double)l.size();
  out->elapsed_s = s->stat_t1 - s->stat_t0;
  out->stencil_updates = s->work;
  out->gstencil_s = out->elapsed_s > 0 ? s->work / out->elapsed_s / 1e9 : 0.0;
  out->peak_hbm_bytes = s->peak;
  out->chunks = s->chunks;
  size_t used = 0, reserved = 0, largest = 0;
  uint32_t live = 0;
  (void)bk_arena_stats(s->arena, &used, &reserved, &largest, &live);
  out->arena_reserved_bytes = (double)reserved;
  out->tuned_classes = (uint32_t)s->tuned.size();
  return BK_OK;
}

int bk_server_reset_stats(bk_server s) {
  if (!s) return sfail(BK_ERR_ARG, "bk_server_reset_stats: null server");
  std::lock_guard<std::mutex> lock(s->mu);
  s->latencies.clear();
  s->work = 0.0;
  s->peak = s->used;
  s->chunks = 0;
  s->stat_t0 = s->stat_t1 = s->now();
  return BK_OK;
}

}  // extern "C"
