// bk_tune.hip -- the autotuner behind bk_tune_sweep (host code).
//
// For one stencil and dtype on a periodic grid of given cell extents it
// builds the grid once per brick layout (8x8x8 in the 8x8 fold, 4x8x16 in
// the 4x16 and the 8x8 fold), times every schedule the library has for that
// layout -- one brick per wave with 1, 2, 4 or 8 bricks per workgroup,
// k-columns of 4 / 8 / 16 bricks and 2x2 column quads of 1 / 2 / 4 / 8
// levels for the stars, quads of 2 / 4 / 6 / 8 levels for the cube -- and returns them
// fastest first.  The reference
// has no tuner: its brick shape and fold are template arguments fixed at
// compile time (brick.h:353-395, stencils/stencils.h:10-12) and its GPU
// kernels one launch shape (3axis.hip.cpp:16-41).  Every candidate computes
// the same bits (fixed per-point evaluation order), so a choice only
// changes the speed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "bk_api.h"
#include "bk_internal.h"

namespace {

// errors go to the library's bk_last_error
int tfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return bk::plan_fail(code, buf);
}

// Propagates a C-ABI status (its message is already in bk_last_error).
#define BK_TRY(call)                \
  do {                              \
    int rc_ = (call);               \
    if (rc_ != BK_OK) return rc_;   \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)bk_free(p);
  }
  int alloc(size_t bytes) { return bk_alloc(bytes, &p); }
  int upload(const void *h, size_t bytes) { return bk_memcpy_h2d(p, h, bytes, nullptr); }
};

struct Plan {
  bk_plan p = nullptr;
  ~Plan() {
    if (p) (void)bk_plan_destroy(p);
  }
};

struct Event {
  bk_event e = nullptr;
  ~Event() {
    if (e) (void)bk_event_destroy(e);
  }
};

// the tuner's grid of one layout: tables in HBM and two fields
struct TuneGrid {
  long nb[3] = {0, 0, 0};
  long padded[3] = {0, 0, 0};
  uint32_t nbricks = 0;
  std::vector<uint32_t> adj, list;
  uint32_t nlist = 0, npairs = 0;
  DevBuf d_adj, d_list, d_pairs, f0, f1;
};

double seconds_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int build_grid(int layout, int dtype, const long cells[3], int ghost, TuneGrid &g) {
  int ext[3];
  BK_TRY(bk_layout_extents(layout, ext));
  g.nb[0] = cells[0] / ext[2];
  g.nb[1] = cells[1] / ext[1];
  g.nb[2] = cells[2] / ext[0];
  for (int a = 0; a < 3; ++a) g.padded[a] = g.nb[a] + 2 * ghost;
  const size_t nbr = (size_t)g.padded[0] * g.padded[1] * g.padded[2];
  if (nbr > 0xffffffffull) return tfail(BK_ERR_ARG, "bk_tune_sweep: grid too large");
  g.nbricks = (uint32_t)nbr;
  std::vector<uint32_t> grid(nbr);
  g.adj.resize(nbr * 27);
  BK_TRY(bk_init_grid(g.padded, ghost ? 0 : 1, grid.data(), g.adj.data()));
  // the sweep order of the library's grids: 32 x 16 (i, j) brick tiles
  uint32_t n = 0;
  BK_TRY(bk_tiled_list(g.nb, ghost, 32, 16, nullptr, &n));
  g.list.resize(n);
  BK_TRY(bk_tiled_list(g.nb, ghost, 32, 16, g.list.data(), &n));
  g.nlist = n;
  std::vector<uint32_t> pairs;
  if (ghost) {
    BK_TRY(bk_ghost_pairs(g.nb, nullptr, &g.npairs));
    pairs.resize(2 * (size_t)g.npairs);
    BK_TRY(bk_ghost_pairs(g.nb, pairs.data(), &g.npairs));
  }
  BK_TRY(g.d_adj.alloc(g.adj.size() * 4));
  BK_TRY(g.d_adj.upload(g.adj.data(), g.adj.size() * 4));
  BK_TRY(g.d_list.alloc((size_t)g.nlist * 4));
  BK_TRY(g.d_list.upload(g.list.data(), (size_t)g.nlist * 4));
  if (g.npairs) {
    BK_TRY(g.d_pairs.alloc(pairs.size() * 4));
    BK_TRY(g.d_pairs.upload(pairs.data(), pairs.size() * 4));
  }
  BK_TRY(bk_storage_alloc(dtype, (long)nbr, 512, &g.f0.p));
  BK_TRY(bk_storage_alloc(dtype, (long)nbr, 512, &g.f1.p));
  BK_TRY(bk_zero(g.f1.p, nbr * 512 * (size_t)bk_dtype_size(dtype), nullptr));
  BK_TRY(bk_fill_uniform_layout(layout, dtype, g.f0.p, 512, g.nb, ghost, 99, -1.0, 1.0,
                                nullptr));
  return bk_device_sync();
}

// the schedules of one layout, in a fixed order
void candidates_of(int layout, int stencil, unsigned flags, std::vector<bk_tune_cand> &out) {
  const bool batch_only = (flags & BK_TUNE_BATCH_SCHEDULES) != 0;
  auto add = [&](int shape, uint32_t kcol, int wg, int shift = 0) {
    bk_tune_cand c;
    std::memset(&c, 0, sizeof(c));
    c.layout = layout;
    c.plan_shape = shape;
    c.kcol = kcol;
    c.wg_bricks = wg;
    c.lane_shift = shift;
    c.ms = 0.0f;
    out.push_back(c);
  };
  if (layout != BK_LAYOUT_8x8x8) {
    if (batch_only) return;
    for (int wg : {1, 2, 4}) add(0, 0, wg);
    return;
  }
  if (stencil == BK_CUBE125) {
    if (!batch_only) add(0, 0, 4);
    for (uint32_t k : {2u, 4u, 6u, 8u}) add(BK_PLAN_QUAD, k, 0);
    return;
  }
  if (batch_only) {
    add(0, 0, 4);
  } else {
    for (int wg : {1, 2, 4, 8}) add(0, 0, wg);
  }
  for (uint32_t k : {4u, 8u, 16u}) add(BK_PLAN_COLUMN, k, 0);
  for (uint32_t k : {1u, 2u, 4u, 8u}) add(BK_PLAN_QUAD, k, 0);
  // i-taps by DPP lane shifts: single-grid quad kernels only
  if (!batch_only)
    for (uint32_t k : {1u, 2u, 4u, 8u}) add(BK_PLAN_QUAD, k, 0, 1);
}

struct TuneKey {
  int stencil, dtype, ghost;
  long c0, c1, c2;
  unsigned flags;
  bool operator<(const TuneKey &o) const {
    return std::tie(stencil, dtype, ghost, c0, c1, c2, flags) <
           std::tie(o.stencil, o.dtype, o.ghost, o.c0, o.c1, o.c2, o.flags);
  }
};

struct TuneCache {
  std::mutex mu;
  std::map<TuneKey, std::vector<bk_tune_cand>> results;
};
TuneCache &tune_cache() {
  static TuneCache c;
  return c;
}

}  // namespace

extern "C" {

int bk_tiled_list(const long nb[3], int ghost, uint32_t tile_i, uint32_t tile_j,
                  uint32_t *h_list, uint32_t *n) {
  if (!nb || !n) return tfail(BK_ERR_ARG, "bk_tiled_list: null pointer");
  if (ghost != 0 && ghost != 1) return tfail(BK_ERR_ARG, "bk_tiled_list: ghost must be 0 or 1");
  for (int a = 0; a < 3; ++a)
    if (nb[a] <= 0) return tfail(BK_ERR_ARG, "bk_tiled_list: nb must be positive");
  const uint64_t total = (uint64_t)nb[0] * nb[1] * nb[2];
  if (total > 0xffffffffull) return tfail(BK_ERR_ARG, "bk_tiled_list: too many bricks");
  *n = (uint32_t)total;
  if (!h_list) return BK_OK;
  const long Ti = tile_i ? (long)tile_i : nb[0], Tj = tile_j ? (long)tile_j : nb[1];
  const long S0 = nb[0] + 2 * ghost, S1 = nb[1] + 2 * ghost;
  uint32_t q = 0;
  for (long j0 = 0; j0 < nb[1]; j0 += Tj)
    for (long i0 = 0; i0 < nb[0]; i0 += Ti)
      for (long k = 0; k < nb[2]; ++k)
        for (long j = j0; j < std::min(j0 + Tj, nb[1]); ++j)
          for (long i = i0; i < std::min(i0 + Ti, nb[0]); ++i)
            h_list[q++] = (uint32_t)((i + ghost) + S0 * ((j + ghost) + S1 * (k + ghost)));
  return BK_OK;
}

int bk_tune_sweep(int stencil, int dtype, const double *h_coeff, double theta, int cond_mode,
                  const long cells[3], int ghost, int steps, int rounds, double warm_s,
                  unsigned flags, bk_tune_cand *h_out, uint32_t max_out, uint32_t *nout,
                  bk_stream s) {
  if (!cells || !nout) return tfail(BK_ERR_ARG, "bk_tune_sweep: null pointer");
  if (bk_stencil_ncoeff(stencil) == 0)
    return tfail(BK_ERR_ARG, "bk_tune_sweep: unknown stencil %d", stencil);
  if (bk_dtype_size(dtype) == 0) return tfail(BK_ERR_ARG, "bk_tune_sweep: unknown dtype %d", dtype);
  if (!h_coeff) return tfail(BK_ERR_ARG, "bk_tune_sweep: null coefficients");
  if (ghost != 0 && ghost != 1) return tfail(BK_ERR_ARG, "bk_tune_sweep: ghost must be 0 or 1");
  if (steps < 1 || rounds < 1 || warm_s < 0.0)
    return tfail(BK_ERR_ARG, "bk_tune_sweep: steps and rounds must be >= 1, warm_s >= 0");
  const unsigned lay_mask = flags & 7u;
  if (!lay_mask) return tfail(BK_ERR_ARG, "bk_tune_sweep: no layout selected");
  // the layouts the cell extents fit
  std::vector<int> layouts;
  for (int l = 0; l < 3; ++l) {
    if (!(lay_mask & (1u << l))) continue;
    int ext[3];
    BK_TRY(bk_layout_extents(l, ext));
    if (cells[0] <= 0 || cells[1] <= 0 || cells[2] <= 0 || cells[0] % ext[2] ||
        cells[1] % ext[1] || cells[2] % ext[0])
      continue;
    layouts.push_back(l);
  }
  std::vector<bk_tune_cand> all;
  for (int l : layouts) candidates_of(l, stencil, flags, all);
  if (all.empty())
    return tfail(BK_ERR_ARG, "bk_tune_sweep: the cell extents fit no selected layout");
  if (!h_out) {
    *nout = (uint32_t)all.size();
    return BK_OK;
  }

  bk_sweep_desc d;
  std::memset(&d, 0, sizeof(d));
  d.stencil = stencil;
  d.dtype = dtype;
  d.h_coeff = h_coeff;
  d.theta = theta;
  d.cond_mode = cond_mode;
  d.in_step = d.out_step = 512;
  Event e0, e1;
  BK_TRY(bk_event_create(&e0.e));
  BK_TRY(bk_event_create(&e1.e));
  std::vector<bk_tune_cand> timed;
  for (int l : layouts) {
    TuneGrid g;
    BK_TRY(build_grid(l, dtype, cells, ghost, g));
    std::vector<bk_tune_cand> mine;
    candidates_of(l, stencil, flags, mine);
    std::vector<Plan> plans(mine.size());
    for (size_t q = 0; q < mine.size(); ++q)
      if (mine[q].plan_shape)
        BK_TRY(bk_plan_create(g.adj.data(), g.nbricks, g.list.data(), g.nlist,
                              mine[q].plan_shape, mine[q].kcol, &plans[q].p));
    d.layout = l;
    d.d_adj = static_cast<const uint32_t *>(g.d_adj.p);
    d.d_list = static_cast<const uint32_t *>(g.d_list.p);
    d.nlist = g.nlist;
    d.d_in = g.f0.p;
    d.d_out = g.f1.p;
    auto run = [&](size_t q, int n) -> int {
      d.plan = plans[q].p;
      d.wg_bricks = mine[q].wg_bricks;
      d.lane_shift = mine[q].lane_shift;
      return bk_sweeps(&d, static_cast<const uint32_t *>(g.d_pairs.p), g.npairs, n, 1, s);
    };
    // untimed: one run each (graph capture), then warm_s seconds of all
    // candidates in turn -- the clock rises over the first fraction of a
    // second of sustained work, which would favour whichever runs last
    for (size_t q = 0; q < mine.size(); ++q) BK_TRY(run(q, steps));
    BK_TRY(bk_stream_sync(s));
    const double t0 = seconds_now();
    while (seconds_now() - t0 < warm_s) {
      for (size_t q = 0; q < mine.size(); ++q) BK_TRY(run(q, steps));
      BK_TRY(bk_stream_sync(s));
    }
    for (auto &c : mine) c.ms = 3.4e38f;
    for (int r = 0; r < rounds; ++r) {
      for (size_t q = 0; q < mine.size(); ++q) {
        BK_TRY(bk_event_record(e0.e, s));
        BK_TRY(run(q, steps));
        BK_TRY(bk_event_record(e1.e, s));
        BK_TRY(bk_event_sync(e1.e));
        float ms = 0.0f;
        BK_TRY(bk_event_elapsed_ms(e0.e, e1.e, &ms));
        mine[q].ms = std::min(mine[q].ms, ms / (float)steps);
      }
    }
    timed.insert(timed.end(), mine.begin(), mine.end());
  }
  std::stable_sort(timed.begin(), timed.end(),
                   [](const bk_tune_cand &a, const bk_tune_cand &b) { return a.ms < b.ms; });
  {
    auto &tc = tune_cache();
    std::lock_guard<std::mutex> lock(tc.mu);
    tc.results[TuneKey{stencil, dtype, ghost, cells[0], cells[1], cells[2], flags}] = timed;
  }
  const uint32_t n = std::min<uint32_t>(max_out, (uint32_t)timed.size());
  std::copy(timed.begin(), timed.begin() + n, h_out);
  *nout = n;
  return BK_OK;
}

int bk_tune_lookup(int stencil, int dtype, const long cells[3], int ghost, unsigned flags,
                   bk_tune_cand *best) {
  if (!cells || !best) return tfail(BK_ERR_ARG, "bk_tune_lookup: null pointer");
  auto &tc = tune_cache();
  std::lock_guard<std::mutex> lock(tc.mu);
  auto it = tc.results.find(TuneKey{stencil, dtype, ghost, cells[0], cells[1], cells[2], flags});
  if (it == tc.results.end() || it->second.empty())
    return tfail(BK_ERR_ARG, "bk_tune_lookup: not tuned");
  *best = it->second.front();
  return BK_OK;
}

}  // extern "C"
