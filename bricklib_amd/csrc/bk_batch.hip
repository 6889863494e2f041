// bk_batch.hip -- batched sweeps over many independent grids.
//
// A batch holds one device table per (stencil, dtype) group and one per
// dtype for the ghost refresh.  bk_batch_set() writes the current requests
// into those tables (pinned staging, then stream-ordered copies into HBM);
// bk_batch_run() launches, per tick, one persistent refresh kernel per dtype
// and one persistent sweep kernel per group -- each serving the
// concatenated work of all its requests -- optionally captured once into a
// hipGraph and replayed.  The launch grids do not depend on the table
// contents, so a captured run stays valid when the host rewrites the
// tables between runs (continuous batching).
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "bk_api.h"
#include "bk_graphs.h"
#include "bk_internal.h"

namespace bk {
int plan_fail(int code, const char *msg);
}

namespace {
// sweep groups: (stencil, dtype, schedule); the schedule of a star request
// is 0 without a plan (one brick per wave), 1 with a column plan (k-marching
// units), 2 with a quad plan (2x2 blocks of k-columns per workgroup)
constexpr int kStencils = 5, kDtypes = 3, kScheds = 3, kGroups = kStencils * kDtypes * kScheds;
static_assert(kGroups <= 64, "the non-empty groups form a 64-bit mask");
constexpr size_t kCoeffSlot = 125;  // doubles per request in the coefficient pool

constexpr int group_of(int stencil, int dtype, int sched) {
  return (stencil * kDtypes + dtype) * kScheds + sched;
}
}  // namespace

struct bk_batch_s {
  int device = 0;
  uint32_t max_reqs = 0;
  bk::BatchTable *d_sweep[kGroups] = {};
  bk::BatchTable *h_sweep[kGroups] = {};
  bk::BatchTable *d_ref[kDtypes] = {};
  bk::BatchTable *h_ref[kDtypes] = {};
  double *d_coeff = nullptr;
  double *h_coeff = nullptr;
  uint64_t sweep_mask = 0;  // non-empty sweep groups
  uint32_t ref_mask = 0;    // non-empty refresh tables
  hipEvent_t staged = nullptr;
  hipStream_t capture = nullptr;
  // captured runs keyed by (tick count, sweep and refresh group masks); an
  // exec is destroyed only once its last launch has completed, so a caller
  // may queue the next run while the previous one is still on the device
  bk::ExecCache<std::tuple<uint32_t, uint64_t, uint32_t>> graphs{64};
  std::mutex mu;
};

namespace {

int fail(int code, const char *msg) { return bk::plan_fail(code, msg); }

int hipfail(hipError_t e, const char *what) {
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  return fail(e == hipErrorOutOfMemory ? BK_ERR_NOMEM : BK_ERR_HIP, buf);
}

#define BKB_HIP(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hipfail(e_, #call);  \
  } while (0)

size_t table_head_bytes() { return offsetof(bk::BatchTable, seg); }

hipError_t enqueue_run(bk_batch_s *b, uint32_t nticks, hipStream_t s) {
  for (uint32_t t = 0; t < nticks; ++t) {
    for (int d = 0; d < kDtypes; ++d)
      if (b->ref_mask >> d & 1u) {
        hipError_t e = bk::launch_ghost_batched(d, b->d_ref[d], t, s);
        if (e != hipSuccess) return e;
      }
    for (int g = 0; g < kGroups; ++g)
      if (b->sweep_mask >> g & 1u) {
        const int stencil = g / (kScheds * kDtypes), dtype = (g / kScheds) % kDtypes;
        const int sched = g % kScheds;
        hipError_t e =
            stencil == BK_CUBE125
                ? bk::launch_cube_batched(dtype, b->d_sweep[g], t, s)
                : sched == 1 ? bk::launch_star_columns_batched(stencil, dtype, b->d_sweep[g], t, s)
                : sched == 2 ? bk::launch_star_quads_batched(stencil, dtype, b->d_sweep[g], t, s)
                             : bk::launch_star_batched(stencil, dtype, b->d_sweep[g], t, s);
        if (e != hipSuccess) return e;
      }
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int bk_batch_destroy(bk_batch b) {
  if (!b) return BK_OK;
  (void)b->graphs.clear();
  for (int g = 0; g < kGroups; ++g) {
    (void)hipFree(b->d_sweep[g]);
    (void)hipHostFree(b->h_sweep[g]);
  }
  for (int d = 0; d < kDtypes; ++d) {
    (void)hipFree(b->d_ref[d]);
    (void)hipHostFree(b->h_ref[d]);
  }
  (void)hipFree(b->d_coeff);
  (void)hipHostFree(b->h_coeff);
  if (b->staged) (void)hipEventDestroy(b->staged);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  delete b;
  return BK_OK;
}

int bk_batch_create(uint32_t max_reqs, bk_batch *out) {
  if (!out) return fail(BK_ERR_ARG, "bk_batch_create: null out pointer");
  *out = nullptr;
  if (max_reqs == 0 || max_reqs > bk::kBatchMaxSegs)
    return fail(BK_ERR_ARG, "bk_batch_create: max_reqs must be in [1, 2048]");
  bk_batch_s *b = new bk_batch_s();
  b->max_reqs = max_reqs;
  (void)hipGetDevice(&b->device);
  auto cleanup = [&](int rc) {
    bk_batch_destroy(b);
    return rc;
  };
  const size_t tb = sizeof(bk::BatchTable);
  for (int g = 0; g < kGroups; ++g) {
    if (hipMalloc(&b->d_sweep[g], tb) != hipSuccess ||
        hipHostMalloc(&b->h_sweep[g], tb, hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(BK_ERR_NOMEM, "bk_batch_create: table allocation failed"));
  }
  for (int d = 0; d < kDtypes; ++d) {
    if (hipMalloc(&b->d_ref[d], tb) != hipSuccess ||
        hipHostMalloc(&b->h_ref[d], tb, hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(BK_ERR_NOMEM, "bk_batch_create: table allocation failed"));
  }
  const size_t cb = sizeof(double) * kCoeffSlot * max_reqs;
  if (hipMalloc(&b->d_coeff, cb) != hipSuccess ||
      hipHostMalloc(&b->h_coeff, cb, hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(BK_ERR_NOMEM, "bk_batch_create: coefficient pool allocation failed"));
  if (hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(BK_ERR_HIP, "bk_batch_create: event/stream creation failed"));
  // empty tables on the device, so a run before the first set is a no-op
  for (int g = 0; g < kGroups; ++g) std::memset(b->h_sweep[g], 0, table_head_bytes());
  for (int d = 0; d < kDtypes; ++d) std::memset(b->h_ref[d], 0, table_head_bytes());
  for (int g = 0; g < kGroups; ++g)
    if (hipMemcpy(b->d_sweep[g], b->h_sweep[g], table_head_bytes(), hipMemcpyHostToDevice) !=
        hipSuccess)
      return cleanup(fail(BK_ERR_HIP, "bk_batch_create: table upload failed"));
  for (int d = 0; d < kDtypes; ++d)
    if (hipMemcpy(b->d_ref[d], b->h_ref[d], table_head_bytes(), hipMemcpyHostToDevice) !=
        hipSuccess)
      return cleanup(fail(BK_ERR_HIP, "bk_batch_create: table upload failed"));
  *out = b;
  return BK_OK;
}

int bk_batch_set(bk_batch b, const bk_batch_req *reqs, uint32_t n, bk_stream stream) {
  if (!b) return fail(BK_ERR_ARG, "bk_batch_set: null batch");
  if (n && !reqs) return fail(BK_ERR_ARG, "bk_batch_set: null request array");
  if (n > b->max_reqs) return fail(BK_ERR_ARG, "bk_batch_set: more requests than max_reqs");
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(b->mu);
  // the staging buffers may still feed the previous set's copies
  BKB_HIP(hipEventSynchronize(b->staged));
  // validate everything before touching the tables
  for (uint32_t r = 0; r < n; ++r) {
    const bk_batch_req &q = reqs[r];
    if (bk_stencil_ncoeff(q.stencil) == 0) return fail(BK_ERR_ARG, "bk_batch_set: unknown stencil");
    if (q.dtype < 0 || q.dtype >= kDtypes) return fail(BK_ERR_ARG, "bk_batch_set: unknown dtype");
    if (!q.d_adj || !q.d_f0 || !q.d_f1 || !q.h_coeff)
      return fail(BK_ERR_ARG, "bk_batch_set: null pointer in a request");
    if (q.d_f0 == q.d_f1) return fail(BK_ERR_ARG, "bk_batch_set: a request's fields alias");
    if (q.step < 512) return fail(BK_ERR_ARG, "bk_batch_set: step smaller than a brick");
    if (q.npairs && !q.d_pairs) return fail(BK_ERR_ARG, "bk_batch_set: null pair list");
    if (q.stencil == BK_CUBE125) {
      int shape = 0;
      if (!q.plan || bk_plan_query(q.plan, nullptr, nullptr, &shape, nullptr) != BK_OK ||
          shape != BK_PLAN_QUAD)
        return fail(BK_ERR_ARG, "bk_batch_set: a cube request needs a BK_PLAN_QUAD plan");
      if (!bk::quad_offsets_fit(bk::plan_view(q.plan).nbricks, q.step))
        return fail(BK_ERR_UNSUPPORTED,
                    "bk_batch_set: the quad schedule addresses at most 2^32 input elements");
    } else if (q.plan) {
      int shape = 0;
      if (bk_plan_query(q.plan, nullptr, nullptr, &shape, nullptr) != BK_OK ||
          (shape != BK_PLAN_COLUMN && shape != BK_PLAN_QUAD))
        return fail(BK_ERR_ARG, "bk_batch_set: a star request's plan must be a BK_PLAN_COLUMN or "
                    "BK_PLAN_QUAD plan");
    } else if (q.nlist && !q.d_list) {
      return fail(BK_ERR_ARG, "bk_batch_set: null brick list");
    }
    if (q.stencil == BK_COND7 && q.cond_mode != BK_COND_REFERENCE &&
        q.cond_mode != BK_COND_THRESHOLD)
      return fail(BK_ERR_ARG, "bk_batch_set: unknown cond mode");
  }
  uint32_t count[kGroups] = {}, rcount[kDtypes] = {};
  for (int g = 0; g < kGroups; ++g) b->h_sweep[g]->prefix[0] = 0;
  for (int d = 0; d < kDtypes; ++d) b->h_ref[d]->prefix[0] = 0;
  for (uint32_t r = 0; r < n; ++r) {
    const bk_batch_req &q = reqs[r];
    const int sched = (q.stencil == BK_CUBE125 || !q.plan) ? 0
                      : bk::plan_view(q.plan).shape == BK_PLAN_QUAD ? 2 : 1;
    const int g = group_of(q.stencil, q.dtype, sched);
    double *slot = b->h_coeff + (size_t)r * kCoeffSlot;
    const int nc = bk_stencil_ncoeff(q.stencil);
    if (q.dtype == BK_F64) {
      for (int c = 0; c < nc; ++c) slot[c] = q.h_coeff[c];
    } else {
      float *fs = reinterpret_cast<float *>(slot);
      for (int c = 0; c < nc; ++c) fs[c] = (float)q.h_coeff[c];
    }
    bk::BatchSeg sg;
    std::memset(&sg, 0, sizeof(sg));
    sg.adj = q.d_adj;
    sg.f0 = q.d_f0;
    sg.f1 = q.d_f1;
    sg.coeff = b->d_coeff + (size_t)r * kCoeffSlot;
    sg.step = q.step;
    sg.parity = q.parity & 1u;
    sg.ticks = q.ticks;
    sg.theta = q.theta;
    sg.cond_mode = q.cond_mode;
    if (q.stencil == BK_CUBE125 || sched != 0) {
      const bk::PlanView P = bk::plan_view(q.plan);
      sg.items = P.anchor;
      sg.len = P.len;
      sg.mask = P.mask;
      sg.nitems = P.nunits;
      sg.kcol = P.kcol;
    } else {
      sg.items = q.d_list;
      sg.nitems = q.nlist;
    }
    bk::BatchTable *t = b->h_sweep[g];
    t->seg[count[g]] = sg;
    t->prefix[count[g] + 1] = t->prefix[count[g]] + sg.nitems;
    ++count[g];
    if (q.npairs) {
      bk::BatchTable *rt = b->h_ref[q.dtype];
      bk::BatchSeg rs = sg;
      rs.items = q.d_pairs;
      rs.len = rs.mask = nullptr;
      rs.nitems = q.npairs;
      rt->seg[rcount[q.dtype]] = rs;
      rt->prefix[rcount[q.dtype] + 1] = rt->prefix[rcount[q.dtype]] + q.npairs;
      ++rcount[q.dtype];
    }
  }
  b->sweep_mask = 0;
  b->ref_mask = 0;
  for (int g = 0; g < kGroups; ++g) {
    bk::BatchTable *t = b->h_sweep[g];
    t->nsegs = count[g];
    if (count[g] && t->prefix[count[g]]) b->sweep_mask |= 1ull << g;
    BKB_HIP(hipMemcpyAsync(b->d_sweep[g], t, table_head_bytes(), hipMemcpyHostToDevice, s));
    if (count[g])
      BKB_HIP(hipMemcpyAsync(&b->d_sweep[g]->seg[0], &t->seg[0], sizeof(bk::BatchSeg) * count[g],
                             hipMemcpyHostToDevice, s));
  }
  for (int d = 0; d < kDtypes; ++d) {
    bk::BatchTable *t = b->h_ref[d];
    t->nsegs = rcount[d];
    if (rcount[d] && t->prefix[rcount[d]]) b->ref_mask |= 1u << d;
    BKB_HIP(hipMemcpyAsync(b->d_ref[d], t, table_head_bytes(), hipMemcpyHostToDevice, s));
    if (rcount[d])
      BKB_HIP(hipMemcpyAsync(&b->d_ref[d]->seg[0], &t->seg[0], sizeof(bk::BatchSeg) * rcount[d],
                             hipMemcpyHostToDevice, s));
  }
  if (n)
    BKB_HIP(hipMemcpyAsync(b->d_coeff, b->h_coeff, sizeof(double) * kCoeffSlot * n,
                           hipMemcpyHostToDevice, s));
  BKB_HIP(hipEventRecord(b->staged, s));
  return BK_OK;
}

int bk_batch_run(bk_batch b, uint32_t nticks, int use_graph, bk_stream stream) {
  if (!b) return fail(BK_ERR_ARG, "bk_batch_run: null batch");
  if (nticks == 0) return BK_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(b->mu);
  if (!use_graph) {
    BKB_HIP(enqueue_run(b, nticks, s));
    return BK_OK;
  }
  const std::tuple<uint32_t, uint64_t, uint32_t> key{nticks, b->sweep_mask, b->ref_mask};
  if (!b->graphs.find(key)) {
    BKB_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_run(b, nticks, b->capture);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->capture, &g);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hipfail(e, "capture of batched run");
    }
    if (e2 != hipSuccess) return hipfail(e2, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hipfail(e, "hipGraphInstantiate");
    BKB_HIP(b->graphs.insert(key, ge));
  }
  BKB_HIP(b->graphs.launch(key, s));
  return BK_OK;
}

}  // extern "C"
