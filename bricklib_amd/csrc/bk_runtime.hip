// bk_runtime.hip -- the C-ABI of libbricks.so (declared in include/bk_api.h).
//
// Host-side entry points: status/error reporting, device-resident storage,
// streams and events, brick-grid construction, sweep dispatch, the
// graph-captured ping-pong driver and the FP64 reductions.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <algorithm>
#include <vector>

#include "bk_api.h"
#include "bk_graphs.h"
#include "bk_internal.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(BK_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

#define BK_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

inline hipStream_t S(bk_stream s) { return static_cast<hipStream_t>(s); }

bool valid_dtype(int dt) { return dt == BK_F64 || dt == BK_F32 || dt == BK_BF16; }

bool valid_layout(int l) {
  return l == BK_LAYOUT_8x8x8 || l == BK_LAYOUT_4x8x16 || l == BK_LAYOUT_4x8x16_F88;
}

// brick extents (k, j, i) of a layout
void layout_ext(int l, int e[3]) {
  e[0] = l == BK_LAYOUT_8x8x8 ? 8 : 4;
  e[1] = 8;
  e[2] = l == BK_LAYOUT_8x8x8 ? 8 : 16;
}

int dt_size(int dt) { return dt == BK_F64 ? 8 : dt == BK_F32 ? 4 : dt == BK_BF16 ? 2 : 0; }

// Per-device FP64 reduction scratch, allocated once and guarded by a mutex;
// the guarded section ends with a stream sync, so calls do not overlap.
struct Scratch {
  std::mutex mu;
  std::map<int, double *> buf;
};
Scratch &reduce_scratch() {
  static Scratch s;
  return s;
}

int get_scratch(double **p) {
  int dev = 0;
  BK_HIP(hipGetDevice(&dev));
  auto &sc = reduce_scratch();
  auto it = sc.buf.find(dev);
  if (it != sc.buf.end()) {
    *p = it->second;
    return BK_OK;
  }
  void *q = nullptr;
  BK_HIP(hipMalloc(&q, sizeof(double) * bk::kReduceBlocks * 8));
  sc.buf[dev] = static_cast<double *>(q);
  *p = static_cast<double *>(q);
  return BK_OK;
}

}  // namespace

// Pass-1 partials of bk_field_stats_noalloc: a file-scope device array, one
// copy per GPU at its own address, held until the process ends (24 KiB).
// Its size does not depend on the input: the first pass is grid-stride over
// a fixed kReduceBlocks blocks, each writing its own three partials, and the
// second pass reads all of them, so every slot is rewritten by every call
// and nothing needs clearing.  Calls that overlap on one GPU share it and
// race.
constexpr size_t kStaticScratchDoubles = (size_t)bk::kReduceBlocks * 3;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {

// The array's device address on each GPU, looked up once per device with
// that device current (its name is a host-side stand-in, not an address).
constexpr int kMaxDevices = 64;
struct StaticScratch {
  std::mutex mu;
  double *ptr[kMaxDevices] = {};
};
StaticScratch &static_scratch() {
  static StaticScratch s;
  return s;
}

int static_scratch_ptr(double **p) {
  int dev = 0;
  BK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(BK_ERR_ARG, "bk_field_stats_noalloc: device %d beyond the %d supported", dev,
                kMaxDevices);
  auto &st = static_scratch();
  std::lock_guard<std::mutex> lock(st.mu);
  if (!st.ptr[dev]) {
    void *q = nullptr;
    BK_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    if (!q) return fail(BK_ERR_HIP, "bk_field_stats_noalloc: no device address for the scratch array");
    st.ptr[dev] = static_cast<double *>(q);
  }
  *p = st.ptr[dev];
  return BK_OK;
}

// Graph cache for bk_sweeps(use_graph): key = every argument byte, with
// plans keyed by their ids (a plan's host handle can be reused by a later
// plan once freed; its id cannot).  At most 64 graphs, evicted least
// recently used once their last launch has completed.
struct GraphCache {
  std::mutex mu;
  bk::ExecCache<std::string> exec{64};
  std::map<int, hipStream_t> capture_stream;
};
GraphCache &graphs() {
  static GraphCache g;
  return g;
}

uint64_t plan_id(const bk_plan_s *p) { return p ? bk::plan_view(p).id : 0; }

// pingpong: the sequence also reads d_out (with out_step) as an input, so
// a bound on input addressing applies to both fields.
int check_desc(const bk_sweep_desc *d, bool pingpong = false) {
  if (!d) return fail(BK_ERR_ARG, "bk_sweep: null descriptor");
  const int nc = bk_stencil_ncoeff(d->stencil);
  if (nc == 0) return fail(BK_ERR_ARG, "bk_sweep: unknown stencil %d", d->stencil);
  if (!valid_dtype(d->dtype)) return fail(BK_ERR_ARG, "bk_sweep: unknown dtype %d", d->dtype);
  if (!d->h_coeff) return fail(BK_ERR_ARG, "bk_sweep: null coefficient pointer");
  if (d->stencil == BK_COND7 && d->cond_mode != BK_COND_REFERENCE &&
      d->cond_mode != BK_COND_THRESHOLD)
    return fail(BK_ERR_ARG, "bk_sweep: unknown cond mode %d", d->cond_mode);
  if (!valid_layout(d->layout)) return fail(BK_ERR_ARG, "bk_sweep: unknown layout %d", d->layout);
  if (d->wg_bricks != 0 && d->wg_bricks != 1 && d->wg_bricks != 2 && d->wg_bricks != 4 &&
      d->wg_bricks != 8)
    return fail(BK_ERR_ARG, "bk_sweep: wg_bricks must be 0, 1, 2, 4 or 8 (got %d)",
                d->wg_bricks);
  if (d->wg_bricks == 8 && (d->layout != BK_LAYOUT_8x8x8 || d->stencil == BK_CUBE125))
    return fail(BK_ERR_UNSUPPORTED, "bk_sweep: 8 bricks per workgroup is built for the 8x8x8 "
                "stars only");
  if (d->layout != BK_LAYOUT_8x8x8 && d->plan)
    return fail(BK_ERR_UNSUPPORTED, "bk_sweep: plans are built for the 8x8x8 layout only");
  if (d->lane_shift != 0 && d->lane_shift != 1)
    return fail(BK_ERR_ARG, "bk_sweep: lane_shift must be 0 or 1 (got %d)", d->lane_shift);
  if (d->lane_shift &&
      (d->stencil == BK_CUBE125 || !d->plan || bk::plan_view(d->plan).shape != BK_PLAN_QUAD))
    return fail(BK_ERR_UNSUPPORTED, "bk_sweep: lane_shift is built into the star and cond "
                "kernels of quad plans only");
  if (!d->plan && d->nlist == 0) return BK_OK;
  if (!d->d_adj || !d->d_in || !d->d_out || (!d->plan && !d->d_list))
    return fail(BK_ERR_ARG, "bk_sweep: null device pointer");
  if (d->d_in == d->d_out)
    return fail(BK_ERR_ARG, "bk_sweep: input and output fields alias (use ping-pong fields)");
  if (d->in_step < 512 || d->out_step < 512)
    return fail(BK_ERR_ARG, "bk_sweep: step smaller than a brick (512 elements)");
  if (d->plan && d->stencil == BK_CUBE125) {
    const bk::PlanView P = bk::plan_view(d->plan);
    const size_t in_step = pingpong ? std::max(d->in_step, d->out_step) : d->in_step;
    if (P.shape == BK_PLAN_QUAD && !bk::quad_offsets_fit(P.nbricks, in_step))
      return fail(BK_ERR_UNSUPPORTED,
                  "bk_sweep: the quad schedule addresses at most 2^32 input elements "
                  "(nbricks * step); use a smaller field or no plan");
  }
  return BK_OK;
}

bool desc_empty(const bk_sweep_desc *d) {
  if (d->plan) {
    uint32_t n = 0;
    bk_plan_query(d->plan, nullptr, nullptr, nullptr, &n);
    return n == 0;
  }
  return d->nlist == 0;
}

void to_launch(const bk_sweep_desc *d, bk::SweepLaunch &L) {
  L.stencil = d->stencil;
  L.dtype = d->dtype;
  L.adj = d->d_adj;
  L.list = d->d_list;
  L.nlist = d->nlist;
  L.in = d->d_in;
  L.in_step = d->in_step;
  L.out = d->d_out;
  L.out_step = d->out_step;
  L.ncoeff = bk_stencil_ncoeff(d->stencil);
  for (int q = 0; q < 125; ++q) L.coeff[q] = q < L.ncoeff ? d->h_coeff[q] : 0.0;
  L.theta = d->theta;
  L.cond_mode = d->cond_mode;
  L.layout = d->layout;
  L.wg_bricks = d->wg_bricks;
  L.lane_shift = d->lane_shift;
}

hipError_t launch_sweep(const bk::SweepLaunch &L, const bk_plan_s *plan, hipStream_t s) {
  if (plan) {
    const bk::PlanView P = bk::plan_view(plan);
    if (P.ncovered == 0) return hipSuccess;
    if (L.stencil == BK_CUBE125 && P.shape == BK_PLAN_QUAD) return bk::launch_cube_quads(L, P, s);
    if (L.stencil != BK_CUBE125 && P.shape == BK_PLAN_COLUMN)
      return bk::launch_star_columns(L, P, s);
    if (L.stencil != BK_CUBE125 && P.shape == BK_PLAN_QUAD) return bk::launch_star_quads(L, P, s);
    bk::SweepLaunch F = L;  // other pairings: per-brick kernel over the plan's bricks
    F.list = P.flat;
    F.nlist = P.ncovered;
    return F.stencil == BK_CUBE125 ? bk::launch_cube(F, s) : bk::launch_star(F, s);
  }
  if (L.nlist == 0) return hipSuccess;
  if (L.layout != BK_LAYOUT_8x8x8) return bk::launch_shape(L, s);
  if (L.stencil == BK_CUBE125) return bk::launch_cube(L, s);
  return bk::launch_star(L, s);
}

// Enqueue the ping-pong sequence of bk_sweeps on stream s.
// The ping-pong sequence: before sweep t the ghost bricks of its input are
// refreshed when t % ghost_iters == 0; sweep t covers the desc's brick list
// (or plan) when it is the last of its refresh group or of the sequence,
// and the `all` list (or plan) -- every brick of the padded grid, ghosts
// included -- otherwise.  ghost_iters == 1 is the refresh-every-sweep loop.
hipError_t enqueue_sweeps(const bk_sweep_desc *d, const uint32_t *pairs, uint32_t npairs,
                          int nsweeps, int ghost_iters, const uint32_t *all, uint32_t nall,
                          bk_plan all_plan, const uint32_t *img_off, const uint32_t *img,
                          hipStream_t s) {
  bk::SweepLaunch L;
  to_launch(d, L);
  // fused refresh: the input's ghosts once, then every sweep also writes
  // its output into the output's ghost bricks
  L.img_off = img_off;
  L.img = img;
  const size_t es = (size_t)dt_size(d->dtype);
  for (int t = 0; t < nsweeps; ++t) {
    const bool fwd = (t % 2) == 0;
    void *src = fwd ? const_cast<void *>(d->d_in) : d->d_out;
    void *dst = fwd ? d->d_out : const_cast<void *>(d->d_in);
    const size_t sst = fwd ? d->in_step : d->out_step;
    const size_t dstep = fwd ? d->out_step : d->in_step;
    if (npairs && (img_off ? t == 0 : t % ghost_iters == 0)) {
      hipError_t e = bk::launch_ghost_refresh(src, sst * es, 512 * es, pairs, npairs, s);
      if (e != hipSuccess) return e;
    }
    const bool last = t % ghost_iters == ghost_iters - 1 || t == nsweeps - 1;
    L.in = src;
    L.in_step = sst;
    L.out = dst;
    L.out_step = dstep;
    L.list = last ? d->d_list : all;
    L.nlist = last ? d->nlist : nall;
    hipError_t e = launch_sweep(L, last ? d->plan : all_plan, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void grid_dims(const long nb[3], long S[3]) {
  for (int a = 0; a < 3; ++a) S[a] = nb[a] + 2;
}

}  // namespace

namespace bk {
int plan_fail(int code, const char *msg) { return fail(code, "%s", msg); }

void graphs_forget_plan(uint64_t plan_id) {
  auto &gc = graphs();
  std::lock_guard<std::mutex> lock(gc.mu);
  (void)gc.exec.forget(plan_id);
}
}  // namespace bk

extern "C" {

int bk_version(void) { return BK_API_VERSION; }

const char *bk_last_error(void) { return g_err.c_str(); }

int bk_stencil_ncoeff(int stencil) {
  switch (stencil) {
    case BK_STAR7: case BK_COND7: return 7;
    case BK_STAR13: return 13;
    case BK_STAR25: return 25;
    case BK_CUBE125: return 125;
    default: return 0;
  }
}

int bk_layout_extents(int layout, int ext[3]) {
  if (!ext) return fail(BK_ERR_ARG, "bk_layout_extents: null pointer");
  if (!valid_layout(layout)) return fail(BK_ERR_ARG, "bk_layout_extents: unknown layout %d", layout);
  layout_ext(layout, ext);
  return BK_OK;
}

int bk_stencil_radius(int stencil) {
  switch (stencil) {
    case BK_STAR7: case BK_COND7: return 1;
    case BK_STAR13: return 2;
    case BK_STAR25: return 4;
    case BK_CUBE125: return 2;
    default: return 0;
  }
}

int bk_dtype_size(int dtype) { return dt_size(dtype); }

int bk_device_count(int *n) {
  if (!n) return fail(BK_ERR_ARG, "bk_device_count: null pointer");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess || c == 0) {
    *n = 0;
    return fail(BK_ERR_NODEV, "no HIP device visible (%s)", hipGetErrorString(e));
  }
  *n = c;
  return BK_OK;
}

int bk_set_device(int dev) {
  BK_HIP(hipSetDevice(dev));
  return BK_OK;
}

int bk_device_sync(void) {
  BK_HIP(hipDeviceSynchronize());
  return BK_OK;
}

int bk_pointer_is_device(const void *p, int *is_device) {
  if (!is_device) return fail(BK_ERR_ARG, "bk_pointer_is_device: null out pointer");
  *is_device = 0;
  if (!p) return BK_OK;
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof(a));
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // plain host memory: not an error for the caller
    return BK_OK;
  }
  *is_device = (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) ? 1 : 0;
  return BK_OK;
}

int bk_alloc(size_t bytes, void **d_ptr) {
  if (!d_ptr) return fail(BK_ERR_ARG, "bk_alloc: null out pointer");
  *d_ptr = nullptr;
  if (bytes == 0) return BK_OK;
  hipError_t e = hipMalloc(d_ptr, bytes);
  if (e == hipErrorOutOfMemory) return fail(BK_ERR_NOMEM, "bk_alloc: %zu bytes: out of memory", bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return BK_OK;
}

int bk_storage_alloc(int dtype, long chunks, size_t step, void **d_dat) {
  if (!d_dat) return fail(BK_ERR_ARG, "bk_storage_alloc: null out pointer");
  *d_dat = nullptr;
  if (!valid_dtype(dtype)) return fail(BK_ERR_ARG, "bk_storage_alloc: unknown dtype %d", dtype);
  if (chunks <= 0 || step == 0)
    return fail(BK_ERR_ARG, "bk_storage_alloc: chunks and step must be positive");
  const size_t es = (size_t)dt_size(dtype);
  if ((size_t)chunks > SIZE_MAX / step / es)
    return fail(BK_ERR_ARG, "bk_storage_alloc: size overflows");
  // hipMalloc returns at least 256-byte aligned memory; round the request
  // to the 2048-byte ALIGN of brick.h so the last chunk never straddles.
  size_t bytes = (size_t)chunks * step * es;
  bytes = (bytes + 2047) & ~(size_t)2047;
  return bk_alloc(bytes, d_dat);
}

// The default stream-ordered pool hands freed memory back to the driver at
// every synchronisation (release threshold 0), so a loop that frees grids
// and allocates new ones pays an unmap + map of every field each round.
// Keep freed memory in the pool instead (per device, once).
namespace {
std::mutex g_pool_mu;
bool g_pool_kept[64] = {};

hipError_t keep_pool_memory() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipSuccess;
  std::lock_guard<std::mutex> lock(g_pool_mu);
  if (g_pool_kept[dev]) return hipSuccess;
  hipMemPool_t pool = nullptr;
  e = hipDeviceGetDefaultMemPool(&pool, dev);
  if (e != hipSuccess) return e;
  uint64_t keep = UINT64_MAX;
  e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  if (e != hipSuccess) return e;
  g_pool_kept[dev] = true;
  return hipSuccess;
}
}  // namespace

int bk_alloc_async(size_t bytes, void **d_ptr, bk_stream s) {
  if (!d_ptr) return fail(BK_ERR_ARG, "bk_alloc_async: null out pointer");
  *d_ptr = nullptr;
  if (bytes == 0) return BK_OK;
  BK_HIP(keep_pool_memory());
  hipError_t e = hipMallocAsync(d_ptr, bytes, S(s));
  if (e == hipErrorOutOfMemory)
    return fail(BK_ERR_NOMEM, "bk_alloc_async: %zu bytes: out of memory", bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
  return BK_OK;
}

int bk_free_async(void *d_ptr, bk_stream s) {
  if (!d_ptr) return BK_OK;
  BK_HIP(hipFreeAsync(d_ptr, S(s)));
  return BK_OK;
}

int bk_free(void *d_ptr) {
  if (!d_ptr) return BK_OK;
  BK_HIP(hipFree(d_ptr));
  return BK_OK;
}

int bk_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, bk_stream s) {
  if (bytes == 0) return BK_OK;
  if (!d_dst || !h_src) return fail(BK_ERR_ARG, "bk_memcpy_h2d: null pointer");
  BK_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, S(s)));
  BK_HIP(hipStreamSynchronize(S(s)));
  return BK_OK;
}

int bk_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, bk_stream s) {
  if (bytes == 0) return BK_OK;
  if (!h_dst || !d_src) return fail(BK_ERR_ARG, "bk_memcpy_d2h: null pointer");
  BK_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, S(s)));
  BK_HIP(hipStreamSynchronize(S(s)));
  return BK_OK;
}

int bk_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, bk_stream s) {
  if (bytes == 0) return BK_OK;
  if (!d_dst || !d_src) return fail(BK_ERR_ARG, "bk_memcpy_d2d: null pointer");
  BK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, S(s)));
  return BK_OK;
}

int bk_stream_copy(void *d_dst, const void *d_src, size_t bytes, bk_stream s) {
  if (bytes == 0) return BK_OK;
  if (!d_dst || !d_src) return fail(BK_ERR_ARG, "bk_stream_copy: null pointer");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(d_dst) % 16 || reinterpret_cast<uintptr_t>(d_src) % 16)
    return fail(BK_ERR_ARG, "bk_stream_copy: pointers and size must be 16-byte aligned");
  const char *a = static_cast<const char *>(d_src);
  const char *b = static_cast<const char *>(d_dst);
  if (a < b + bytes && b < a + bytes) return fail(BK_ERR_ARG, "bk_stream_copy: ranges overlap");
  BK_HIP(bk::launch_stream_copy(d_dst, d_src, bytes, S(s)));
  return BK_OK;
}

int bk_valu_probe(int dtype, int chains, int waves_per_simd, uint32_t iters, void *d_out,
                  size_t out_elems, double *flops, bk_stream s) {
  if (dtype != BK_F64 && dtype != BK_F32)
    return fail(BK_ERR_ARG, "bk_valu_probe: dtype must be BK_F64 or BK_F32");
  if (chains != 1 && chains != 8) return fail(BK_ERR_ARG, "bk_valu_probe: chains must be 1 or 8");
  if (waves_per_simd < 1 || waves_per_simd > 8)
    return fail(BK_ERR_ARG, "bk_valu_probe: waves_per_simd in [1, 8]");
  if (!d_out || !flops) return fail(BK_ERR_ARG, "bk_valu_probe: null pointer");
  int dev = 0, cus = 0;
  BK_HIP(hipGetDevice(&dev));
  BK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const uint64_t nblocks = (uint64_t)waves_per_simd * (uint64_t)cus * 4u;  // 4 SIMDs per CU
  if (out_elems < nblocks * 64) {
    char msg[160];
    snprintf(msg, sizeof msg, "bk_valu_probe: d_out holds %zu elements, the launch stores %llu",
             out_elems, (unsigned long long)(nblocks * 64));
    return fail(BK_ERR_ARG, msg);
  }
  *flops = 2.0 * 64.0 * (double)iters * (double)nblocks * 64.0;
  BK_HIP(bk::launch_valu_probe(dtype, chains, (uint32_t)nblocks, iters, d_out, S(s)));
  return BK_OK;
}

int bk_zero(void *d_dst, size_t bytes, bk_stream s) {
  if (bytes == 0) return BK_OK;
  if (!d_dst) return fail(BK_ERR_ARG, "bk_zero: null pointer");
  BK_HIP(bk::launch_zero(d_dst, bytes, S(s)));
  return BK_OK;
}

int bk_stream_create(bk_stream *s) {
  if (!s) return fail(BK_ERR_ARG, "bk_stream_create: null pointer");
  hipStream_t h;
  BK_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = h;
  return BK_OK;
}

int bk_stream_destroy(bk_stream s) {
  if (!s) return BK_OK;
  BK_HIP(hipStreamDestroy(S(s)));
  return BK_OK;
}

int bk_stream_sync(bk_stream s) {
  BK_HIP(hipStreamSynchronize(S(s)));
  return BK_OK;
}

int bk_event_create(bk_event *e) {
  if (!e) return fail(BK_ERR_ARG, "bk_event_create: null pointer");
  hipEvent_t h;
  BK_HIP(hipEventCreate(&h));
  *e = h;
  return BK_OK;
}

int bk_event_destroy(bk_event e) {
  if (!e) return BK_OK;
  BK_HIP(hipEventDestroy(static_cast<hipEvent_t>(e)));
  return BK_OK;
}

int bk_event_record(bk_event e, bk_stream s) {
  if (!e) return fail(BK_ERR_ARG, "bk_event_record: null event");
  BK_HIP(hipEventRecord(static_cast<hipEvent_t>(e), S(s)));
  return BK_OK;
}

int bk_event_sync(bk_event e) {
  if (!e) return fail(BK_ERR_ARG, "bk_event_sync: null event");
  BK_HIP(hipEventSynchronize(static_cast<hipEvent_t>(e)));
  return BK_OK;
}

int bk_event_elapsed_ms(bk_event start, bk_event stop, float *ms) {
  if (!start || !stop || !ms) return fail(BK_ERR_ARG, "bk_event_elapsed_ms: null pointer");
  BK_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return BK_OK;
}

// ---------------------------------------------------------------------------
// grid helpers (host)
// ---------------------------------------------------------------------------
int bk_init_grid(const long dims[3], int periodic, uint32_t *h_grid, uint32_t *h_adj) {
  if (!dims || !h_grid || !h_adj) return fail(BK_ERR_ARG, "bk_init_grid: null pointer");
  for (int a = 0; a < 3; ++a)
    if (dims[a] <= 0) return fail(BK_ERR_ARG, "bk_init_grid: dims must be positive");
  const long s1 = dims[0], s2 = dims[0] * dims[1];
  const long size = s2 * dims[2];
  if (size > (long)UINT32_MAX) return fail(BK_ERR_ARG, "bk_init_grid: more than 2^32 bricks");
  for (long p = 0; p < size; ++p) h_grid[p] = (uint32_t)p;
  for (long k = 0; k < dims[2]; ++k)
    for (long j = 0; j < dims[1]; ++j)
      for (long i = 0; i < dims[0]; ++i) {
        const long pos = i + j * s1 + k * s2;
        uint32_t *a = h_adj + (size_t)pos * 27;
        for (int kd = 0; kd < 3; ++kd)
          for (int jd = 0; jd < 3; ++jd)
            for (int id = 0; id < 3; ++id) {
              long v;
              if (periodic) {
                const long ii = (i + id - 1 + dims[0]) % dims[0];
                const long jj = (j + jd - 1 + dims[1]) % dims[1];
                const long kk = (k + kd - 1 + dims[2]) % dims[2];
                v = ii + jj * s1 + kk * s2;
              } else {
                // the reference accepts any in-range linear position
                // (bricksetup.h:42-47), which wraps along rows at edges
                const long q = pos + (id - 1) + (jd - 1) * s1 + (kd - 1) * s2;
                v = (q >= 0 && q < size) ? q : 0;
              }
              a[kd * 9 + jd * 3 + id] = (uint32_t)v;
            }
      }
  return BK_OK;
}

int bk_ghost_pairs(const long nb[3], uint32_t *h_pairs, uint32_t *npairs) {
  if (!nb || !npairs) return fail(BK_ERR_ARG, "bk_ghost_pairs: null pointer");
  for (int a = 0; a < 3; ++a)
    if (nb[a] <= 0) return fail(BK_ERR_ARG, "bk_ghost_pairs: nb must be positive");
  long S[3];
  grid_dims(nb, S);
  uint32_t n = 0;
  for (long k = 0; k < S[2]; ++k)
    for (long j = 0; j < S[1]; ++j)
      for (long i = 0; i < S[0]; ++i) {
        const bool ghost = i == 0 || i == S[0] - 1 || j == 0 || j == S[1] - 1 || k == 0 ||
                           k == S[2] - 1;
        if (!ghost) continue;
        if (h_pairs) {
          const long si = (i - 1 + nb[0]) % nb[0] + 1;
          const long sj = (j - 1 + nb[1]) % nb[1] + 1;
          const long sk = (k - 1 + nb[2]) % nb[2] + 1;
          h_pairs[2 * n] = (uint32_t)(i + S[0] * (j + S[1] * k));
          h_pairs[2 * n + 1] = (uint32_t)(si + S[0] * (sj + S[1] * sk));
        }
        ++n;
      }
  *npairs = n;
  return BK_OK;
}

int bk_ghost_images(const long nb[3], uint32_t *h_off, uint32_t *h_img, uint32_t *nimg) {
  if (!nb || !nimg) return fail(BK_ERR_ARG, "bk_ghost_images: null pointer");
  uint32_t np = 0;
  int rc = bk_ghost_pairs(nb, nullptr, &np);
  if (rc != BK_OK) return rc;
  *nimg = np;
  if (!h_off && !h_img) return BK_OK;
  if (!h_off || !h_img) return fail(BK_ERR_ARG, "bk_ghost_images: null output array");
  long S[3];
  grid_dims(nb, S);
  const size_t nbricks = (size_t)S[0] * S[1] * S[2];
  std::vector<uint32_t> pairs(2 * (size_t)np);
  bk_ghost_pairs(nb, pairs.data(), &np);
  // counting sort of the ghost bricks by their periodic image
  std::fill(h_off, h_off + nbricks + 1, 0u);
  for (uint32_t q = 0; q < np; ++q) ++h_off[pairs[2 * q + 1] + 1];
  for (size_t b = 0; b < nbricks; ++b) h_off[b + 1] += h_off[b];
  std::vector<uint32_t> fillp(h_off, h_off + nbricks);
  for (uint32_t q = 0; q < np; ++q) h_img[fillp[pairs[2 * q + 1]]++] = pairs[2 * q];
  return BK_OK;
}

int bk_interior_list(const long nb[3], uint32_t *h_list, uint32_t *n) {
  if (!nb || !n) return fail(BK_ERR_ARG, "bk_interior_list: null pointer");
  for (int a = 0; a < 3; ++a)
    if (nb[a] <= 0) return fail(BK_ERR_ARG, "bk_interior_list: nb must be positive");
  long S[3];
  grid_dims(nb, S);
  uint32_t c = 0;
  for (long k = 1; k <= nb[2]; ++k)
    for (long j = 1; j <= nb[1]; ++j)
      for (long i = 1; i <= nb[0]; ++i) {
        if (h_list) h_list[c] = (uint32_t)(i + S[0] * (j + S[1] * k));
        ++c;
      }
  *n = c;
  return BK_OK;
}

// ---------------------------------------------------------------------------
// sweeps
// ---------------------------------------------------------------------------
int bk_sweep(const bk_sweep_desc *d, bk_stream s) {
  int rc = check_desc(d);
  if (rc != BK_OK) return rc;
  if (desc_empty(d)) return BK_OK;
  bk::SweepLaunch L;
  to_launch(d, L);
  BK_HIP(launch_sweep(L, d->plan, S(s)));
  return BK_OK;
}

int bk_ghost_refresh(int dtype, void *d_dat, size_t step, size_t brick_elems,
                     const uint32_t *d_pairs, uint32_t npairs, bk_stream s) {
  if (!valid_dtype(dtype)) return fail(BK_ERR_ARG, "bk_ghost_refresh: unknown dtype %d", dtype);
  if (npairs == 0) return BK_OK;
  if (!d_dat || !d_pairs) return fail(BK_ERR_ARG, "bk_ghost_refresh: null pointer");
  if (brick_elems == 0 || brick_elems > step)
    return fail(BK_ERR_ARG, "bk_ghost_refresh: brick_elems must be in [1, step]");
  const size_t es = (size_t)dt_size(dtype);
  if ((brick_elems * es) % 2 != 0) return fail(BK_ERR_ARG, "bk_ghost_refresh: odd byte count");
  BK_HIP(bk::launch_ghost_refresh(d_dat, step * es, brick_elems * es, d_pairs, npairs, S(s)));
  return BK_OK;
}

namespace {

int sweeps_impl(const bk_sweep_desc *d, const uint32_t *d_pairs, uint32_t npairs, int nsweeps,
                int ghost_iters, const uint32_t *d_all, uint32_t nall, bk_plan all_plan,
                const uint32_t *img_off, const uint32_t *img, int use_graph, hipStream_t s) {
  if (nsweeps == 0 || desc_empty(d)) return BK_OK;
  if (!use_graph) {
    BK_HIP(enqueue_sweeps(d, d_pairs, npairs, nsweeps, ghost_iters, d_all, nall, all_plan,
                          img_off, img, s));
    return BK_OK;
  }
  // key: every argument that shapes the captured launches
  std::string key;
  auto put = [&key](const void *p, size_t n) { key.append(static_cast<const char *>(p), n); };
  int dev = 0;
  BK_HIP(hipGetDevice(&dev));
  put(&dev, sizeof(dev));
  put(&d->stencil, sizeof(d->stencil));
  put(&d->dtype, sizeof(d->dtype));
  put(&d->d_adj, sizeof(d->d_adj));
  put(&d->d_list, sizeof(d->d_list));
  put(&d->nlist, sizeof(d->nlist));
  put(&d->d_in, sizeof(d->d_in));
  put(&d->in_step, sizeof(d->in_step));
  put(&d->d_out, sizeof(d->d_out));
  put(&d->out_step, sizeof(d->out_step));
  const uint64_t pid = plan_id(d->plan), apid = plan_id(all_plan);
  put(&pid, sizeof(pid));
  put(d->h_coeff, sizeof(double) * (size_t)bk_stencil_ncoeff(d->stencil));
  put(&d->theta, sizeof(d->theta));
  put(&d->cond_mode, sizeof(d->cond_mode));
  put(&d->layout, sizeof(d->layout));
  put(&d->wg_bricks, sizeof(d->wg_bricks));
  put(&d->lane_shift, sizeof(d->lane_shift));
  put(&d_pairs, sizeof(d_pairs));
  put(&npairs, sizeof(npairs));
  put(&nsweeps, sizeof(nsweeps));
  put(&ghost_iters, sizeof(ghost_iters));
  put(&d_all, sizeof(d_all));
  put(&nall, sizeof(nall));
  put(&apid, sizeof(apid));
  put(&img_off, sizeof(img_off));
  put(&img, sizeof(img));
  auto &gc = graphs();
  std::lock_guard<std::mutex> lock(gc.mu);
  if (!gc.exec.find(key)) {
    hipStream_t cs;
    auto cit = gc.capture_stream.find(dev);
    if (cit == gc.capture_stream.end()) {
      BK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      gc.capture_stream[dev] = cs;
    } else {
      cs = cit->second;
    }
    BK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue_sweeps(d, d_pairs, npairs, nsweeps, ghost_iters, d_all, nall,
                                  all_plan, img_off, img, cs);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(e, "capture of sweep sequence");
    }
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);  // the executable graph keeps what it needs
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    BK_HIP(gc.exec.insert(key, ge, pid, apid));
  }
  BK_HIP(gc.exec.launch(key, s));
  return BK_OK;
}

}  // namespace

int bk_sweeps(const bk_sweep_desc *d, const uint32_t *d_pairs, uint32_t npairs, int nsweeps,
              int use_graph, bk_stream s) {
  int rc = check_desc(d, nsweeps > 1);
  if (rc != BK_OK) return rc;
  if (nsweeps < 0) return fail(BK_ERR_ARG, "bk_sweeps: negative sweep count");
  if (npairs && !d_pairs) return fail(BK_ERR_ARG, "bk_sweeps: null pair list");
  return sweeps_impl(d, d_pairs, npairs, nsweeps, 1, d->d_list, d->nlist, d->plan, nullptr,
                     nullptr, use_graph, S(s));
}

int bk_sweeps_fused(const bk_sweep_desc *d, const uint32_t *d_img_off, const uint32_t *d_img,
                    const uint32_t *d_pairs, uint32_t npairs, int nsweeps, int use_graph,
                    bk_stream s) {
  int rc = check_desc(d, nsweeps > 1);
  if (rc != BK_OK) return rc;
  if (nsweeps < 0) return fail(BK_ERR_ARG, "bk_sweeps_fused: negative sweep count");
  if (npairs && !d_pairs) return fail(BK_ERR_ARG, "bk_sweeps_fused: null pair list");
  if (!d_img_off || !d_img) return fail(BK_ERR_ARG, "bk_sweeps_fused: null ghost-image table");
  if (d->stencil == BK_CUBE125 || d->layout != BK_LAYOUT_8x8x8 ||
      (d->plan && bk::plan_view(d->plan).shape != BK_PLAN_QUAD))
    return fail(BK_ERR_UNSUPPORTED,
                "bk_sweeps_fused: the fused refresh runs in the star/cond kernels over 8x8x8 "
                "bricks one per wave or by quad plans (no column plan, no cube); use "
                "bk_sweeps");
  return sweeps_impl(d, d_pairs, npairs, nsweeps, 1, d->d_list, d->nlist, d->plan, d_img_off,
                     d_img, use_graph, S(s));
}

int bk_sweeps_deep(const bk_sweep_desc *d, const uint32_t *d_all, uint32_t nall,
                   bk_plan all_plan, const uint32_t *d_pairs, uint32_t npairs, int nsweeps,
                   int ghost_iters, int use_graph, bk_stream s) {
  int rc = check_desc(d, nsweeps > 1);
  if (rc != BK_OK) return rc;
  if (nsweeps < 0) return fail(BK_ERR_ARG, "bk_sweeps_deep: negative sweep count");
  if (npairs && !d_pairs) return fail(BK_ERR_ARG, "bk_sweeps_deep: null pair list");
  const int R = bk_stencil_radius(d->stencil);
  int ext[3];
  layout_ext(valid_layout(d->layout) ? d->layout : BK_LAYOUT_8x8x8, ext);
  const int depth = std::min(ext[0], std::min(ext[1], ext[2]));  // the ghost zone's thinnest side
  if (ghost_iters < 1 || R * ghost_iters > depth)
    return fail(BK_ERR_ARG,
                "bk_sweeps_deep: ghost_iters must be in [1, %d / radius] (%d sweeps of radius %d "
                "exceed the %d-cell ghost zone)",
                depth, ghost_iters, R, depth);
  bk_sweep_desc all = *d;
  all.d_list = d_all;
  all.nlist = nall;
  all.plan = all_plan;
  rc = check_desc(&all, nsweeps > 1);
  if (rc != BK_OK) return rc;
  return sweeps_impl(d, d_pairs, npairs, nsweeps, ghost_iters, d_all, nall, all_plan, nullptr,
                     nullptr, use_graph, S(s));
}

int bk_graph_cache_clear(void) {
  auto &gc = graphs();
  std::lock_guard<std::mutex> lock(gc.mu);
  BK_HIP(gc.exec.clear());
  return BK_OK;
}

// ---------------------------------------------------------------------------
// seeded fill and FP64 reductions
// ---------------------------------------------------------------------------
int bk_fill_uniform_layout(int layout, int dtype, void *d_dat, size_t step, const long nb[3],
                           int ghost, uint64_t seed, double lo, double hi, bk_stream s) {
  if (!valid_layout(layout)) return fail(BK_ERR_ARG, "bk_fill_uniform: unknown layout %d", layout);
  if (!valid_dtype(dtype)) return fail(BK_ERR_ARG, "bk_fill_uniform: unknown dtype %d", dtype);
  if (!d_dat || !nb) return fail(BK_ERR_ARG, "bk_fill_uniform: null pointer");
  for (int a = 0; a < 3; ++a)
    if (nb[a] <= 0) return fail(BK_ERR_ARG, "bk_fill_uniform: nb must be positive");
  if (ghost != 0 && ghost != 1) return fail(BK_ERR_ARG, "bk_fill_uniform: ghost must be 0 or 1");
  if (step < 512) return fail(BK_ERR_ARG, "bk_fill_uniform: step smaller than a brick");
  BK_HIP(bk::launch_fill_uniform(layout, dtype, d_dat, step, nb, ghost, seed, lo, hi, S(s)));
  return BK_OK;
}

int bk_fill_uniform_grid(int dtype, void *d_dat, size_t step, const long nb[3], int ghost,
                         uint64_t seed, double lo, double hi, bk_stream s) {
  return bk_fill_uniform_layout(BK_LAYOUT_8x8x8, dtype, d_dat, step, nb, ghost, seed, lo, hi, s);
}

int bk_fill_uniform(int dtype, void *d_dat, size_t step, const long nb[3], uint64_t seed,
                    double lo, double hi, bk_stream s) {
  return bk_fill_uniform_grid(dtype, d_dat, step, nb, 1, seed, lo, hi, s);
}

int bk_field_stats(int dtype, const void *d_dat, size_t step, const uint32_t *d_list,
                   uint32_t nlist, size_t brick_elems, double h_out[3], bk_stream s) {
  if (!valid_dtype(dtype)) return fail(BK_ERR_ARG, "bk_field_stats: unknown dtype %d", dtype);
  if (!h_out) return fail(BK_ERR_ARG, "bk_field_stats: null output");
  if (nlist && (!d_dat || !d_list)) return fail(BK_ERR_ARG, "bk_field_stats: null pointer");
  if (brick_elems == 0 || brick_elems > step)
    return fail(BK_ERR_ARG, "bk_field_stats: brick_elems must be in [1, step]");
  auto &sc = reduce_scratch();
  std::lock_guard<std::mutex> lock(sc.mu);
  double *buf = nullptr;
  int rc = get_scratch(&buf);
  if (rc != BK_OK) return rc;
  BK_HIP(bk::launch_field_stats(dtype, d_dat, step, d_list, nlist, brick_elems, buf, nullptr,
                                S(s)));
  BK_HIP(hipMemcpyAsync(h_out, buf, 3 * sizeof(double), hipMemcpyDeviceToHost, S(s)));
  BK_HIP(hipStreamSynchronize(S(s)));
  return BK_OK;
}

int bk_field_stats_noalloc(int dtype, const void *d_dat, size_t step, const uint32_t *d_list,
                           uint32_t nlist, size_t brick_elems, double *d_out, bk_stream s) {
  static_assert((size_t)bk::kReduceBlocks * 3 <= kStaticScratchDoubles,
                "pass-1 partials exceed the static scratch array");
  if (!valid_dtype(dtype))
    return fail(BK_ERR_ARG, "bk_field_stats_noalloc: unknown dtype %d", dtype);
  if (!d_out) return fail(BK_ERR_ARG, "bk_field_stats_noalloc: null output");
  if (nlist && (!d_dat || !d_list))
    return fail(BK_ERR_ARG, "bk_field_stats_noalloc: null pointer");
  if (brick_elems == 0 || brick_elems > step)
    return fail(BK_ERR_ARG, "bk_field_stats_noalloc: brick_elems must be in [1, step]");
  double *buf = nullptr;
  int rc = static_scratch_ptr(&buf);
  if (rc != BK_OK) return rc;
  BK_HIP(bk::launch_field_stats(dtype, d_dat, step, d_list, nlist, brick_elems, buf, d_out, S(s)));
  return BK_OK;
}

int bk_field_compare(int dtype_a, const void *d_a, size_t step_a, int dtype_r, const void *d_r,
                     size_t step_r, const uint32_t *d_list, uint32_t nlist, size_t brick_elems,
                     double floor_, double atol, double rtol, double h_out[5], bk_stream s) {
  if (!valid_dtype(dtype_a) || !valid_dtype(dtype_r))
    return fail(BK_ERR_ARG, "bk_field_compare: unknown dtype");
  if (!h_out) return fail(BK_ERR_ARG, "bk_field_compare: null output");
  if (nlist && (!d_a || !d_r || !d_list)) return fail(BK_ERR_ARG, "bk_field_compare: null pointer");
  if (brick_elems == 0 || brick_elems > step_a || brick_elems > step_r)
    return fail(BK_ERR_ARG, "bk_field_compare: brick_elems must be in [1, step]");
  auto &sc = reduce_scratch();
  std::lock_guard<std::mutex> lock(sc.mu);
  double *buf = nullptr;
  int rc = get_scratch(&buf);
  if (rc != BK_OK) return rc;
  BK_HIP(bk::launch_field_compare(dtype_a, d_a, step_a, dtype_r, d_r, step_r, d_list, nlist,
                                  brick_elems, floor_, atol, rtol, buf, S(s)));
  BK_HIP(hipMemcpyAsync(h_out, buf, 5 * sizeof(double), hipMemcpyDeviceToHost, S(s)));
  BK_HIP(hipStreamSynchronize(S(s)));
  return BK_OK;
}

// ---------------------------------------------------------------------------
// array <-> brick
// ---------------------------------------------------------------------------
static int check_copy(const char *who, int dtype, const long dims[3], const long padding[3],
                      const long ghost[3], const void *arr, const uint32_t *grid,
                      const void *dat, size_t step) {
  if (!valid_dtype(dtype)) return fail(BK_ERR_ARG, "%s: unknown dtype %d", who, dtype);
  if (!dims || !padding || !ghost || !arr || !grid || !dat)
    return fail(BK_ERR_ARG, "%s: null pointer", who);
  for (int a = 0; a < 3; ++a) {
    if (dims[a] < 0 || padding[a] < 0 || ghost[a] < 0)
      return fail(BK_ERR_ARG, "%s: negative extent", who);
    if ((dims[a] + 2 * ghost[a]) % 8 != 0 || ghost[a] % 8 != 0)
      return fail(BK_ERR_ARG, "%s: dims + 2*ghost and ghost must be multiples of 8", who);
  }
  if (step < 512) return fail(BK_ERR_ARG, "%s: step smaller than a brick", who);
  return BK_OK;
}

int bk_copy_to_brick(int dtype, const long dims[3], const long padding[3], const long ghost[3],
                     const void *d_arr, const uint32_t *d_grid, void *d_dat, size_t step,
                     bk_stream s) {
  int rc = check_copy("bk_copy_to_brick", dtype, dims, padding, ghost, d_arr, d_grid, d_dat, step);
  if (rc != BK_OK) return rc;
  BK_HIP(bk::launch_copy_brick(1, dtype, dims, padding, ghost, const_cast<void *>(d_arr), d_grid,
                               d_dat, step, S(s)));
  return BK_OK;
}

int bk_copy_from_brick(int dtype, const long dims[3], const long padding[3], const long ghost[3],
                       void *d_arr, const uint32_t *d_grid, const void *d_dat, size_t step,
                       bk_stream s) {
  int rc =
      check_copy("bk_copy_from_brick", dtype, dims, padding, ghost, d_arr, d_grid, d_dat, step);
  if (rc != BK_OK) return rc;
  BK_HIP(bk::launch_copy_brick(0, dtype, dims, padding, ghost, d_arr, d_grid,
                               const_cast<void *>(d_dat), step, S(s)));
  return BK_OK;
}

}  // extern "C"
