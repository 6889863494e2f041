// bk_internal.h -- host-side declarations shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "bk_api.h"

namespace bk {

// Normalised sweep description handed from the C-ABI to the launchers.
struct SweepLaunch {
  int stencil;
  int dtype;
  const uint32_t *adj;
  const uint32_t *list;
  uint32_t nlist;
  const void *in;
  size_t in_step;
  void *out;
  size_t out_step;
  double coeff[125];
  int ncoeff;
  double theta;
  int cond_mode;
  // fused refresh (per-brick star kernel only): the ghost images of brick b
  // are img[img_off[b] .. img_off[b+1]); each also receives b's output
  const uint32_t *img_off = nullptr;
  const uint32_t *img = nullptr;
  int layout = BK_LAYOUT_8x8x8;  // brick shape and fold of both fields
  int wg_bricks = 0;             // bricks per workgroup (0: the kernel's default)
  int lane_shift = 0;            // quad star kernels: i-taps by DPP lane shifts
};

// Device view of an execution plan (see bk_plan.hip).
struct PlanView {
  int shape;
  uint32_t kcol;
  uint32_t nunits;
  uint32_t ncovered;
  const uint32_t *anchor;
  const uint32_t *len;
  const uint32_t *mask;
  const uint32_t *flat;
  const uint32_t *start;
  uint32_t nbricks;  // bricks of the adjacency the plan was built on
  uint64_t id;       // unique per plan (graph-cache key), never 0
};
const PlanView plan_view(const bk_plan_s *p);

// The quad schedule addresses its input through 32-bit element offsets:
// brick * step + element must stay below 2^32.
inline bool quad_offsets_fit(uint32_t nbricks, uint64_t step) {
  return (uint64_t)nbricks * step <= 0x100000000ull;
}

// Per-brick kernels over L.list (any list).
hipError_t launch_star(const SweepLaunch &L, hipStream_t s);
// Per-brick kernels for the 4x8x16 layouts (any stencil).
hipError_t launch_shape(const SweepLaunch &L, hipStream_t s);
hipError_t launch_cube(const SweepLaunch &L, hipStream_t s);
// Marching kernels over a plan: stars over k-columns, the cube over 2x2
// column quads streamed plane by plane.
hipError_t launch_star_columns(const SweepLaunch &L, const PlanView &P, hipStream_t s);
// Stars over 2x2 column quads: four waves share one LDS tile of the block.
hipError_t launch_star_quads(const SweepLaunch &L, const PlanView &P, hipStream_t s);
hipError_t launch_cube_quads(const SweepLaunch &L, const PlanView &P, hipStream_t s);

int plan_fail(int code, const char *msg);

// ---- batched sweeps (bk_batch.hip) ---------------------------------------
// A group table lists the requests (segments) one persistent kernel serves:
// the concatenation of their work items (bricks for the stars, quad-plan
// units for the cube, (dst, src) ghost pairs for the refresh).  Tables live
// in HBM; the host rewrites them between runs, the kernels read them.
constexpr uint32_t kBatchMaxSegs = 2048;

struct BatchSeg {
  const uint32_t *adj;    // the request's adjacency (its own brick numbering)
  const uint32_t *items;  // stars: bricks or column-plan anchors; cube: plan anchors;
                          // refresh: pairs
  const uint32_t *len;    // plans: bricks (column) or levels (quad) per unit
  const uint32_t *mask;   // plans: store masks [unit][kcol]
  void *f0;               // the request's two fields
  void *f1;
  const void *coeff;      // coefficients in the accumulate type
  uint64_t step;          // elements per brick, both fields
  uint32_t nitems;
  uint32_t kcol;
  uint32_t parity;  // the field holding the input at tick 0 (0: f0)
  uint32_t ticks;   // the request takes part in ticks [0, ticks) of a run
  double theta;
  int32_t cond_mode;
  int32_t pad;
};

struct BatchTable {
  uint32_t nsegs;
  uint32_t pad;
  uint32_t prefix[kBatchMaxSegs + 1];  // items before segment s
  BatchSeg seg[kBatchMaxSegs];
};

hipError_t launch_star_batched(int stencil, int dtype, const BatchTable *tab, uint32_t tick,
                               hipStream_t s);
hipError_t launch_star_columns_batched(int stencil, int dtype, const BatchTable *tab,
                                       uint32_t tick, hipStream_t s);
hipError_t launch_star_quads_batched(int stencil, int dtype, const BatchTable *tab,
                                     uint32_t tick, hipStream_t s);
hipError_t launch_cube_batched(int dtype, const BatchTable *tab, uint32_t tick, hipStream_t s);
// quad units the persistent batched cube kernel holds at once (one per wave)
uint32_t cube_batched_capacity(int dtype);
// quad units the persistent batched star kernel holds at once (one per workgroup)
uint32_t star_quads_batched_capacity(int stencil, int dtype);
hipError_t launch_ghost_batched(int dtype, const BatchTable *tab, uint32_t tick, hipStream_t s);

hipError_t launch_ghost_refresh(void *dat, size_t step_bytes,
                                size_t brick_bytes, const uint32_t *pairs,
                                uint32_t npairs, hipStream_t s);
hipError_t launch_zero(void *dst, size_t bytes, hipStream_t s);
hipError_t launch_stream_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t launch_valu_probe(int dtype, int chains, uint32_t nblocks, uint32_t iters, void *out,
                             hipStream_t s);
hipError_t launch_fill_uniform(int layout, int dtype, void *dat, size_t step,
                               const long nb[3], int ghost, uint64_t seed,
                               double lo, double hi, hipStream_t s);
// Deterministic two-pass FP64 reductions; `scratch` holds kReduceBlocks*nout
// doubles (pass-1 partials).  Results land in scratch[0 .. nout), or in
// out[0 .. nout) for launch_field_stats when `out` is not null.
constexpr int kReduceBlocks = 1024;
hipError_t launch_field_stats(int dtype, const void *dat, size_t step,
                              const uint32_t *list, uint32_t nlist,
                              size_t brick_elems, double *scratch, double *out,
                              hipStream_t s);
hipError_t launch_field_compare(int dta, const void *a, size_t step_a,
                                int dtr, const void *r, size_t step_r,
                                const uint32_t *list, uint32_t nlist,
                                size_t brick_elems, double floor_,
                                double atol, double rtol, double *scratch,
                                hipStream_t s);
hipError_t launch_copy_brick(int to_brick, int dtype, const long dims[3],
                             const long padding[3], const long ghost[3],
                             void *arr, const uint32_t *grid, void *dat,
                             size_t step, hipStream_t s);

}  // namespace bk
