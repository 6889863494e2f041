/*
 * bk_api.h -- C-ABI of the MI355X brick-stencil library (libbricks.so).
 *
 * This is the drop-in boundary.  bricklib itself is header-only C++ with no
 * foreign-function interface; its GPU work is reached through the template
 * headers (brick.h, bricksetup.h, brick-gpu.h) and kernels written against
 * them.  Every entry point below names the reference interface it stands in
 * for (file:line in CtopCsUtahEdu/bricklib).  The C++ headers shipped beside
 * this file (brick.h, bricksetup.h, brickcompare.h, brick-hip.h) keep the
 * reference's template API and forward device work here.
 *
 * Conventions
 *   - Every function returns an int status (BK_OK == 0).  Nothing throws or
 *     exits across this boundary; bk_last_error() gives a message for the
 *     calling thread's last failure.
 *   - Pointers named d_* are device (HBM) pointers, h_* are host pointers.
 *   - Brick data: element e (0 <= e < brick_elems) of brick b of a field is
 *     at  field[(size_t)b * step + e]  where `field` is the storage base
 *     already offset to that field (BrickStorage + offset, brick.h:390-394).
 *   - Adjacency: uint32 adj[nbricks][27]; slot kd*9 + jd*3 + id with
 *     d = offset + 1 in {0,1,2}; slot 13 is the brick itself
 *     (brick.h:96-127, accessor brick.h:235-246).
 *   - All launches are stream-ordered on the stream argument (NULL = the
 *     device's null stream).  No entry point that launches kernels
 *     allocates device memory, except bk_sweeps with use_graph != 0, which
 *     builds and caches an executable graph on first use.
 */
#ifndef BK_API_H
#define BK_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BK_API_VERSION 3

enum bk_status {
  BK_OK = 0,
  BK_ERR_ARG = 1,         /* invalid argument (null pointer, bad size, bad enum) */
  BK_ERR_NOMEM = 2,       /* device or host allocation failed */
  BK_ERR_HIP = 3,         /* HIP runtime error (message in bk_last_error) */
  BK_ERR_UNSUPPORTED = 4, /* combination not built (e.g. BF16 cube) */
  BK_ERR_NODEV = 5        /* no HIP device visible */
};

/* Element type of a brick field.  bElem (vecscatter.h:12-14) is double in
 * the reference; FP32 and BF16 storage are additions. */
enum bk_dtype { BK_F64 = 0, BK_F32 = 1, BK_BF16 = 2 };

/* Stencil kinds.  Coefficient layouts (ncoeff values, double):
 *   STAR7/13/25 (radius R = 1/2/4): c[0] centre, then for r = 1..R the six
 *     taps (i+r, i-r, j+r, j-r, k+r, k-r) at c[6(r-1)+1 .. 6(r-1)+6]
 *     -- the order of stencils/7pt.py:19-26, mpi13pt.py:18-31,
 *     mpi25pt.py:20-45 (7pt: coeff[1] is i+1 ... coeff[6] is k-1).
 *   CUBE125 (radius 2): c[(dk+2)*25 + (dj+2)*5 + (di+2)], 125 values
 *     (the 5x5x5 offsets of mpi125pt.py:12-162 with one coefficient each).
 *   COND7: the 7 STAR7 coefficients; p = sum c_n * max(x_n, 0) and
 *     out = (s > t) ? p : -p, where (s, t) = (p, 0) in the reference form
 *     (cond.py:23-32, 3axis.cpp:132-137) and (centre, theta) in the
 *     threshold form.                                                    */
enum bk_stencil {
  BK_STAR7 = 0,
  BK_STAR13 = 1,
  BK_STAR25 = 2,
  BK_CUBE125 = 3,
  BK_COND7 = 4
};

enum bk_cond_mode { BK_COND_REFERENCE = 0, BK_COND_THRESHOLD = 1 };

/* Brick shape and vector fold of a field: Brick<Dim<BK,BJ,BI>,Dim<FJ,FI>>
 * (brick.h:353-395).  Element (k, j, i) of a brick sits at
 *   vec*64 + (j % FJ)*FI + i % FI,  vec = (k*(BJ/FJ) + j/FJ)*(BI/FI) + i/FI
 * (the accessor's fold decomposition, brick.h:235-246, 276-284).  Every
 * shape holds 512 elements; grids count bricks per axis as usual, so a
 * 4x8x16 grid of n^3 cells has n/16 x n/8 x n/4 bricks (i, j, k).      */
enum bk_layout {
  BK_LAYOUT_8x8x8 = 0,      /* Dim<8,8,8>,  Dim<8,8>:  k*64 + j*8 + i          */
  BK_LAYOUT_4x8x16 = 1,     /* Dim<4,8,16>, Dim<4,16>: k*128 + j*16 + i        */
  BK_LAYOUT_4x8x16_F88 = 2  /* Dim<4,8,16>, Dim<8,8>:  k*128 + (i/8)*64 + j*8 + i%8 */
};

typedef void *bk_stream; /* hipStream_t */
typedef void *bk_event;  /* hipEvent_t */
typedef struct bk_plan_s *bk_plan; /* execution plan (bk_plan_create) */
typedef struct bk_batch_s *bk_batch; /* batched sweeps (bk_batch_create) */
typedef struct bk_arena_s *bk_arena; /* field arena (bk_arena_create) */

enum bk_plan_shape { BK_PLAN_COLUMN = 1, BK_PLAN_QUAD = 4 };

/* ---- library, errors, device ------------------------------------------ */
int bk_version(void);
const char *bk_last_error(void);
/* Number of coefficients a stencil takes (0 for an unknown kind). */
int bk_stencil_ncoeff(int stencil);
/* Brick extents of a layout: ext[0] = BK (k), ext[1] = BJ, ext[2] = BI;
 * BK_ERR_ARG for an unknown layout. */
int bk_layout_extents(int layout, int ext[3]);
/* Radius of a stencil (max |offset| along an axis), 0 for an unknown kind. */
int bk_stencil_radius(int stencil);
int bk_dtype_size(int dtype);
int bk_device_count(int *n);
int bk_set_device(int dev);
int bk_device_sync(void);
/* *is_device = 1 when p is device (or managed) memory a kernel may read,
 * 0 for host memory.  No counterpart in the reference (its accessor kernels
 * simply take what they are given); the C++ headers use it to refuse host
 * pointers before a launch rather than fault the GPU. */
int bk_pointer_is_device(const void *p, int *is_device);

/* ---- memory: device-resident BrickStorage ----------------------------- */
/* BrickStorage::allocate(chunks, step) (brick.h:68-75) with the data in HBM:
 * chunks * step elements of `dtype`, 2048-byte aligned like the reference's
 * ALIGN (brick.h:15).  Fails with BK_ERR_ARG when the byte count overflows. */
int bk_storage_alloc(int dtype, long chunks, size_t step, void **d_dat);
int bk_free(void *d_ptr);
int bk_alloc(size_t bytes, void **d_ptr);
/* Stream-ordered allocation and release (hipMallocAsync / hipFreeAsync):
 * no device-wide synchronisation, for loops that admit and retire grids
 * while other grids are being swept.  Memory freed this way stays in the
 * device's default pool for reuse (release threshold set to the maximum on
 * the first bk_alloc_async of each device). */
int bk_alloc_async(size_t bytes, void **d_ptr, bk_stream s);
int bk_free_async(void *d_ptr, bk_stream s);
/* Field arena for grids that come and go (no counterpart in the reference,
 * whose grids live for the whole run).  Blocks are carved best-fit out of
 * slabs of at least slab_bytes (hipMalloc on the current device, kept until
 * bk_arena_destroy; capacity bounds their total, 0 = no bound), 2 KiB
 * aligned; a freed block merges with its free neighbours.  No driver call
 * per block.  A block freed while queued kernels still use it may be handed
 * out again at once: its next user must work on the same stream (or wait
 * for it).  Thread-safe. */
int bk_arena_create(size_t slab_bytes, size_t capacity, bk_arena *out);
int bk_arena_destroy(bk_arena a);
int bk_arena_alloc(bk_arena a, size_t bytes, void **d_ptr);
int bk_arena_free(bk_arena a, void *d_ptr);
int bk_arena_stats(bk_arena a, size_t *used, size_t *reserved, size_t *largest_free,
                   uint32_t *live_blocks);
/* Host<->device copies: movBrickStorage (brick-gpu.h:84-103) and
 * movBrickInfo (brick-gpu.h:42-55) are built on these.  Sizes are size_t
 * (the reference computes them in `unsigned`, brick-gpu.h:47,89). */
int bk_memcpy_h2d(void *d_dst, const void *h_src, size_t bytes, bk_stream s);
int bk_memcpy_d2h(void *h_dst, const void *d_src, size_t bytes, bk_stream s);
int bk_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes, bk_stream s);
/* Streaming device copy (16-byte aligned, non-overlapping).  A measurement
 * utility with no counterpart in the reference: the benchmark uses it as
 * the achievable-HBM-bandwidth yardstick for the memory-bound sweeps. */
int bk_stream_copy(void *d_dst, const void *d_src, size_t bytes, bk_stream s);
/* Register-only multiply-add loop: waves_per_simd one-wave workgroups per
 * SIMD of the current device, each lane running `iters` iterations of 64
 * fused multiply-adds (dtype BK_F64 or BK_F32) over `chains` (1 or 8)
 * independent accumulators with a wave-uniform coefficient, then storing
 * one value per lane to d_out (capacity out_elems elements: at least
 * waves_per_simd * CUs * 4 * 64).  *flops receives the launch's flop count
 * for the caller's own event timing.  A measurement utility with no
 * counterpart in the reference: the benchmark uses it as the sustained-VALU
 * yardstick for the compute-bound 125-point cube. */
int bk_valu_probe(int dtype, int chains, int waves_per_simd, uint32_t iters, void *d_out,
                  size_t out_elems, double *flops, bk_stream s);
/* Zero-fill with a kernel (no hipMemset node, so it is safe in graphs). */
int bk_zero(void *d_dst, size_t bytes, bk_stream s);

/* ---- streams and events (hipEvent timing, stencils.hip.h:13-31) ------- */
int bk_stream_create(bk_stream *s);
int bk_stream_destroy(bk_stream s);
int bk_stream_sync(bk_stream s);
int bk_event_create(bk_event *e);
int bk_event_destroy(bk_event e);
int bk_event_record(bk_event e, bk_stream s);
int bk_event_sync(bk_event e);
int bk_event_elapsed_ms(bk_event start, bk_event stop, float *ms);

/* ---- grid numbering and adjacency (host) ------------------------------- */
/* init_grid<3> (bricksetup.h:73-89): bricks of a dims[0] x dims[1] x
 * dims[2] brick grid (i fastest, dims contiguous-first) numbered
 * grid[pos] = pos, and adj[pos][27] filled from that numbering.  With
 * periodic == 0 a neighbour outside the grid is brick 0, exactly as the
 * reference (bricksetup.h:42-47, which also wraps edge rows); with
 * periodic != 0 every neighbour coordinate is taken modulo dims.
 * h_grid: dims[0]*dims[1]*dims[2] uint32, h_adj: 27 * that many uint32. */
int bk_init_grid(const long dims[3], int periodic, uint32_t *h_grid,
                 uint32_t *h_adj);
/* Periodic ghost-brick pairs for a grid of nb[0] x nb[1] x nb[2] interior
 * bricks surrounded by one layer of ghost bricks ((nb+2)^3 numbered as
 * bk_init_grid).  Writes (dst, src) pairs: every ghost brick and the
 * interior brick that is its periodic image (the one-rank periodic exchange
 * of BrickDecomp::exchange, brick-mpi.h:466-495, over the periodic
 * Cartesian communicator of weak/args.cpp:107-108).  h_pairs may be NULL to
 * query *npairs. */
int bk_ghost_pairs(const long nb[3], uint32_t *h_pairs, uint32_t *npairs);
/* The inverse of bk_ghost_pairs: for every brick b of the padded grid
 * ((nb+2)^3 numbered as bk_init_grid), the ghost bricks whose periodic image
 * it is, as CSR -- h_img[h_off[b] .. h_off[b+1]) -- in pair order.  h_off
 * has nbricks + 1 entries, h_img *nimg (= the pair count).  Both NULL to
 * query *nimg.  Feeds the fused refresh of bk_sweeps_fused. */
int bk_ghost_images(const long nb[3], uint32_t *h_off, uint32_t *h_img,
                    uint32_t *nimg);
/* The BrickDecomp layout (BrickDecomp::initialize, brick-mpi.h:322-460,
 * with the skin3d_good surface order, src/brick-mpi.cpp:25-52) of one
 * subdomain of nb bricks per axis (i, j, k) with a ghost zone of
 * ghost_bricks bricks: ids grouped as inner bricks, skin, ghost, each
 * neighbour's ghost bricks contiguous.  Bricks of brick_bytes are grouped
 * to whole pages of page_bytes (ids start at page/brick when that exceeds 1,
 * regions padded to it), as the reference does for mmap.  Outputs:
 *   h_grid   (nb+2g)^3 ids, i fastest (the grid array bDecomp[k][j][i]);
 *   h_adj    nbricks x 27 adjacency, out-of-range neighbours = brick 0;
 *   sep_pos  first skin id, first ghost id, end of ghost ids;
 *   h_regions one entry per contiguous exchange run: the ghost bricks
 *            [ghost_pos, +len) owned by neighbour `neighbor` and the skin
 *            bricks [skin_pos, +len) the opposite neighbour receives; on a
 *            periodic single subdomain the first run is a copy of the second.
 * Any output array may be NULL; call with NULLs to size them (*nbricks,
 * *nregions). */
typedef struct bk_decomp_region {
  int32_t neighbor[3];  /* -1, 0, +1 per axis (i, j, k) */
  uint32_t ghost_pos;
  uint32_t skin_pos;
  uint32_t len;
  uint32_t skin_st;     /* skin-list positions [skin_st, skin_ed) */
  uint32_t skin_ed;
  uint32_t first_pad;   /* padding bricks at the run's start / end */
  uint32_t last_pad;
} bk_decomp_region;
int bk_decomp_layout(const long nb[3], int ghost_bricks, size_t brick_bytes,
                     size_t page_bytes, uint32_t *h_grid, uint32_t *h_adj,
                     uint32_t *nbricks, uint32_t sep_pos[3],
                     bk_decomp_region *h_regions, uint32_t *nregions);
/* The same layout for any surface order, as BrickDecomp::initialize takes
 * a skin list (brick-mpi.h:322; the reference ships skin3d_good,
 * skin3d_normal and skin3d_bad, src/brick-mpi.cpp:25-110).  h_skin holds
 * nskin entries in order, three int32 per entry, one per axis (i, j, k):
 * +1, -1 or 0 (the reference's BitSet {1, -3} is (+1, 0, -1)).  Each of
 * the 26 non-empty regions must appear once; empty entries (0, 0, 0) may
 * appear anywhere, as in skin3d_normal and skin3d_bad: they take no ids but
 * end runs and move padding, as in the reference.  26 <= nskin <= 1024,
 * BK_ERR_ARG otherwise.  h_skin_size (nskin entries, may be NULL) receives
 * the reference's skin_size: the ids each entry takes, padding included. */
int bk_decomp_layout_skin(const long nb[3], int ghost_bricks, size_t brick_bytes,
                          size_t page_bytes, const int32_t *h_skin, int nskin, uint32_t *h_grid,
                          uint32_t *h_adj, uint32_t *nbricks, uint32_t sep_pos[3],
                          bk_decomp_region *h_regions, uint32_t *nregions,
                          uint32_t *h_skin_size);
/* Interior brick list in grid order (k slowest, i fastest) for the same
 * padded grid: the bricks the sweep loops of stencils/3axis.cpp:51-66 visit.
 * h_list may be NULL to query *n. */
int bk_interior_list(const long nb[3], uint32_t *h_list, uint32_t *n);

/* The interior bricks of a grid of nb bricks per axis with `ghost` (0 or 1)
 * ghost layers, numbered as bk_init_grid over nb + 2*ghost, in the order
 * the sweeps visit them best: the (i, j) plane cut into tile_i x tile_j
 * brick tiles, each swept through all k before the next (k-neighbours are
 * then tile_i*tile_j list positions apart instead of a whole brick plane;
 * 0 means the whole axis, so 0, 0 is grid order with k slowest).  The
 * reference loops over grid[k][j][i] (stencils/3axis.cpp:51-66); any order
 * gives the same results.  h_list may be NULL to query *n. */
int bk_tiled_list(const long nb[3], int ghost, uint32_t tile_i, uint32_t tile_j,
                  uint32_t *h_list, uint32_t *n);

/* ---- execution plans --------------------------------------------------- */
/* A plan is a schedule for sweeping the bricks of a list: it groups them,
 * through the host copy of the adjacency, into units one wave marches
 * through along k -- k-columns of up to kcol bricks (BK_PLAN_COLUMN) or
 * up to kcol levels of 2x2 blocks of columns (BK_PLAN_QUAD; members
 * anchor, adj[anchor][14], adj[anchor][16], adj[adj[anchor][14]][16]).
 * Every listed brick is written exactly once by a sweep that uses the plan;
 * bricks a unit computes but that are not listed are not written.  Columns
 * link b to adj[b][k+1] only where adj[adj[b][k+1]][k-1] == b.  A quad plan
 * requires every listed brick's 27 neighbours to be where the quad's 4x4
 * column table puts them (true on periodic grids and on the interior of a
 * padded grid, not on the outer layer of a non-periodic one): otherwise
 * BK_ERR_UNSUPPORTED.  A cube sweep through a quad plan addresses its input
 * with 32-bit element offsets: nbricks * step must not exceed 2^32
 * (BK_ERR_UNSUPPORTED from bk_sweep / bk_batch_set otherwise).  The
 * reference has no counterpart: its sweep is a loop over grid[k][j][i]
 * (stencils/3axis.cpp:51-66, weak/main.hip.cpp:274-279) and the plan is
 * how that loop is scheduled on the GPU.  Allocates device memory.
 * Kernel choice follows the plan: star stencils run a column plan with
 * kcol <= 4 as k-stacks (one workgroup loads the stacked bricks together
 * and shares their k-faces in LDS) and with kcol > 4 as k-marches (one wave
 * walks the column, prefetching the next brick); star stencils run a quad
 * plan as 2x2 blocks of k-marching columns, one workgroup per block whose
 * four waves share an LDS tile of the block (only its outer faces are
 * fetched); the cube streams a quad plan plane by plane; a sweep without a
 * plan runs one brick per wave.  Results do not depend on the choice. */
int bk_plan_create(const uint32_t *h_adj, uint32_t nbricks,
                   const uint32_t *h_list, uint32_t nlist, int shape,
                   uint32_t kcol, bk_plan *plan);
int bk_plan_destroy(bk_plan plan);
/* The same schedule built on the host only (no device memory): h_anchor and
 * h_len hold nunits <= nlist entries, h_mask nunits*kcol, h_flat ncovered
 * (== the number of distinct listed bricks).  Any array may be NULL to
 * query the counts first. */
int bk_plan_build_host(const uint32_t *h_adj, uint32_t nbricks,
                       const uint32_t *h_list, uint32_t nlist, int shape,
                       uint32_t kcol, uint32_t *h_anchor, uint32_t *h_len,
                       uint32_t *h_mask, uint32_t *h_flat, uint32_t *nunits,
                       uint32_t *ncovered);
int bk_plan_query(bk_plan plan, uint32_t *nunits, uint32_t *kcol, int *shape,
                  uint32_t *ncovered);
/* Copy the plan's arrays to the host (any pointer may be NULL):
 * anchor[nunits], len[nunits], mask[nunits*kcol], flat[ncovered]. */
int bk_plan_download(bk_plan plan, uint32_t *h_anchor, uint32_t *h_len,
                     uint32_t *h_mask, uint32_t *h_flat);

/* ---- the sweep -------------------------------------------------------- */
typedef struct bk_sweep_desc {
  int stencil;            /* enum bk_stencil */
  int dtype;              /* enum bk_dtype of both fields */
  const uint32_t *d_adj;  /* [nbricks_total][27] adjacency in HBM */
  const uint32_t *d_list; /* brick indices to compute, in HBM */
  uint32_t nlist;         /* entries in d_list */
  const void *d_in;       /* input field base (already offset) */
  size_t in_step;         /* elements between consecutive bricks of d_in */
  void *d_out;            /* output field base (already offset) */
  size_t out_step;
  const double *h_coeff;  /* host coefficients, bk_stencil_ncoeff values */
  double theta;           /* COND7 threshold (threshold form) */
  int cond_mode;          /* enum bk_cond_mode */
  bk_plan plan;           /* optional: when set, the sweep covers the plan's
                             bricks and d_list / nlist are ignored */
  int layout;             /* enum bk_layout of both fields (0: 8x8x8) */
  int wg_bricks;          /* bricks (waves) per workgroup of the per-brick
                             kernels: 0 (default, 4), 1, 2, 4 or 8 (8: the
                             8x8x8 stars only) */
  int lane_shift;         /* 1: the star/cond kernels of a quad plan take the
                             i-direction taps from DPP lane shifts of the
                             wave's registers (same bits as 0, which reads
                             them from the LDS tile); quad plans only */
} bk_sweep_desc;

/* One stencil sweep: for every brick b in d_list (or in the plan) and
 * every element of b,
 * out[b] = stencil(in) read through the adjacency.  This is the loop of
 * stencils/3axis.cpp:51-66 (7pt), :123-140 (cond), and the brick(...)
 * kernel of weak/main.hip.cpp:32-40 (13/25/125-point via
 * stencils/mpi*.py), for 8x8x8 bricks in the 8x8 vector fold or, with
 * d->layout, 4x8x16 bricks in the 4x16 or 8x8 fold (no plan: one brick per
 * wave, d->wg_bricks bricks per workgroup).
 * Accumulation order per point (fixed, so results are reproducible):
 * acc = c_t0 * x_t0, then acc = fma(c_t, x_t, acc) over the remaining taps,
 * in the coefficient order above for stars and (dk, dj, di) lexicographic
 * for the cube.  FP64 accumulates in FP64; FP32 and BF16 in FP32. */
int bk_sweep(const bk_sweep_desc *d, bk_stream s);

/* Periodic ghost refresh on the device: for each (dst, src) pair, copy the
 * brick_elems elements of brick src onto brick dst in the field d_dat.
 * Replaces the MPI exchange of weak/main.hip.cpp:249-269 for one rank. */
int bk_ghost_refresh(int dtype, void *d_dat, size_t step, size_t brick_elems,
                     const uint32_t *d_pairs, uint32_t npairs, bk_stream s);

/* nsweeps ping-pong sweeps with a periodic ghost refresh of the current
 * input before each sweep (weak/main.hip.cpp:274-279 with the exchange of
 * :249-269 done on the device).  Sweep 0 reads d_in and writes d_out,
 * sweep 1 reads d_out and writes d_in, ...; the result is in d_out when
 * nsweeps is odd and in d_in when it is even.  With use_graph != 0 the
 * whole sequence is captured once into a hipGraph (cached per argument set)
 * and replayed. */
int bk_sweeps(const bk_sweep_desc *d, const uint32_t *d_pairs, uint32_t npairs,
              int nsweeps, int use_graph, bk_stream s);
/* The communication-avoiding form of the same loop (weak/main.cpp:26-36,
 * :203-211, where brick_stencil(..., skip = 0) sweeps every brick of the
 * padded grid and the last sweep per exchange uses skip = 1): the ghost
 * bricks are refreshed once per ghost_iters sweeps; the sweeps in between
 * cover every brick in d_all (or all_plan), ghost bricks included, and the
 * last sweep of each group and of the sequence covers d->d_list (or
 * d->plan).  Stale ghost values move inwards by the stencil radius per
 * sweep, so the interior is exact -- bit for bit the bk_sweeps result --
 * when radius * ghost_iters <= 8 (the one-brick ghost zone; 4 for the
 * 4x8x16 layouts, whose bricks are 4 deep along k); larger values
 * are rejected.  ghost_iters = 1 is bk_sweeps.  all_plan: a column plan
 * over d_all, or NULL (one brick per wave); bk_plan_create rejects quad
 * plans over the outer layer of a non-periodic grid. */
int bk_sweeps_deep(const bk_sweep_desc *d, const uint32_t *d_all, uint32_t nall,
                   bk_plan all_plan, const uint32_t *d_pairs, uint32_t npairs,
                   int nsweeps, int ghost_iters, int use_graph, bk_stream s);
/* bk_sweeps with the refresh fused into the sweep: the ghost bricks of
 * d_in are refreshed once (d_pairs), then every sweep also stores each
 * output brick into its ghost images (d_img_off / d_img from
 * bk_ghost_images, uploaded), so the next sweep's input is already
 * refreshed.  The same bits as bk_sweeps, with one ghost write per sweep
 * instead of a refresh's read + write.  Star and cond stencils without a
 * plan (one brick per wave) or with a quad plan; otherwise
 * BK_ERR_UNSUPPORTED. */
int bk_sweeps_fused(const bk_sweep_desc *d, const uint32_t *d_img_off,
                    const uint32_t *d_img, const uint32_t *d_pairs, uint32_t npairs,
                    int nsweeps, int use_graph, bk_stream s);
/* Drop every cached graph (frees their device resources). */
int bk_graph_cache_clear(void);

/* ---- autotuner ---------------------------------------------------------- */
/* One timed schedule: brick layout, plan (0: none, one brick per wave) and
 * its length, bricks per workgroup of the per-brick kernel, and the best
 * measured milliseconds per sweep (refresh included on a ghost layout). */
typedef struct bk_tune_cand {
  int layout;      /* enum bk_layout */
  int plan_shape;  /* 0, BK_PLAN_COLUMN or BK_PLAN_QUAD */
  uint32_t kcol;   /* plan length (0 without a plan) */
  int wg_bricks;   /* bricks per workgroup without a plan (0 with one) */
  int lane_shift;  /* bk_sweep_desc.lane_shift of the schedule */
  float ms;
} bk_tune_cand;

/* flags of bk_tune_sweep: the layouts to try (bit l = enum bk_layout l;
 * at least one), and a restriction to the schedules the batched kernels
 * run (8x8x8 only: per brick with 4 bricks per workgroup, k-columns or
 * 2x2 brick blocks for the stars, 2x2 column quads for the cube). */
#define BK_TUNE_LAYOUT_8x8x8 0x1u
#define BK_TUNE_LAYOUT_4x8x16 0x2u
#define BK_TUNE_LAYOUT_4x8x16_F88 0x4u
#define BK_TUNE_ALL_LAYOUTS 0x7u
#define BK_TUNE_BATCH_SCHEDULES 0x100u

/* The autotuner (no counterpart in the reference, whose brick shape, fold
 * and launch shape are fixed at compile time: brick.h:353-395,
 * stencils/stencils.h:10-12, 3axis.hip.cpp:16-41).  On a periodic grid of
 * cells[0] x cells[1] x cells[2] cells (i, j, k) with `ghost` (0: wrapped
 * adjacency, 1: one ghost-brick layer refreshed before every sweep) it
 * builds the grid in every selected layout the extents fit, runs every
 * schedule of that layout -- per brick with 1, 2, 4 or 8 bricks per
 * workgroup (8: 8x8x8 stars only; 4: the cube on 8x8x8), k-columns of
 * 4/8/16 bricks and 2x2 column quads of 1/2/4/8 levels (8x8x8 stars; each
 * also with lane_shift = 1 unless BK_TUNE_BATCH_SCHEDULES), quads of 2/4/6/8
 * levels (8x8x8 cube) -- untimed for warm_s seconds, then times `steps` ping-pong sweeps
 * per candidate (hipGraph replays) in `rounds` interleaved rounds, keeping
 * each candidate's best.  h_out receives up to max_out candidates fastest
 * first (h_out[0] is the choice); with h_out NULL *nout is the candidate
 * count and nothing runs.  Allocates and frees its own grids; the result
 * is also kept for bk_tune_lookup.  All schedules compute the same bits. */
int bk_tune_sweep(int stencil, int dtype, const double *h_coeff, double theta,
                  int cond_mode, const long cells[3], int ghost, int steps,
                  int rounds, double warm_s, unsigned flags, bk_tune_cand *h_out,
                  uint32_t max_out, uint32_t *nout, bk_stream s);
/* The fastest candidate of an earlier bk_tune_sweep with the same stencil,
 * dtype, cells, ghost and flags; BK_ERR_ARG when there is none. */
int bk_tune_lookup(int stencil, int dtype, const long cells[3], int ghost,
                   unsigned flags, bk_tune_cand *best);

/* ---- batched sweeps over many grids ------------------------------------ */
/* One request of a batch: an independent grid with its own fields,
 * adjacency, stencil, dtype and coefficients.  A batch runs ticks; at tick t
 * every request with t < ticks refreshes the ghost bricks of its input
 * field (pairs from bk_ghost_pairs, may be none) and sweeps it into the
 * other field: the input is f0 when (parity + t) is even, f1 otherwise.
 * Stars sweep the bricks of d_list, or, when `plan` is set, the bricks of
 * that plan by the k-marching column schedule (BK_PLAN_COLUMN) or in 2x2
 * brick blocks (BK_PLAN_QUAD); the cube sweeps the bricks of `plan`, which
 * must be a BK_PLAN_QUAD plan.  Either way the results are those of bk_sweep.  This is the weak-scaling loop of
 * weak/main.hip.cpp:243-284 for many subdomains at once, the way the
 * strong-scaling driver's one launch over many subdomains does
 * (strong/main.cu:85-99), with grids of any size, stencil and dtype. */
typedef struct bk_batch_req {
  int stencil;
  int dtype;
  const uint32_t *d_adj;
  const uint32_t *d_list; /* stars without a plan: bricks to compute */
  uint32_t nlist;
  bk_plan plan;           /* cube: BK_PLAN_QUAD plan; stars: a column or quad plan, or NULL */
  const uint32_t *d_pairs; /* ghost (dst, src) pairs, or NULL */
  uint32_t npairs;
  void *d_f0;
  void *d_f1;
  size_t step;            /* elements per brick in both fields */
  const double *h_coeff;  /* copied by bk_batch_set */
  double theta;
  int cond_mode;
  uint32_t parity;        /* 0: f0 holds the input at tick 0 */
  uint32_t ticks;         /* ticks of the next bk_batch_run it takes part in */
} bk_batch_req;

/* A batch serving up to max_reqs requests (<= 2048): one persistent kernel
 * per (stencil, dtype, schedule) group and per tick, each over the concatenated work
 * of all requests in the group, plus one persistent refresh kernel per
 * dtype.  Allocates its tables (HBM and pinned host memory). */
int bk_batch_create(uint32_t max_reqs, bk_batch *batch);
int bk_batch_destroy(bk_batch batch);
/* Replace the batch's requests (stream-ordered copies of the tables and
 * coefficients; the host arrays may be reused when the call returns).  The
 * device pointers in the requests must stay valid until the runs that use
 * them have completed. */
int bk_batch_set(bk_batch batch, const bk_batch_req *reqs, uint32_t n, bk_stream s);
/* Run nticks ticks.  With use_graph != 0 the launches are captured once per
 * (nticks, set of non-empty groups) and replayed; the tables are read when
 * the kernels run, so later bk_batch_set calls need no new capture. */
int bk_batch_run(bk_batch batch, uint32_t nticks, int use_graph, bk_stream s);

/* ---- serving loop ------------------------------------------------------ */
/* Continuous batching of stencil solves on one GPU (no counterpart in the
 * reference; its closest precedent is the strong-scaling driver's one
 * launch over many subdomains, strong/main.cu:85-99).  Requests of any grid
 * size, stencil, dtype and sweep count are admitted as they arrive against
 * an HBM budget (two fields each, from a field arena, resident until the
 * request completes), swept together chunk by chunk through bk_batch_run on
 * one or two lanes (streams), and retired as they complete, freeing their
 * memory for the next arrivals. */
typedef struct bk_server_s *bk_server;

typedef struct bk_serve_config {
  double hbm_budget;    /* bytes of fields resident at once */
  uint32_t chunk;       /* ticks per batched run between admissions (16) */
  uint32_t max_active;  /* requests resident at once, <= 2048 */
  int lanes;            /* 1 or 2 batches on their own streams */
  int use_graph;        /* replay each chunk's launches as a hipGraph */
  int ghost;            /* 0: wrapped adjacency (no refresh), 1: ghost layer */
  int pipeline;         /* queue the next chunk before waiting for the last */
  int tune;             /* schedules from bk_tune_sweep per (size class,
                           stencil, dtype); 0: fixed defaults -- cubes in
                           quads of 2 levels, stars per star_plan; a quad
                           schedule of L > 1 levels runs L/2 in a chunk where
                           the lane's units of that stencil and dtype number
                           fewer than four times what the batched kernel
                           holds at once */
  size_t slab_bytes;    /* arena slab size (0: 1 GiB) */
  int star_plan;        /* fixed default for stars and cond: 0 one brick
                           per wave, BK_PLAN_QUAD 2x2 brick blocks (levels
                           per stencil and dtype), BK_PLAN_COLUMN k-columns
                           of 8 */
} bk_serve_config;

typedef struct bk_serve_req {
  long nb[3];            /* interior bricks per axis (i, j, k) */
  int stencil;           /* enum bk_stencil */
  int dtype;             /* enum bk_dtype */
  uint32_t steps;        /* sweeps */
  uint64_t seed;         /* input: bk_fill_uniform_grid(seed, -1, 1) */
  const double *h_coeff; /* copied by bk_server_submit */
  double theta;
  int cond_mode;
  double arrival;        /* seconds on the server clock (< 0: now) */
  int keep;              /* keep the result field until bk_server_release */
} bk_serve_req;

typedef struct bk_serve_info {
  int state;            /* 0 pending, 1 active, 2 done (fields kept), 3 done */
  double arrival, admitted, finished; /* server clock, seconds */
  uint32_t sweeps_done;
  size_t field_bytes;
  const void *d_result; /* state 2: the field holding the result */
  int plan_shape;       /* the schedule of its last chunk: 0 (per brick) or a
                           plan shape, and the plan's levels per unit */
  uint32_t kcol;
} bk_serve_info;

typedef struct bk_serve_stats {
  uint32_t requests;    /* completed since the last reset */
  double p50_s, p99_s, mean_s; /* latency = finished - arrival */
  double elapsed_s;     /* from the run's start to the last completion */
  double stencil_updates; /* sum of interior cells x sweeps */
  double gstencil_s;
  double peak_hbm_bytes;
  double arena_reserved_bytes;
  uint64_t chunks;
  uint32_t tuned_classes;
} bk_serve_stats;

/* Threads: submit, wait, info, result, release, stats and reset_stats may
 * be called from any thread at any time.  prewarm, run, start and stop are
 * the server's driver: call them from one thread, and prewarm and run only
 * while the server's own thread is not running.  result and release of the
 * same request must not overlap (release returns the fields result reads). */
int bk_server_create(const bk_serve_config *cfg, bk_server *out);
/* Stops the thread (bk_server_stop), frees every field and table. */
int bk_server_destroy(bk_server s);
/* Queue a request (thread-safe); *id numbers requests from 0 in submission
 * order.  A request whose fields exceed the budget is refused. */
int bk_server_submit(bk_server s, const bk_serve_req *r, uint32_t *id);
/* Build the shared tables (adjacency, brick list, ghost pairs, plans) of
 * these requests' grid shapes, and with tuning on time their size classes,
 * ahead of serving (the server's shape catalogue).  Nothing is queued. */
int bk_server_prewarm(bk_server s, const bk_serve_req *r, uint32_t n);
/* Serve every queued request to completion on the calling thread, the
 * clock starting now: with realtime != 0 a request is admitted once the
 * clock passes its arrival, otherwise all are present at once (a
 * saturating run). */
int bk_server_run(bk_server s, int realtime);
/* Or serve on the server's own thread: start it, submit from any thread,
 * wait for a request, stop (returns once everything queued is done). */
int bk_server_start(bk_server s);
int bk_server_stop(bk_server s);
int bk_server_wait(bk_server s, uint32_t id, double timeout_s);
int bk_server_info(bk_server s, uint32_t id, bk_serve_info *out);
/* Copy a completed kept request's result field (all its bricks, numbered
 * as bk_init_grid over nb + 2*ghost) to the host; bytes = field_bytes. */
int bk_server_result(bk_server s, uint32_t id, void *h_out, size_t bytes);
/* Return a kept request's fields to the arena. */
int bk_server_release(bk_server s, uint32_t id);
int bk_server_stats(bk_server s, bk_serve_stats *out);
int bk_server_reset_stats(bk_server s);

/* ---- seeded input and FP64 reductions ---------------------------------- */
/* Fill the interior of a padded brick grid (nb interior bricks per axis,
 * one ghost layer) with the seeded field value(x,y,z) = lo + (hi-lo) * u,
 * u = the 53-bit uniform of hash(seed, x + n0*(y + n1*z)) (n = 8*nb),
 * converted to dtype with round-to-nearest-even.  Ghost bricks untouched. */
int bk_fill_uniform(int dtype, void *d_dat, size_t step, const long nb[3],
                    uint64_t seed, double lo, double hi, bk_stream s);
/* The same values for a grid with `ghost` (0 or 1) ghost-brick layers:
 * ghost = 1 is bk_fill_uniform; ghost = 0 fills an nb[0] x nb[1] x nb[2]
 * brick grid numbered as bk_init_grid (the layout of a periodic grid whose
 * adjacency wraps, bk_init_grid with periodic != 0, which needs no ghost
 * refresh).  Interior cell (x, y, z) gets the same value in both layouts. */
int bk_fill_uniform_grid(int dtype, void *d_dat, size_t step, const long nb[3],
                         int ghost, uint64_t seed, double lo, double hi,
                         bk_stream s);
/* bk_fill_uniform_grid for any layout: nb interior bricks per axis of the
 * layout's shape, cell (x, y, z) of the nb[0]*BI x nb[1]*BJ x nb[2]*BK
 * interior gets the value bk_fill_uniform gives that cell. */
int bk_fill_uniform_layout(int layout, int dtype, void *d_dat, size_t step,
                           const long nb[3], int ghost, uint64_t seed, double lo,
                           double hi, bk_stream s);
/* FP64 reductions over the bricks in d_list (brick_elems each):
 *   out[0] = sum x, out[1] = sum x^2, out[2] = max |x|
 * written to the host array h_out after a stream sync. */
int bk_field_stats(int dtype, const void *d_dat, size_t step,
                   const uint32_t *d_list, uint32_t nlist, size_t brick_elems,
                   double h_out[3], bk_stream s);
/* bk_field_stats without any allocation: no counterpart in the reference
 * (it has no GPU reductions); for callers that cannot allocate device memory
 * -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The same kernels and the same results as bk_field_stats, but
 * the pass-1 partials live in a static 24 KiB device array of the library
 * (one per GPU, whatever the input size) and the three results are written
 * to device memory d_out[3], stream-ordered on s with no host wait.  Calls
 * that overlap on one GPU share that array and race: order them on one
 * stream (or wait between them). */
int bk_field_stats_noalloc(int dtype, const void *d_dat, size_t step,
                           const uint32_t *d_list, uint32_t nlist,
                           size_t brick_elems, double *d_out, bk_stream s);
/* FP64 comparison of field a against reference field r (dtypes may differ)
 * over d_list:  out[0] = max |a-r|,  out[1] = max |a-r| / max(|r|, floor),
 * out[2] = sum (a-r)^2,  out[3] = sum r^2,
 * out[4] = count of points with |a-r| > atol + rtol * |r| (compareBrick's
 * rule of brickcompare.h:35-39 with separate tolerances).  */
int bk_field_compare(int dtype_a, const void *d_a, size_t step_a, int dtype_r,
                     const void *d_r, size_t step_r, const uint32_t *d_list,
                     uint32_t nlist, size_t brick_elems, double floor,
                     double atol, double rtol, double h_out[5], bk_stream s);

/* ---- array <-> brick conversion on the device -------------------------- */
/* copyToBrick / copyFromBrick (bricksetup.h:172-221) on the device for
 * 8x8x8 bricks: the array is contiguous-first with extents
 * dims[a] + 2*(padding[a] + ghost[a]); the bricks of d_grid (a uint32 grid
 * of (dims[a] + 2*ghost[a]) / 8 bricks per axis) cover dims + 2*ghost.
 * Array and bricks share `dtype`. */
int bk_copy_to_brick(int dtype, const long dims[3], const long padding[3],
                     const long ghost[3], const void *d_arr,
                     const uint32_t *d_grid, void *d_dat, size_t step,
                     bk_stream s);
int bk_copy_from_brick(int dtype, const long dims[3], const long padding[3],
                       const long ghost[3], void *d_arr,
                       const uint32_t *d_grid, const void *d_dat, size_t step,
                       bk_stream s);

#ifdef __cplusplus
}
#endif

#endif /* BK_API_H */
