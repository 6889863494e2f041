// bk_star.hip -- star stencil sweeps (7/13/25-point) and the cond stencil
// over 8x8x8 bricks in the 8x8 vector fold, for gfx950.
//
// Replaces the per-brick kernels of stencils/3axis.hip.cpp:16-41 (7pt),
// :200-227 (cond) and the codegen brick kernel of weak/main.hip.cpp:32-40.
//
// Layout: element (k, j, i) of brick b is field[b*step + k*64 + j*8 + i]
// (the Brick<Dim<8,8,8>,Dim<8,8>> accessor, brick.h:235-246, 276-284).
//
// One wave64 per brick; a workgroup of four waves takes four consecutive
// entries of the brick list.  Lane L = j*8 + i owns the column (j, i) and
// walks k = 0..7:
//   - the brick's eight k-planes are eight coalesced 512-byte (FP64) loads,
//     kept in registers, and so are the R planes of the k-1 / k+1
//     neighbours: every k-direction tap is a register;
//   - the i/j taps come from an LDS tile [k][8+2R][8+2R] per wave holding
//     the brick plus the R-deep faces of its four i/j neighbours (one load
//     instruction per face layer);
//   - each point is evaluated as acc = c0*x0, then acc = fma(c, x, acc)
//     over (i+r, i-r, j+r, j-r, k+r, k-r) for r = 1..R -- the tap order of
//     the reference's stencil files, fixed so results are reproducible.
// Neighbour bricks are found through the adjacency table in HBM.
#include <type_traits>

#include "bk_device.h"
#include "bk_internal.h"

namespace bk {
namespace {

template <class T, int R, bool COND>
struct StarArgs {
  using A = typename acc_of<T>::type;
  const uint32_t *__restrict__ adj;
  const uint32_t *__restrict__ list;
  const T *__restrict__ in;
  T *__restrict__ out;
  uint64_t in_step;
  uint64_t out_step;
  uint32_t nlist;
  int cond_mode;
  // plan (column kernel only)
  const uint32_t *__restrict__ anchor;
  const uint32_t *__restrict__ len;
  const uint32_t *__restrict__ mask;
  const uint32_t *__restrict__ flat;
  const uint32_t *__restrict__ start;
  uint32_t nunits;
  uint32_t kcol;
  // fused refresh: ghost images of brick b, img[img_off[b] .. img_off[b+1])
  const uint32_t *__restrict__ img_off;
  const uint32_t *__restrict__ img;
  A theta;
  A c[6 * R + 1];
};

template <class A, bool COND>
__device__ __forceinline__ A tapv(A x) {
  if constexpr (COND) return relu_ref(x);
  else return x;
}

// Lane shifts inside the 16-lane DPP rows (two i-rows of the 8x8 plane
// fold): row_shl:n gives lane l the value of lane l + n, row_shr:n that of
// lane l - n, for 1 <= n <= 7 (lanes whose source leaves the row read 0 and
// are never used below).  64-bit values move as two 32-bit halves.
template <bool UP>
__device__ __forceinline__ uint32_t row_shift32(uint32_t v, int n) {
  // n is a constant after unrolling: the switch folds to one DPP move
  switch (n) {
#define BK_ROWSH(N)                                                              \
  case N:                                                                        \
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, UP ? 0x100 + N : 0x110 + N, \
                                              0xf, 0xf, true);
    BK_ROWSH(1) BK_ROWSH(2) BK_ROWSH(3) BK_ROWSH(4) BK_ROWSH(5) BK_ROWSH(6) BK_ROWSH(7)
#undef BK_ROWSH
    default: return 0u;
  }
}

template <bool UP>
__device__ __forceinline__ float row_shift(float x, int n) {
  return __builtin_bit_cast(float, row_shift32<UP>(__builtin_bit_cast(uint32_t, x), n));
}

template <bool UP>
__device__ __forceinline__ double row_shift(double x, int n) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = row_shift32<UP>((uint32_t)u, n), hi = row_shift32<UP>((uint32_t)(u >> 32), n);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The i + r (UP) or i - r tap of lane (j, i) of an 8x8 plane x, 1 <= r <= 7,
// from registers: inside the brick a shift of x by r lanes; past its edge
// the neighbour brick's cell, taken from nb = that brick's cells at the
// same (j, i) (the i + 8 or i - 8 cells), shifted by 8 - r the other way.
// Both shifts run with every lane active and a select picks one: a DPP move
// reads a source lane that EXEC disables as 0, so neither may sit in a
// branch on i.
template <bool UP, class A>
__device__ __forceinline__ A itap_shift(A x, A nb, int i, int r) {
  const A in_brick = row_shift<UP>(x, r);
  const A past_edge = row_shift<!UP>(nb, 8 - r);
  return (UP ? i + r < 8 : i - r >= 0) ? in_brick : past_edge;
}

// LDS tile of one brick and its R-deep i/j faces: planes k of P x P cells
// (P = 8 + 2R) of ES bytes, element (k, y, x) at at(k, y, x).  For R = 4
// (16-wide rows) bit 3 of x is flipped so that the rows one read pass
// serves land on distinct banks: a pass of 4-byte reads serves 32 lanes =
// four consecutive rows of 8 cells, and the flip by bit 1 of y (xor the
// plane's parity) separates rows y and y+2, which a 16-cell stride would put
// on the same banks; a pass of 8-byte reads serves 16 lanes = two rows, and
// the flip by bit 0 of y separates them.  The 25-point sweep ran 13% (FP64)
// and 23% (FP32) faster with the first form, and FP64 another 8% with the
// second.  R = 1, 2 keep P-wide rows: padding them to 16 for the swizzle
// cost more in occupancy than it saved.
template <int R, int ES>
struct Tile {
  static constexpr int P = 8 + 2 * R;
  static constexpr bool kSwizzle = R == 4;
  static constexpr int RS = kSwizzle ? 16 : P;  // row stride
  static constexpr int PL = P * RS;             // plane stride
  __device__ static constexpr int flip(int k, int y) {
    return ES >= 8 ? (y & 1) : (((y >> 1) ^ k) & 1);
  }
  __device__ static constexpr int at(int k, int y, int x) {
    return k * PL + y * RS + (kSwizzle ? (x ^ (flip(k, y) << 3)) : x);
  }
};

// The R-deep i/j faces of a brick's four in-plane neighbours, R values per
// lane per face.  Lane -> (c = lane & 7, rr = (lane >> 3) % R,
// kq = (lane >> 3) / R); value q belongs to plane q * (8 / R) + kq.  A
// half-wave then covers whole rows of R-deep face runs: contiguous in
// memory (R doubles of an i-face row, 8 of a j-face row) and on distinct
// LDS banks.
template <class T, int R>
struct Faces {
  using A = typename acc_of<T>::type;
  A jm[R], jp[R], im[R], ip[R];

  __device__ __forceinline__ void load(const T *__restrict__ pjm, const T *__restrict__ pjp,
                                       const T *__restrict__ pim, const T *__restrict__ pip,
                                       int lane) {
    const int c = lane & 7, rr = (lane >> 3) % R, kq = (lane >> 3) / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = q * (8 / R) + kq;
      jm[q] = widen(pjm[k * 64 + (8 - R + rr) * 8 + c]);
      jp[q] = widen(pjp[k * 64 + rr * 8 + c]);
      im[q] = widen(pim[k * 64 + c * 8 + (8 - R + rr)]);
      ip[q] = widen(pip[k * 64 + c * 8 + rr]);
    }
  }

  // the values of planes 4H .. 4H+3 into tile planes 0 .. 3 (half tiles)
  template <int H>
  __device__ __forceinline__ void store_half(A *tile, int lane) const {
    using TL = Tile<R, sizeof(A)>;
    const int c = lane & 7, rr = (lane >> 3) % R, kq = (lane >> 3) / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      // radius 2 and 4: a value's half is fixed by q; radius 1: by kq
      if constexpr (R > 1) {
        if ((q * (8 / R)) / 4 != H) continue;
      } else {
        if ((kq >> 2) != H) continue;
      }
      const int k = q * (8 / R) + kq - 4 * H;
      tile[TL::at(k, rr, R + c)] = jm[q];
      tile[TL::at(k, 8 + R + rr, R + c)] = jp[q];
      tile[TL::at(k, R + c, rr)] = im[q];
      tile[TL::at(k, R + c, 8 + R + rr)] = ip[q];
    }
  }

  // into planes k0 .. k0+7 of the tile
  __device__ __forceinline__ void store(A *tile, int k0, int lane) const {
    using TL = Tile<R, sizeof(A)>;
    const int c = lane & 7, rr = (lane >> 3) % R, kq = (lane >> 3) / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = k0 + q * (8 / R) + kq;
      tile[TL::at(k, rr, R + c)] = jm[q];
      tile[TL::at(k, 8 + R + rr, R + c)] = jp[q];
      tile[TL::at(k, R + c, rr)] = im[q];
      tile[TL::at(k, R + c, 8 + R + rr)] = ip[q];
    }
  }
};

// One brick of the per-brick schedule, in two phases around a workgroup
// barrier: load() reads the brick's planes into registers and stages the
// brick plus its i/j faces in this wave's LDS tile; compute() evaluates and
// stores the 512 points.  Shared by the single-grid and the batched kernels.
template <class T, int R, bool COND>
struct StarBrick {
  using A = typename acc_of<T>::type;
  static constexpr int P = 8 + 2 * R;
  A v[8], vkm[R], vkp[R];

  // Lane L owns the column (j, i) = (L >> 3, L & 7), element e = L of every
  // plane.  (The transposed assignment i = L >> 3, j = L & 7 removes the
  // two-way LDS read conflicts of radius 1 in FP64 but ran the 7-point
  // sweep 5% slower: lanes then address 64-byte-apart cells.)
  // The two phases of load(): the brick's planes, k-faces and i/j faces
  // into registers, then the planes and faces into this wave's LDS tile.
  __device__ __forceinline__ void load_regs(const uint32_t *__restrict__ adj,
                                            const T *__restrict__ in, size_t st, uint32_t b,
                                            Faces<T, R> &f, int lane) {
    const int e = lane;
    const uint32_t *ab = adj + (size_t)b * 27;
    const T *pc = in + (size_t)b * st;
    const T *pkm = in + (size_t)uniform(ab[4]) * st;
    const T *pkp = in + (size_t)uniform(ab[22]) * st;
    const T *pjm = in + (size_t)uniform(ab[10]) * st;
    const T *pjp = in + (size_t)uniform(ab[16]) * st;
    const T *pim = in + (size_t)uniform(ab[12]) * st;
    const T *pip = in + (size_t)uniform(ab[14]) * st;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = widen(pc[k * 64 + e]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      vkm[r] = widen(pkm[(8 - R + r) * 64 + e]);
      vkp[r] = widen(pkp[r * 64 + e]);
    }
    f.load(pjm, pjp, pim, pip, lane);
  }
  __device__ __forceinline__ void stage(A *tile, const Faces<T, R> &f, int lane) const {
    using TL = Tile<R, sizeof(A)>;
    const int j = lane >> 3, i = lane & 7;
#pragma unroll
    for (int k = 0; k < 8; ++k) tile[TL::at(k, R + j, R + i)] = v[k];
    f.store(tile, 0, lane);
  }
  __device__ __forceinline__ void load(const uint32_t *__restrict__ adj, const T *__restrict__ in,
                                       size_t st, uint32_t b, A *tile, int lane) {
    Faces<T, R> f;
    load_regs(adj, in, st, b, f, lane);
    stage(tile, f, lane);
  }
  // planes 4H .. 4H+3 and their faces into a 4-plane tile
  template <int H>
  __device__ __forceinline__ void stage_half(A *tile, const Faces<T, R> &f, int lane) const {
    using TL = Tile<R, sizeof(A)>;
    const int j = lane >> 3, i = lane & 7;
#pragma unroll
    for (int k = 0; k < 4; ++k) tile[TL::at(k, R + j, R + i)] = v[4 * H + k];
    f.template store_half<H>(tile, lane);
  }

  // radius 1 and 2 keep the brick's 8 output planes in registers for the
  // image stores after the plane loop (3 more VGPRs for the 7-point FP64
  // kernel); radius 4 stores its images plane by plane (keeping them would
  // take it from 79 to 115 VGPRs)
  static constexpr bool kImagesAfter = R <= 2;
  A o[kImagesAfter ? 8 : 1];

  // out: the output field, ost its step; planes K0 .. K0+NK-1 of b's points
  // (their i/j taps at tile plane k - TK0) go to brick b and, for radius 4,
  // to its nimg ghost images imgs[0..nimg) (the fused refresh; nimg = 0
  // without it)
  template <int K0, int NK, int TK0>
  __device__ __forceinline__ void compute_planes(const A (&c)[6 * R + 1], A theta, int cond_mode,
                                                 T *__restrict__ out, size_t ost, uint32_t b,
                                                 const A *tile, int lane,
                                                 const uint32_t *__restrict__ imgs,
                                                 uint32_t nimg) {
    T *__restrict__ po = out + (size_t)b * ost;
    using TL = Tile<R, sizeof(A)>;
    const int j = lane >> 3, i = lane & 7, e = lane;
#pragma unroll
    for (int k = K0; k < K0 + NK; ++k) {
      auto X = [&](int kk) -> A {
        return kk < 0 ? vkm[kk + R] : (kk >= 8 ? vkp[kk - 8] : v[kk]);
      };
      const int t = k - TK0;
      A acc = c[0] * tapv<A, COND>(v[k]);
#pragma unroll
      for (int r = 1; r <= R; ++r) {
        acc = fmadd(c[6 * r - 5], tapv<A, COND>(tile[TL::at(t, R + j, R + i + r)]), acc);
        acc = fmadd(c[6 * r - 4], tapv<A, COND>(tile[TL::at(t, R + j, R + i - r)]), acc);
        acc = fmadd(c[6 * r - 3], tapv<A, COND>(tile[TL::at(t, R + j + r, R + i)]), acc);
        acc = fmadd(c[6 * r - 2], tapv<A, COND>(tile[TL::at(t, R + j - r, R + i)]), acc);
        acc = fmadd(c[6 * r - 1], tapv<A, COND>(X(k + r)), acc);
        acc = fmadd(c[6 * r], tapv<A, COND>(X(k - r)), acc);
      }
      if constexpr (COND) {
        const bool sel = cond_mode == 0 ? (acc > A(0)) : (v[k] > theta);
        acc = sel ? acc : -acc;
      }
      store_nt(&po[k * 64 + e], acc);
      if constexpr (kImagesAfter) {
        o[k] = acc;
      } else {
        for (uint32_t q = 0; q < nimg; ++q)
          store_nt(&out[(size_t)uniform(imgs[q]) * ost + k * 64 + e], acc);
      }
    }
  }

  // the ghost images of radius 1 and 2: each image's id read once, its 8
  // planes stored in a row (7-point FP64 fused sweep 475.7 -> 467.0 us per
  // step against storing them plane by plane)
  __device__ __forceinline__ void store_images(T *__restrict__ out, size_t ost, int lane,
                                               const uint32_t *__restrict__ imgs,
                                               uint32_t nimg) const {
    if constexpr (kImagesAfter) {
      for (uint32_t q = 0; q < nimg; ++q) {
        T *__restrict__ pq = out + (size_t)uniform(imgs[q]) * ost + lane;
#pragma unroll
        for (int k = 0; k < 8; ++k) store_nt(&pq[k * 64], o[k]);
      }
    }
  }

  // all 8 planes from a full tile
  __device__ __forceinline__ void compute(const A (&c)[6 * R + 1], A theta, int cond_mode,
                                          T *__restrict__ out, size_t ost, uint32_t b,
                                          const A *tile, int lane,
                                          const uint32_t *__restrict__ imgs = nullptr,
                                          uint32_t nimg = 0) {
    compute_planes<0, 8, 0>(c, theta, cond_mode, out, ost, b, tile, lane, imgs, nimg);
    store_images(out, ost, lane, imgs, nimg);
  }
};

// FP64 stages the per-brick LDS tile in two halves of four planes (all
// eight planes are in registers from the start; only the staging is split):
// a full FP64 tile held the 7-point star at 6 waves per SIMD by LDS (25.6 KB
// per workgroup), the 13-point at 4 (36.9 KB) and the 25-point at 2 (64 KB).
template <class T, int R>
__host__ __device__ constexpr bool lds_half_tiles() {
  return sizeof(typename acc_of<T>::type) == 8;
}

// WPB waves (bricks) per workgroup: 4 by default; the tuner also times 1,
// 2 and 8 (the workgroup count per CU and the barrier group change, the
// per-wave work does not).
template <class T, int R, bool COND, int WPB = 4>
__global__ __launch_bounds__(64 * WPB) void star_sweep_lds(const StarArgs<T, R, COND> a) {
  using A = typename acc_of<T>::type;
  constexpr bool kHalf = lds_half_tiles<T, R>();
  __shared__ A tile[WPB][(kHalf ? 4 : 8) * Tile<R, sizeof(A)>::PL];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * (uint32_t)WPB + (uint32_t)w;
  const bool active = slot < a.nlist;

  StarBrick<T, R, COND> sb;
  uint32_t b = 0;
  A c[6 * R + 1];
  const uint32_t *imgs = nullptr;
  uint32_t nimg = 0;
  auto prepare = [&]() {
#pragma unroll
    for (int q = 0; q < 6 * R + 1; ++q) c[q] = a.c[q];
    if (a.img_off) {
      const uint32_t i0 = uniform(a.img_off[b]);
      nimg = uniform(a.img_off[b + 1]) - i0;
      imgs = a.img + i0;
    }
  };
  if constexpr (!kHalf) {
    if (active) {
      b = uniform(a.list[slot]);
      sb.load(a.adj, a.in, a.in_step, b, tile[w], lane);
    }
    __syncthreads();
    if (!active) return;
    prepare();
    sb.compute(c, a.theta, a.cond_mode, a.out, a.out_step, b, tile[w], lane, imgs, nimg);
  } else {
    // every wave passes all three barriers, active or not
    Faces<T, R> f;
    if (active) {
      b = uniform(a.list[slot]);
      sb.load_regs(a.adj, a.in, a.in_step, b, f, lane);
      sb.template stage_half<0>(tile[w], f, lane);
    }
    __syncthreads();
    if (active) {
      prepare();
      sb.template compute_planes<0, 4, 0>(c, a.theta, a.cond_mode, a.out, a.out_step, b, tile[w],
                                          lane, imgs, nimg);
    }
    __syncthreads();  // this wave's reads of the first half are done
    if (active) sb.template stage_half<1>(tile[w], f, lane);
    __syncthreads();
    if (active) {
      sb.template compute_planes<4, 4, 4>(c, a.theta, a.cond_mode, a.out, a.out_step, b,
                                          tile[w], lane, imgs, nimg);
      sb.store_images(a.out, a.out_step, lane, imgs, nimg);
    }
  }
}


// BF16 storage in the row form: one wave per brick, lane L = k*8 + j holds
// the brick's i-row (k, j) -- eight 2-byte cells -- as one 16-byte load and
// writes it back as one 16-byte store.  (The column form above moves 2 bytes
// per lane per plane for BF16: eight times the memory instructions per byte
// of this form.)  The i taps are the lane's own cells plus R cells of the
// same row of the i-1 / i+1 neighbours (one 4- or 8-byte load each); the j
// and k taps are whole rows, of other lanes or of the j/k neighbours' R-deep
// faces, exchanged through a per-wave LDS tile of rows at (kk + R, jj + R),
// kk, jj = -R .. 7 + R (corners unused).  Each point's taps are applied in
// the column form's order (c0 x, then i+r, i-r, j+r, j-r, k+r, k-r for
// r = 1..R, FP32 FMAs, round to nearest even on store), so the bits are the
// same.  Needs 16-byte aligned fields and brick steps (the host checks).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
// cell i of a row of eight BF16 cells
__device__ __forceinline__ float bf16_cell(const u32x4 &row, int i) {
  return (i & 1) ? bf16_hi(row[i >> 1]) : bf16_lo(row[i >> 1]);
}
// p[i] of a table the kernel never writes (adjacency, plans, image lists),
// as a scalar load: in a loop that also stores, a plain load of it would be
// a vector load waited for on its own.
__device__ __forceinline__ uint32_t ld_table(const uint32_t *p, size_t i) {
  return ((kconst_ptr<uint32_t>)p)[i];
}

__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  bf16 a, b;
  narrow(a, lo);
  narrow(b, hi);
  return (uint32_t)a.u | ((uint32_t)b.u << 16);
}

template <int R>
struct RowTile {
  static constexpr int P = 8 + 2 * R;
  __device__ static constexpr int at(int kk, int jj) { return (kk + R) * P + jj + R; }
};

// Bricks per wave of the row form: BPW consecutive list entries, all their
// loads issued before the first is computed (more bytes in flight per wave).
// One is the default: at 512^3 two ran the 7-point in 129-132 us against
// 131 (the 25-point 203-217 against 194) and four in 136-141 (25-point 290).
#ifndef BK_ROWS_BPW
#define BK_ROWS_BPW 1
#endif

// A tile row of eight BF16 cells.
struct BRow {
  u32x4 v;
  __device__ __forceinline__ float cell(int i) const { return bf16_cell(v, i); }
};

// The row form's arithmetic for lane (k, j): its own row of eight cells,
// the i edges past the brick (as loaded: west cells 6, 7 / east cells 0, 1
// for R <= 2, west 4 .. 7 / east 0 .. 3 for R = 4) and rows(dk, dj), the
// row at (k + dk, j + dj) (a BRow); the taps in the column form's order.
// Returns the eight results rounded to BF16.
template <int R, bool COND, class Edge, class Rows>
__device__ __forceinline__ u32x4 row_star(const StarArgs<bf16, R, COND> &a, const u32x4 &own,
                                          const Edge &we, const Edge &ee, Rows rows) {
  float c[6 * R + 1];
#pragma unroll
  for (int q = 0; q < 6 * R + 1; ++q) c[q] = a.c[q];
  // west cells 8-R .. 7 and east cells 0 .. R-1
  float xw[R], xe[R];
  if constexpr (R <= 2) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      xw[q] = (q + 2 - R) ? bf16_hi(we) : bf16_lo(we);
      xe[q] = q ? bf16_hi(ee) : bf16_lo(ee);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      xw[q] = (q & 1) ? bf16_hi(we[q >> 1]) : bf16_lo(we[q >> 1]);
      xe[q] = (q & 1) ? bf16_hi(ee[q >> 1]) : bf16_lo(ee[q >> 1]);
    }
  }
  float x[8], acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    x[i] = bf16_cell(own, i);
    acc[i] = c[0] * tapv<float, COND>(x[i]);
  }
  auto X = [&](int i) -> float { return i < 0 ? xw[i + R] : (i >= 8 ? xe[i - 8] : x[i]); };
#pragma unroll
  for (int r = 1; r <= R; ++r) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      acc[i] = fmadd(c[6 * r - 5], tapv<float, COND>(X(i + r)), acc[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      acc[i] = fmadd(c[6 * r - 4], tapv<float, COND>(X(i - r)), acc[i]);
    const auto yp = rows(0, r), ym = rows(0, -r), zp = rows(r, 0), zm = rows(-r, 0);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      acc[i] = fmadd(c[6 * r - 3], tapv<float, COND>(yp.cell(i)), acc[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      acc[i] = fmadd(c[6 * r - 2], tapv<float, COND>(ym.cell(i)), acc[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      acc[i] = fmadd(c[6 * r - 1], tapv<float, COND>(zp.cell(i)), acc[i]);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      acc[i] = fmadd(c[6 * r], tapv<float, COND>(zm.cell(i)), acc[i]);
    // radius 4 one radius at a time: left to itself the compiler reads
    // every radius's rows first and then runs each cell's whole chain (the
    // 2x2-block kernel in 111 VGPRs, 4 waves per SIMD: 25-point 211-220 us
    // against 177-182 in 72); an empty asm that takes the accumulators and
    // clobbers memory ends this radius's taps before the next rows are read
    if constexpr (R == 4)
      __asm__ volatile("" : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]),
                       "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]), "+v"(acc[7])::"memory");
  }
  if constexpr (COND) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bool sel = a.cond_mode == 0 ? (acc[i] > 0.0f) : (x[i] > a.theta);
      acc[i] = sel ? acc[i] : -acc[i];
    }
  }
  u32x4 o;
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = pack_bf16(acc[2 * q], acc[2 * q + 1]);
  return o;
}

// Lane's row o of brick b, and of its nimg ghost images (the fused refresh):
// ids img[0 .. 7) read ahead, any further ones from a.img[i0 + q].
template <int R, bool COND>
__device__ __forceinline__ void row_store(const StarArgs<bf16, R, COND> &a, const u32x4 &o,
                                          uint32_t b, int lane, uint32_t nimg, uint32_t i0,
                                          const uint32_t (&img)[7]) {
  uint16_t *out = reinterpret_cast<uint16_t *>(a.out);
  const size_t ost = a.out_step;
  __builtin_nontemporal_store(o, reinterpret_cast<u32x4 *>(out + (size_t)b * ost + lane * 8));
#pragma unroll
  for (int q = 0; q < 7; ++q)
    if ((uint32_t)q < nimg)
      __builtin_nontemporal_store(o,
                                  reinterpret_cast<u32x4 *>(out + (size_t)img[q] * ost + lane * 8));
  for (uint32_t q = 7; q < nimg; ++q)
    __builtin_nontemporal_store(
        o, reinterpret_cast<u32x4 *>(out + (size_t)uniform(a.img[i0 + q]) * ost + lane * 8));
}

// the fused refresh's image ids of brick b: count, first index, the first
// seven ids (scalar loads)
template <int R, bool COND>
__device__ __forceinline__ void row_images(const StarArgs<bf16, R, COND> &a, uint32_t b,
                                           uint32_t &nimg, uint32_t &i0, uint32_t (&img)[7]) {
  nimg = 0;
  i0 = 0;
  if (a.img_off) {
    i0 = ld_table(a.img_off, b);
    nimg = ld_table(a.img_off, (size_t)b + 1) - i0;
#pragma unroll
    for (int q = 0; q < 7; ++q)
      if ((uint32_t)q < nimg) img[q] = ld_table(a.img, (size_t)i0 + q);
  }
}

template <int R, bool COND>
struct RowBrick {
  static constexpr int NF = (32 * R + 63) / 64;  // face rows per lane
  using TL = RowTile<R>;
  uint32_t b = 0, nimg = 0, i0 = 0;
  uint32_t img[7];  // image ids read ahead (a periodic grid's corner brick has 7)
  u32x4 own;
  u32x4 fr[NF];
  int fdst[NF];
  // the i taps past the brick's edges, as loaded: west cells 6, 7 and east
  // cells 0, 1 (R <= 2) or west 4 .. 7 and east 0 .. 3 (R = 4)
  using Edge = typename std::conditional<R <= 2, uint32_t, u32x2>::type;
  Edge we, ee;

  // every global load of brick b (list entry): its rows, the 32R face rows
  // (face f / 8R: k-1, k+1, j-1, j+1; row q = f % 8R of it), the i edges and
  // the fused refresh's image ids
  __device__ __forceinline__ void load(const StarArgs<bf16, R, COND> &a, uint32_t bb, int lane) {
    b = bb;
    const uint32_t *ab = a.adj + (size_t)b * 27;
    const uint16_t *in = reinterpret_cast<const uint16_t *>(a.in);
    const size_t st = a.in_step;
    own = *reinterpret_cast<const u32x4 *>(in + (size_t)b * st + lane * 8);
    const uint16_t *pim = in + (size_t)uniform(ab[12]) * st + lane * 8;
    const uint16_t *pip = in + (size_t)uniform(ab[14]) * st + lane * 8;
    const uint16_t *pkm = in + (size_t)uniform(ab[4]) * st;
    const uint16_t *pkp = in + (size_t)uniform(ab[22]) * st;
    const uint16_t *pjm = in + (size_t)uniform(ab[10]) * st;
    const uint16_t *pjp = in + (size_t)uniform(ab[16]) * st;
#pragma unroll
    for (int n = 0; n < NF; ++n) {
      const int f = n * 64 + lane;
      fdst[n] = -1;
      if (f < 32 * R) {
        const int face = f / (8 * R), q = f % (8 * R);
        int sk, sj, kk, jj;
        const uint16_t *src;
        if (face < 2) {  // k faces: rows (kk, jj), kk = -R .. -1 or 8 .. 7 + R
          sj = jj = q & 7;
          sk = face == 0 ? 8 - R + (q >> 3) : q >> 3;
          kk = face == 0 ? -R + (q >> 3) : 8 + (q >> 3);
          src = face == 0 ? pkm : pkp;
        } else {  // j faces: rows (kk, jj), jj = -R .. -1 or 8 .. 7 + R
          sk = kk = q / R;
          sj = face == 2 ? 8 - R + q % R : q % R;
          jj = face == 2 ? -R + q % R : 8 + q % R;
          src = face == 2 ? pjm : pjp;
        }
        fr[n] = *reinterpret_cast<const u32x4 *>(src + (sk * 8 + sj) * 8);
        fdst[n] = TL::at(kk, jj);
      }
    }
    we = *reinterpret_cast<const Edge *>(pim + (R <= 2 ? 6 : 4));
    ee = *reinterpret_cast<const Edge *>(pip);
    row_images(a, b, nimg, i0, img);
  }

  __device__ __forceinline__ void stage(u32x4 *t, int lane) const {
    t[TL::at(lane >> 3, lane & 7)] = own;
#pragma unroll
    for (int n = 0; n < NF; ++n)
      if (fdst[n] >= 0) t[fdst[n]] = fr[n];
  }

  __device__ __forceinline__ void compute(const StarArgs<bf16, R, COND> &a, const u32x4 *t,
                                          int lane) const {
    const int k = lane >> 3, j = lane & 7;
    const u32x4 o = row_star<R, COND>(
        a, own, we, ee, [&](int dk, int dj) { return BRow{t[TL::at(k + dk, j + dj)]}; });
    row_store<R, COND>(a, o, b, lane, nimg, i0, img);
  }
};

// WPB waves per workgroup, BPW bricks per wave: wave w of workgroup g takes
// list entries (g' * WPB + w) * BPW + u, u < BPW (g' the XCD-remapped g).
template <int R, bool COND, int WPB = 4, int BPW = BK_ROWS_BPW>
__global__ __launch_bounds__(64 * WPB) void star_sweep_rows(const StarArgs<bf16, R, COND> a) {
  using TL = RowTile<R>;
  __shared__ u32x4 tile[WPB][BPW][TL::P * TL::P];

  const int w = (int)uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const uint32_t slot0 =
      (xcd_remap(blockIdx.x, gridDim.x) * (uint32_t)WPB + (uint32_t)w) * (uint32_t)BPW;

  RowBrick<R, COND> rb[BPW];
#pragma unroll
  for (int u = 0; u < BPW; ++u)
    if (slot0 + u < a.nlist) rb[u].load(a, uniform(a.list[slot0 + u]), lane);
#pragma unroll
  for (int u = 0; u < BPW; ++u)
    if (slot0 + u < a.nlist) rb[u].stage(tile[w][u], lane);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPW; ++u)
    if (slot0 + u < a.nlist) rb[u].compute(a, tile[w][u], lane);
}

// BF16 quad plans in the row form: the four waves of a workgroup take the
// 2x2 block of bricks of a quad unit (member mm = wave: i-offset mm & 1,
// j-offset mm >> 1), lane (k, j) of each member its brick's i-row (k, j).
// The block's rows share one LDS tile of rows (kk, jj), kk = -R .. 7+R,
// jj = -R .. 15+R, in two halves (the left and the right member's 16-byte
// row parts, each half's rows contiguous so that a wave's row accesses are
// consecutive 16-byte slots), so the j taps across the block's middle and
// the i edges between the two members of a row come from the tile: each
// member loads its own 64 rows, its two k faces and its outer j face (24R
// rows) and its outer i edge.  A unit's levels are swept one after the
// other, one barrier pair each.
template <int R>
struct QuadRowTile {
  static constexpr int NK = 8 + 2 * R, NJ = 16 + 2 * R;
  __device__ static constexpr int at(int kk, int jj, int half) {
    return half * NK * NJ + (kk + R) * NJ + jj + R;
  }
};

template <int R, bool COND>
__global__ __launch_bounds__(256) void star_sweep_quads_rows(const StarArgs<bf16, R, COND> a) {
  using QT = QuadRowTile<R>;
  using Edge = typename std::conditional<R <= 2, uint32_t, u32x2>::type;
  constexpr int NF = (24 * R + 63) / 64;  // face rows per lane
  __shared__ u32x4 tile[QT::NK * QT::NJ * 2];
  const uint32_t u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= a.nunits) return;  // uniform over the workgroup
  const int mm = (int)uniform(threadIdx.x >> 6);
  const int mi = mm & 1, mj = mm >> 1;
  const int lane = threadIdx.x & 63;
  const int k = lane >> 3, jy = 8 * mj + (lane & 7);  // this lane's tile row
  const uint16_t *in = reinterpret_cast<const uint16_t *>(a.in);
  const size_t st = a.in_step;
  const int L = (int)ld_table(a.len, u);
  // this member's level-0 brick and its outer i and j neighbours
  uint32_t b = ld_table(a.anchor, u);
  if (mi) b = ld_table(a.adj, (size_t)b * 27 + 14);
  if (mj) b = ld_table(a.adj, (size_t)b * 27 + 16);
  const int si = mi ? 14 : 12, sj = mj ? 16 : 10;
  // above the first level the k-1 face is the previous level's top R
  // planes, still in the registers of lanes k >= 8 - R
  u32x4 prev = {0u, 0u, 0u, 0u};
  for (int l = 0; l < L; ++l) {
    const size_t ab = (size_t)b * 27;
    const bool store_me = ((ld_table(a.mask, (size_t)u * a.kcol + l) >> mm) & 1u) != 0;
    // every global load before the first LDS store
    const u32x4 own = *reinterpret_cast<const u32x4 *>(in + (size_t)b * st + lane * 8);
    // face f / 8R: k-1, k+1, the outer j neighbour; row q = f % 8R of it
    const uint16_t *pkm = in + (size_t)ld_table(a.adj, ab + 4) * st;
    const uint16_t *pkp = in + (size_t)ld_table(a.adj, ab + 22) * st;
    const uint16_t *pjo = in + (size_t)ld_table(a.adj, ab + sj) * st;
    u32x4 fr[NF];
    int fdst[NF];
#pragma unroll
    for (int n = 0; n < NF; ++n) {
      const int f = n * 64 + lane;
      fdst[n] = -1;
      if (f < 24 * R && (l == 0 || f >= 8 * R)) {
        const int face = f / (8 * R), q = f % (8 * R);
        int sk, sjj, kk, jj;
        const uint16_t *src;
        if (face < 2) {  // k faces: rows (kk, 8 mj + jj)
          sjj = q & 7;
          jj = 8 * mj + sjj;
          sk = face == 0 ? 8 - R + (q >> 3) : q >> 3;
          kk = face == 0 ? -R + (q >> 3) : 8 + (q >> 3);
          src = face == 0 ? pkm : pkp;
        } else {  // the outer j face: rows jj = -R .. -1 (mj = 0) or 16 .. 15 + R
          sk = kk = q / R;
          sjj = mj ? q % R : 8 - R + q % R;
          jj = mj ? 16 + q % R : -R + q % R;
          src = pjo;
        }
        fr[n] = *reinterpret_cast<const u32x4 *>(src + (sk * 8 + sjj) * 8);
        fdst[n] = QT::at(kk, jj, mi);
      }
    }
    // the outer i edge: the west neighbour's cells 6, 7 (4 .. 7) or the east
    // neighbour's cells 0, 1 (0 .. 3) of the same row
    const uint16_t *pio = in + (size_t)ld_table(a.adj, ab + si) * st + lane * 8;
    const Edge eo = *reinterpret_cast<const Edge *>(pio + (mi ? 0 : (R <= 2 ? 6 : 4)));
    uint32_t nimg = 0, i0 = 0, img[7];
    if (store_me) row_images(a, b, nimg, i0, img);
    tile[QT::at(k, jy, mi)] = own;
    if (l > 0 && k >= 8 - R) tile[QT::at(k - 8, jy, mi)] = prev;
#pragma unroll
    for (int n = 0; n < NF; ++n)
      if (fdst[n] >= 0) tile[fdst[n]] = fr[n];
    __syncthreads();
    prev = own;
    if (store_me) {
      // the inner i edge from the other member's half of this row
      const u32x4 nb = tile[QT::at(k, jy, 1 - mi)];
      Edge ei;
      if constexpr (R <= 2) {
        ei = mi ? nb[3] : nb[0];
      } else {
        ei = mi ? u32x2{nb[2], nb[3]} : u32x2{nb[0], nb[1]};
      }
      const u32x4 o = row_star<R, COND>(a, own, mi ? ei : eo, mi ? eo : ei, [&](int dk, int dj) {
        return BRow{tile[QT::at(k + dk, jy + dj, mi)]};
      });
      row_store<R, COND>(a, o, b, lane, nimg, i0, img);
    }
    if (l + 1 < L) {
      __syncthreads();  // every wave's reads of this level's tile are done
      b = ld_table(a.adj, ab + 22);
    }
  }
}

// The FP64 radius-4 column kernel stages its LDS tile in two halves of four
// planes: a full 16 KB tile per one-wave workgroup held it at 2.5 waves per
// SIMD (10 workgroups per CU), below the 3 its registers allow.
template <class T, int R>
__host__ __device__ constexpr bool lds_half_columns() {
  return sizeof(typename acc_of<T>::type) == 8 && R == 4;
}

// Column-marching schedule: one wave per plan unit, a run of up to kcol
// bricks b, adj[b][22], ... along k.  While brick l is computed, brick l+1
// is already being loaded into registers; its plane 0 is the k+1 face of
// brick l, and the top R planes of brick l become the k-1 face of brick
// l+1, so no k-face is read twice.  i/j faces come through the LDS tile as
// in the per-brick kernel.  Same per-point arithmetic.  One unit: L bricks
// from b up the column, store masks mk[0..L); tile is this wave's LDS tile.
// Shared by the single-grid and the batched kernels.
template <class T, int R, bool COND>
__device__ __forceinline__ void column_unit(const uint32_t *__restrict__ adj,
                                            const T *__restrict__ in, T *__restrict__ out,
                                            size_t st, size_t ost, uint32_t b, uint32_t L,
                                            const uint32_t *__restrict__ mk,
                                            const typename acc_of<T>::type (&c)[6 * R + 1],
                                            typename acc_of<T>::type theta, int cond_mode,
                                            typename acc_of<T>::type *tile, int lane) {
  using A = typename acc_of<T>::type;
  using TL = Tile<R, sizeof(A)>;
  const int j = lane >> 3, i = lane & 7;
  A cur[8], nxt[8], km[R];
  {
    const T *p = in + (size_t)b * st;
#pragma unroll
    for (int k = 0; k < 8; ++k) cur[k] = widen(p[k * 64 + lane]);
    const T *q = in + (size_t)adj[(size_t)b * 27 + 4] * st;
#pragma unroll
    for (int r = 0; r < R; ++r) km[r] = widen(q[(8 - R + r) * 64 + lane]);
  }
  for (uint32_t l = 0; l < L; ++l) {
    const uint32_t *ab = adj + (size_t)b * 27;
    const uint32_t up = ab[22];
    const bool more = l + 1 < L;
    // Face loads first: loads retire in order, so the faces can be waited
    // for while the prefetch of the next brick (issued after) stays in
    // flight through the compute below.
    Faces<T, R> f;
    f.load(in + (size_t)ab[10] * st, in + (size_t)ab[16] * st, in + (size_t)ab[12] * st,
           in + (size_t)ab[14] * st, lane);
    const T *pu = in + (size_t)up * st;
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) nxt[k] = widen(pu[k * 64 + lane]);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) nxt[r] = widen(pu[r * 64 + lane]);
    }
    const bool store_out = (mk[l] & 1u) != 0;
    T *po = out + (size_t)b * ost;
    // output plane k of this brick, its i/j taps at tile plane t
    auto plane = [&](int k, int t) {
      auto X = [&](int kk) -> A {
        return kk < 0 ? km[kk + R] : (kk >= 8 ? nxt[kk - 8] : cur[kk]);
      };
      A acc = c[0] * tapv<A, COND>(cur[k]);
#pragma unroll
      for (int r = 1; r <= R; ++r) {
        acc = fmadd(c[6 * r - 5], tapv<A, COND>(tile[TL::at(t, R + j, R + i + r)]), acc);
        acc = fmadd(c[6 * r - 4], tapv<A, COND>(tile[TL::at(t, R + j, R + i - r)]), acc);
        acc = fmadd(c[6 * r - 3], tapv<A, COND>(tile[TL::at(t, R + j + r, R + i)]), acc);
        acc = fmadd(c[6 * r - 2], tapv<A, COND>(tile[TL::at(t, R + j - r, R + i)]), acc);
        acc = fmadd(c[6 * r - 1], tapv<A, COND>(X(k + r)), acc);
        acc = fmadd(c[6 * r], tapv<A, COND>(X(k - r)), acc);
      }
      if constexpr (COND) {
        const bool sel = cond_mode == 0 ? (acc > A(0)) : (cur[k] > theta);
        acc = sel ? acc : -acc;
      }
      store_nt(&po[k * 64 + lane], acc);
    };
    if constexpr (!lds_half_columns<T, R>()) {
#pragma unroll
      for (int k = 0; k < 8; ++k) tile[TL::at(k, R + j, R + i)] = cur[k];
      f.store(tile, 0, lane);
      __syncthreads();
      if (store_out) {
#pragma unroll
        for (int k = 0; k < 8; ++k) plane(k, k);
      }
    } else {
      // the tile in two halves of four planes
#pragma unroll
      for (int k = 0; k < 4; ++k) tile[TL::at(k, R + j, R + i)] = cur[k];
      f.template store_half<0>(tile, lane);
      __syncthreads();
      if (store_out) {
#pragma unroll
        for (int k = 0; k < 4; ++k) plane(k, k);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k) tile[TL::at(k, R + j, R + i)] = cur[4 + k];
      f.template store_half<1>(tile, lane);
      __syncthreads();
      if (store_out) {
#pragma unroll
        for (int k = 4; k < 8; ++k) plane(k, k - 4);
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) km[r] = cur[8 - R + r];
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
    }
    b = up;
    __syncthreads();
  }
}

template <class T, int R, bool COND>
__global__ __launch_bounds__(64) void star_sweep_columns(const StarArgs<T, R, COND> a) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[(lds_half_columns<T, R>() ? 4 : 8) * Tile<R, sizeof(A)>::PL];
  const uint32_t u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= a.nunits) return;
  A c[6 * R + 1];
#pragma unroll
  for (int q = 0; q < 6 * R + 1; ++q) c[q] = a.c[q];
  column_unit<T, R, COND>(a.adj, a.in, a.out, a.in_step, a.out_step, a.anchor[u], a.len[u],
                          a.mask + (size_t)u * a.kcol, c, a.theta, a.cond_mode, tile,
                          threadIdx.x);
}

// Batched column schedule: one persistent launch serves every request of a
// group table whose bricks come as a column plan; work item = one plan unit
// of one request.
template <class T, int R, bool COND>
__global__ __launch_bounds__(64) void star_sweep_columns_batched(const BatchTable *tab_,
                                                                 uint32_t tick) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[(lds_half_columns<T, R>() ? 4 : 8) * Tile<R, sizeof(A)>::PL];
  const kconst_ptr<BatchTable> tab = (kconst_ptr<BatchTable>)tab_;
  const uint32_t nsegs = tab->nsegs;
  const uint32_t total = tab->prefix[nsegs];
  SegCursor cur;
  for (uint32_t u = xcd_remap(blockIdx.x, gridDim.x); u < total; u += gridDim.x) {
    cur.seek(tab->prefix, nsegs, u);
    const auto &sg = tab->seg[cur.s];
    if (tick >= sg.ticks) continue;
    const uint32_t ul = u - cur.lo;
    const bool odd = ((sg.parity + tick) & 1u) != 0;
    const T *in = static_cast<const T *>(odd ? sg.f1 : sg.f0);
    T *out = static_cast<T *>(odd ? sg.f0 : sg.f1);
    const kconst_ptr<A> cp = (kconst_ptr<A>)sg.coeff;
    A c[6 * R + 1];
#pragma unroll
    for (int q = 0; q < 6 * R + 1; ++q) c[q] = cp[q];
    column_unit<T, R, COND>(sg.adj, in, out, sg.step, sg.step, sg.items[ul], sg.len[ul],
                            sg.mask + (size_t)ul * sg.kcol, c, (A)sg.theta, sg.cond_mode, tile,
                            threadIdx.x);
  }
}


// Quad-column schedule: one workgroup of four waves per quad-plan unit, a
// 2x2 block of k-columns (members m0 = anchor, m1 = i+1, m2 = j+1,
// m3 = i+1,j+1) marched level by level.  Each wave owns one member column
// and keeps its planes in registers as the k-march does (the next level's
// planes prefetched during compute, k-faces never read twice); the four
// members' planes share one LDS tile of the whole 16x16 block plus its
// R-deep outer faces, so the faces between the members come from LDS and
// each wave fetches only its two outer faces: half the in-plane face bytes
// of the per-column schedules.  Same per-point arithmetic.
// Rows of 16 + 2R cells.  (Padding the rows of radius 1 and 2 to 24 cells,
// which puts the rows one read pass serves on distinct LDS banks, removed
// most bank-conflict cycles but not time: 7-point FP64 397.4 against 394.6
// us, 13-point 416.0 against 412.9, FP32 7-point 213.0 against 213.7; the
// larger tile costs occupancy.)
template <int R, int ES>
struct QuadTile {
  static constexpr int P = 16 + 2 * R;  // rows and columns of the block tile
  static constexpr int RS = P;          // row stride
  static constexpr int PL = P * RS;     // plane stride
  __device__ static constexpr int at(int k, int y, int x) { return k * PL + y * RS + x; }
};

// Tile planes staged per pass of a quad level: 8 (one pass), or 4 (two
// passes, half the LDS per workgroup, so twice the workgroups per CU where
// LDS is what limits them).  Bit R of BK_QUAD_HALF_R64 (8-byte cells) or
// BK_QUAD_HALF_R32 (4-byte cells) selects two passes for radius R.  Measured
// at 512^3 (kernel ms, one pass / two): FP32 25-point 4 levels 0.250 / 0.239,
// 2 levels 0.264 / 0.241; FP64 25-point 0.468 / 0.482 (the extra barriers,
// no more waves: registers hold it at 4 per SIMD either way), FP64 13-point
// 0.412 / 0.433, FP32 13-point 0.210 / 0.216.
#ifndef BK_QUAD_HALF_R64
#define BK_QUAD_HALF_R64 0x0
#endif
#ifndef BK_QUAD_HALF_R32
#define BK_QUAD_HALF_R32 0x10
#endif
template <class A, int R>
constexpr int quad_tile_planes() {
  return (((sizeof(A) == 8 ? BK_QUAD_HALF_R64 : BK_QUAD_HALF_R32) >> R) & 1) ? 4 : 8;
}

// The two outer faces one level ahead: a unit's first faces are issued
// before its brick planes, and each next level's faces during this level's
// compute (2R more values per lane) instead of at the start of the next
// level, where the tile fill waits for them.  Bit R of BK_QUAD_PF_R64 /
// BK_QUAD_PF_R32.  Measured at 512^3 (kernel ms, at the start / ahead):
// FP64 7-point 1 level 0.397 / 0.392, cond 0.398 / 0.392, 25-point 2 levels
// 0.465 / 0.458 (with 4 waves per SIMD kept, below), FP32 7-point 0.206 /
// 0.202, 25-point 2 levels 0.238 / 0.227.
#ifndef BK_QUAD_PF_R64
#define BK_QUAD_PF_R64 0x16
#endif
#ifndef BK_QUAD_PF_R32
#define BK_QUAD_PF_R32 0x16
#endif
template <class A, int R>
constexpr bool quad_face_prefetch() {
  return (((sizeof(A) == 8 ? BK_QUAD_PF_R64 : BK_QUAD_PF_R32) >> R) & 1) != 0;
}

// One quad unit: the four waves of a workgroup (member mm = wave index)
// march L levels from the unit's anchor; mk[l] holds level l's store mask
// (bit mm: store member mm).  Every wave of the workgroup must call it with
// the same unit (it synchronises the workgroup twice per pass of a level
// through the tile).  Shared by
// the single-grid and the batched kernels.  SHIFT: the i-direction taps come
// from DPP lane shifts of the wave's own plane and of the two cells eight
// columns away (one tile read each per plane) instead of 2R tile reads; the
// same values, so the same bits.
// PF: the outer faces one level ahead (quad_face_prefetch; the
// single-grid kernel -- the persistent batched kernel keeps those registers
// for its occupancy).
template <class T, int R, bool COND, bool SHIFT = false, bool PF = false>
__device__ __forceinline__ void quad_unit(const uint32_t *__restrict__ adj,
                                          const T *__restrict__ in, T *__restrict__ out,
                                          size_t st, size_t ost, uint32_t anchor, int L,
                                          const uint32_t *__restrict__ mk,
                                          const typename acc_of<T>::type (&c)[6 * R + 1],
                                          typename acc_of<T>::type theta, int cond_mode,
                                          const uint32_t *__restrict__ img_off,
                                          const uint32_t *__restrict__ img,
                                          typename acc_of<T>::type *tile) {
  using A = typename acc_of<T>::type;
  using QT = QuadTile<R, sizeof(A)>;
  constexpr int KH = quad_tile_planes<A, R>(), NH = 8 / KH;
  const int mm = (int)uniform(threadIdx.x >> 6);  // member: i-offset mm & 1, j-offset mm >> 1
  const int lane = threadIdx.x & 63;
  const int j = lane >> 3, i = lane & 7;
  const int mi = mm & 1, mj = mm >> 1;

  // this wave's level-0 brick
  uint32_t b = anchor;
  if (mi) b = adj[(size_t)b * 27 + 14];
  if (mj) b = adj[(size_t)b * 27 + 16];
  b = uniform(b);
  // this member's outer neighbours: i-1 or i+1, j-1 or j+1
  const int si = mi ? 14 : 12, sj = mj ? 16 : 10;
  // face lanes: column c, depth rr, plane group kq (as Faces)
  const int fc = lane & 7, frr = (lane >> 3) % R, fkq = (lane >> 3) / R;
  // tile coordinates of this member's cells and of its outer faces
  const int ty = R + 8 * mj + j, tx = R + 8 * mi + i;
  const int fix = mi ? R + 16 + frr : frr;  // the i-face's tile column
  const int fjy = mj ? R + 16 + frr : frr;  // the j-face's tile row

  // the two outer faces of brick bb (its i- and j-neighbours' R-deep layers)
  auto load_faces = [&](uint32_t bb, A(&ffi)[R], A(&ffj)[R]) {
    const uint32_t *abb = adj + (size_t)bb * 27;
    const T *pi = in + (size_t)uniform(abb[si]) * st;
    const T *pj = in + (size_t)uniform(abb[sj]) * st;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = q * (8 / R) + fkq;
      ffi[q] = widen(pi[k * 64 + fc * 8 + (mi ? frr : 8 - R + frr)]);
      ffj[q] = widen(pj[k * 64 + (mj ? frr : 8 - R + frr) * 8 + fc]);
    }
  };

  A cur[8], nxt[8], km[R], fi[R], fj[R];
  if constexpr (PF) load_faces(b, fi, fj);
  {
    const T *p = in + (size_t)b * st;
#pragma unroll
    for (int k = 0; k < 8; ++k) cur[k] = widen(p[k * 64 + lane]);
    const T *q = in + (size_t)uniform(adj[(size_t)b * 27 + 4]) * st;
#pragma unroll
    for (int r = 0; r < R; ++r) km[r] = widen(q[(8 - R + r) * 64 + lane]);
  }
  for (int l = 0; l < L; ++l) {
    const uint32_t *ab = adj + (size_t)b * 27;
    const uint32_t up = uniform(ab[22]);
    const bool more = l + 1 < L;
    // the two outer faces first (waited for before the tile is filled;
    // with PF loaded one level ahead), then the prefetch of the next level,
    // which stays in flight through the compute below
    if constexpr (!PF) load_faces(b, fi, fj);
    const T *pu = in + (size_t)up * st;
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) nxt[k] = widen(pu[k * 64 + lane]);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) nxt[r] = widen(pu[r * 64 + lane]);
    }
    A fni[R], fnj[R];
    if constexpr (PF) {
      if (more) load_faces(up, fni, fnj);
    }
    const bool store_me = ((mk[l] >> mm) & 1u) != 0;
    T *po = out + (size_t)b * ost;
    // fused refresh: b's ghost images receive its output too
    const uint32_t *imgs = nullptr;
    uint32_t nimg = 0;
    if (img_off && store_me) {
      const uint32_t i0 = uniform(img_off[b]);
      nimg = uniform(img_off[b + 1]) - i0;
      imgs = img + i0;
    }
    // the level's planes in NH passes of KH planes through the tile
#pragma unroll
    for (int h = 0; h < NH; ++h) {
#pragma unroll
      for (int kk = 0; kk < KH; ++kk) tile[QT::at(kk, ty, tx)] = cur[h * KH + kk];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int k = q * (8 / R) + fkq;  // for R = 1 the plane, and so the pass, varies by lane
        if (k / KH != h) continue;
        tile[QT::at(k - h * KH, R + 8 * mj + fc, fix)] = fi[q];
        tile[QT::at(k - h * KH, fjy, R + 8 * mi + fc)] = fj[q];
      }
      __syncthreads();
      if (store_me) {
#pragma unroll
        for (int kk = 0; kk < KH; ++kk) {
          const int k = h * KH + kk;
          auto X = [&](int z) -> A {
            return z < 0 ? km[z + R] : (z >= 8 ? nxt[z - 8] : cur[z]);
          };
          A acc = c[0] * tapv<A, COND>(cur[k]);
          // SHIFT: the cells 8 columns right and left (the i-neighbour
          // bricks' same lane; only lanes i < R of the first and i >= 8 - R
          // of the second are used, the others clamped into the tile's row)
          A east = A(0), west = A(0);
          if constexpr (SHIFT) {
            east = tile[QT::at(kk, ty, tx + 8 < QT::P ? tx + 8 : QT::P - 1)];
            west = tile[QT::at(kk, ty, tx >= 8 ? tx - 8 : 0)];
          }
#pragma unroll
          for (int r = 1; r <= R; ++r) {
            if constexpr (SHIFT) {
              acc = fmadd(c[6 * r - 5], tapv<A, COND>(itap_shift<true>(cur[k], east, i, r)), acc);
              acc = fmadd(c[6 * r - 4], tapv<A, COND>(itap_shift<false>(cur[k], west, i, r)), acc);
            } else {
              acc = fmadd(c[6 * r - 5], tapv<A, COND>(tile[QT::at(kk, ty, tx + r)]), acc);
              acc = fmadd(c[6 * r - 4], tapv<A, COND>(tile[QT::at(kk, ty, tx - r)]), acc);
            }
            acc = fmadd(c[6 * r - 3], tapv<A, COND>(tile[QT::at(kk, ty + r, tx)]), acc);
            acc = fmadd(c[6 * r - 2], tapv<A, COND>(tile[QT::at(kk, ty - r, tx)]), acc);
            acc = fmadd(c[6 * r - 1], tapv<A, COND>(X(k + r)), acc);
            acc = fmadd(c[6 * r], tapv<A, COND>(X(k - r)), acc);
          }
          if constexpr (COND) {
            const bool sel = cond_mode == 0 ? (acc > A(0)) : (cur[k] > theta);
            acc = sel ? acc : -acc;
          }
          store_nt(&po[k * 64 + lane], acc);
          for (uint32_t q = 0; q < nimg; ++q)
            store_nt(&out[(size_t)uniform(imgs[q]) * ost + k * 64 + lane], acc);
        }
      }
      __syncthreads();  // every wave's reads of this pass's tile are done
    }
#pragma unroll
    for (int r = 0; r < R; ++r) km[r] = cur[8 - R + r];
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
      if constexpr (PF) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          fi[q] = fni[q];
          fj[q] = fnj[q];
        }
      }
    }
    b = up;
  }
}

// Waves per SIMD the compiler must keep room for (its register budget is
// 512 / N VGPRs per lane); 1 leaves the choice to it.  Radius 4 in FP64 with
// the face prefetch would otherwise take 152 VGPRs (3 waves per SIMD: 25-point
// 0.502 ms against 0.458 at 4 waves in 110 VGPRs); 5 or 6 waves spill
// (0.54-0.79 ms).
#ifndef BK_QUAD_MINW_R4_64
#define BK_QUAD_MINW_R4_64 4
#endif
#ifndef BK_QUAD_MINW_OTHER
#define BK_QUAD_MINW_OTHER 1
#endif
template <class A, int R>
constexpr int quad_min_waves() {
  return sizeof(A) == 8 && R == 4 ? BK_QUAD_MINW_R4_64 : BK_QUAD_MINW_OTHER;
}

template <class T, int R, bool COND, bool SHIFT>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(quad_min_waves<typename acc_of<T>::type, R>())))
void star_sweep_quads(const StarArgs<T, R, COND> a) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[quad_tile_planes<A, R>() * QuadTile<R, sizeof(A)>::PL];
  const uint32_t u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= a.nunits) return;  // uniform over the workgroup
  A c[6 * R + 1];
#pragma unroll
  for (int q = 0; q < 6 * R + 1; ++q) c[q] = a.c[q];
  quad_unit<T, R, COND, SHIFT, quad_face_prefetch<A, R>()>(
      a.adj, a.in, a.out, a.in_step, a.out_step, a.anchor[u], (int)a.len[u],
      a.mask + (size_t)u * a.kcol, c, a.theta, a.cond_mode, a.img_off, a.img, tile);
}

// Batched quad schedule: one persistent launch serves every request of a
// group table whose bricks come as a quad plan; work item = one quad unit of
// one request, taken by a whole workgroup.
template <class T, int R, bool COND>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(quad_min_waves<typename acc_of<T>::type, R>())))
void star_sweep_quads_batched(const BatchTable *tab_, uint32_t tick) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[quad_tile_planes<A, R>() * QuadTile<R, sizeof(A)>::PL];
  const kconst_ptr<BatchTable> tab = (kconst_ptr<BatchTable>)tab_;
  const uint32_t nsegs = tab->nsegs;
  const uint32_t total = tab->prefix[nsegs];
  SegCursor cur;
  for (uint32_t u = xcd_remap(blockIdx.x, gridDim.x); u < total; u += gridDim.x) {
    cur.seek(tab->prefix, nsegs, u);
    const auto &sg = tab->seg[cur.s];
    if (tick >= sg.ticks) continue;  // the same unit for every wave: uniform
    const uint32_t ul = u - cur.lo;
    const bool odd = ((sg.parity + tick) & 1u) != 0;
    const T *in = static_cast<const T *>(odd ? sg.f1 : sg.f0);
    T *out = static_cast<T *>(odd ? sg.f0 : sg.f1);
    const kconst_ptr<A> cp = (kconst_ptr<A>)sg.coeff;
    A c[6 * R + 1];
#pragma unroll
    for (int q = 0; q < 6 * R + 1; ++q) c[q] = cp[q];
    quad_unit<T, R, COND>(sg.adj, in, out, sg.step, sg.step, sg.items[ul], (int)sg.len[ul],
                          sg.mask + (size_t)ul * sg.kcol, c, (A)sg.theta, sg.cond_mode, nullptr,
                          nullptr, tile);
  }
}

// K-stack variant: one workgroup of four waves per column unit of up to
// four bricks stacked along k (plan kcol <= 4).  The four waves load their
// bricks at the same time (the memory parallelism of the per-brick kernel)
// into one tall LDS tile [32][8+2R][8+2R]; the k-faces between the stacked
// bricks are read from that tile instead of from memory, so only the unit's
// bottom and top faces are fetched.  Same per-point arithmetic.
template <class T, int R, bool COND>
__global__ __launch_bounds__(256) void star_sweep_kstack(const StarArgs<T, R, COND> a) {
  using A = typename acc_of<T>::type;
  using TL = Tile<R, sizeof(A)>;
  constexpr int W = 4;
  __shared__ A tile[W * 8 * TL::PL];
  __shared__ A kb[R][64], ka[R][64];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int j = lane >> 3, i = lane & 7;
  const uint32_t u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= a.nunits) return;  // uniform over the workgroup
  const int L = (int)a.len[u];
  const bool active = w < L;
  const size_t st = a.in_step;

  A v[8];
  uint32_t b = 0;
  if (active) {
    b = uniform(a.flat[a.start[u] + (uint32_t)w]);
    const uint32_t *ab = a.adj + (size_t)b * 27;
    const T *pc = a.in + (size_t)b * st;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = widen(pc[k * 64 + lane]);
    Faces<T, R> f;
    f.load(a.in + (size_t)uniform(ab[10]) * st, a.in + (size_t)uniform(ab[16]) * st,
           a.in + (size_t)uniform(ab[12]) * st, a.in + (size_t)uniform(ab[14]) * st, lane);
    if (w == 0) {
      const T *q = a.in + (size_t)uniform(ab[4]) * st;
#pragma unroll
      for (int r = 0; r < R; ++r) kb[r][lane] = widen(q[(8 - R + r) * 64 + lane]);
    }
    if (w == L - 1) {
      const T *q = a.in + (size_t)uniform(ab[22]) * st;
#pragma unroll
      for (int r = 0; r < R; ++r) ka[r][lane] = widen(q[r * 64 + lane]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) tile[TL::at(w * 8 + k, R + j, R + i)] = v[k];
    f.store(tile, w * 8, lane);
  }
  __syncthreads();
  if (!active || !(a.mask[(size_t)u * a.kcol + w] & 1u)) return;

  const int top = L * 8;
  T *po = a.out + (size_t)b * a.out_step;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    auto X = [&](int kk) -> A {
      if (kk >= 0 && kk < 8) return v[kk];
      const int gk = w * 8 + kk;  // plane in the stack
      if (gk < 0) return kb[gk + R][lane];
      if (gk >= top) return ka[gk - top][lane];
      return tile[TL::at(gk, R + j, R + i)];
    };
    A acc = a.c[0] * tapv<A, COND>(v[k]);
#pragma unroll
    for (int r = 1; r <= R; ++r) {
      acc = fmadd(a.c[6 * r - 5], tapv<A, COND>(tile[TL::at(w * 8 + k, R + j, R + i + r)]), acc);
      acc = fmadd(a.c[6 * r - 4], tapv<A, COND>(tile[TL::at(w * 8 + k, R + j, R + i - r)]), acc);
      acc = fmadd(a.c[6 * r - 3], tapv<A, COND>(tile[TL::at(w * 8 + k, R + j + r, R + i)]), acc);
      acc = fmadd(a.c[6 * r - 2], tapv<A, COND>(tile[TL::at(w * 8 + k, R + j - r, R + i)]), acc);
      acc = fmadd(a.c[6 * r - 1], tapv<A, COND>(X(k + r)), acc);
      acc = fmadd(a.c[6 * r], tapv<A, COND>(X(k - r)), acc);
    }
    if constexpr (COND) {
      const bool sel = a.cond_mode == 0 ? (acc > A(0)) : (v[k] > a.theta);
      acc = sel ? acc : -acc;
    }
    store_nt(&po[k * 64 + lane], acc);
  }
}

template <class T, int R, bool COND>
hipError_t launch_cols_t(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  using A = typename acc_of<T>::type;
  StarArgs<T, R, COND> a;
  a.adj = L.adj;
  a.list = nullptr;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = 0;
  a.cond_mode = L.cond_mode;
  a.anchor = P.anchor;
  a.len = P.len;
  a.mask = P.mask;
  a.flat = P.flat;
  a.start = P.start;
  a.nunits = P.nunits;
  a.kcol = P.kcol;
  a.img_off = a.img = nullptr;
  a.theta = (A)L.theta;
  for (int q = 0; q < 6 * R + 1; ++q) a.c[q] = (A)L.coeff[q];
  if (P.nunits == 0) return hipSuccess;
  if (P.kcol <= 4)
    hipLaunchKernelGGL((star_sweep_kstack<T, R, COND>), dim3(P.nunits), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((star_sweep_columns<T, R, COND>), dim3(P.nunits), dim3(64), 0, s, a);
  return hipGetLastError();
}

// BF16 fields whose 16-byte rows are 16-byte aligned: the row-form kernels
bool rows_aligned(const SweepLaunch &L) {
  return (reinterpret_cast<uintptr_t>(L.in) & 15u) == 0 &&
         (reinterpret_cast<uintptr_t>(L.out) & 15u) == 0 && L.in_step % 8 == 0 &&
         L.out_step % 8 == 0;
}

template <class T, int R, bool COND>
hipError_t launch_quads_t(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  using A = typename acc_of<T>::type;
  StarArgs<T, R, COND> a;
  a.adj = L.adj;
  a.list = nullptr;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = 0;
  a.cond_mode = L.cond_mode;
  a.anchor = P.anchor;
  a.len = P.len;
  a.mask = P.mask;
  a.flat = P.flat;
  a.start = P.start;
  a.nunits = P.nunits;
  a.kcol = P.kcol;
  a.img_off = L.img_off;
  a.img = L.img;
  a.theta = (A)L.theta;
  for (int q = 0; q < 6 * R + 1; ++q) a.c[q] = (A)L.coeff[q];
  if (P.nunits == 0) return hipSuccess;
  if constexpr (std::is_same<T, bf16>::value) {
    // the row form where fields and brick steps keep 16-byte rows aligned
    if (!L.lane_shift && rows_aligned(L)) {
      hipLaunchKernelGGL((star_sweep_quads_rows<R, COND>), dim3(P.nunits), dim3(256), 0, s, a);
      return hipGetLastError();
    }
  }
  if (L.lane_shift)
    hipLaunchKernelGGL((star_sweep_quads<T, R, COND, true>), dim3(P.nunits), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((star_sweep_quads<T, R, COND, false>), dim3(P.nunits), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <class T>
hipError_t launch_quads_dt(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  switch (L.stencil) {
    case BK_STAR7: return launch_quads_t<T, 1, false>(L, P, s);
    case BK_STAR13: return launch_quads_t<T, 2, false>(L, P, s);
    case BK_STAR25: return launch_quads_t<T, 4, false>(L, P, s);
    case BK_COND7: return launch_quads_t<T, 1, true>(L, P, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t launch_cols_dt(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  switch (L.stencil) {
    case BK_STAR7: return launch_cols_t<T, 1, false>(L, P, s);
    case BK_STAR13: return launch_cols_t<T, 2, false>(L, P, s);
    case BK_STAR25: return launch_cols_t<T, 4, false>(L, P, s);
    case BK_COND7: return launch_cols_t<T, 1, true>(L, P, s);
    default: return hipErrorInvalidValue;
  }
}

// Batched variant: one persistent launch serves every request of a group
// table (same stencil and dtype).  Work item = one brick of one request;
// the workgroup's four waves take four consecutive items per round, find
// their request by binary search over the table's prefix counts, and run
// the per-brick body with that request's adjacency, fields (ping-pong by
// parity and tick) and coefficients.
template <class T, int R, bool COND>
__global__ __launch_bounds__(256) void star_sweep_batched(const BatchTable *tab_, uint32_t tick) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[4][(lds_half_tiles<T, R>() ? 4 : 8) * Tile<R, sizeof(A)>::PL];
  const kconst_ptr<BatchTable> tab = (kconst_ptr<BatchTable>)tab_;
  const uint32_t nsegs = tab->nsegs;
  const uint32_t total = tab->prefix[nsegs];
  // the wave index through readfirstlane: everything derived from the item
  // (segment, brick, pointers, coefficients) is then wave-uniform and lives
  // in scalar registers, with scalar loads and uniform branches
  const uint32_t w = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const uint32_t nwg = gridDim.x;
  SegCursor cur;
  for (uint32_t base = xcd_remap(blockIdx.x, nwg) * 4u; base < total; base += nwg * 4u) {
    const uint32_t item = base + w;
    bool active = item < total;
    if (active) {
      cur.seek(tab->prefix, nsegs, item);
      active = tick < tab->seg[cur.s].ticks;
    }
    StarBrick<T, R, COND> sb;
    uint32_t b = 0;
    if constexpr (!lds_half_tiles<T, R>()) {
      if (active) {
        const auto &sg = tab->seg[cur.s];
        const bool odd = ((sg.parity + tick) & 1u) != 0;
        const T *in = static_cast<const T *>(odd ? sg.f1 : sg.f0);
        b = sg.items[item - cur.lo];
        sb.load(sg.adj, in, sg.step, b, tile[w], lane);
      }
      __syncthreads();
      if (active) {
        const auto &sg = tab->seg[cur.s];
        const bool odd = ((sg.parity + tick) & 1u) != 0;
        T *out = static_cast<T *>(odd ? sg.f0 : sg.f1);
        const kconst_ptr<A> cp = (kconst_ptr<A>)sg.coeff;
        A c[6 * R + 1];
#pragma unroll
        for (int q = 0; q < 6 * R + 1; ++q) c[q] = cp[q];
        sb.compute(c, (A)sg.theta, sg.cond_mode, out, sg.step, b, tile[w], lane);
      }
    } else {
      // the tile in two halves of four planes; every wave passes all the
      // barriers, active or not
      Faces<T, R> f;
      if (active) {
        const auto &sg = tab->seg[cur.s];
        const bool odd = ((sg.parity + tick) & 1u) != 0;
        const T *in = static_cast<const T *>(odd ? sg.f1 : sg.f0);
        b = sg.items[item - cur.lo];
        sb.load_regs(sg.adj, in, sg.step, b, f, lane);
        sb.template stage_half<0>(tile[w], f, lane);
      }
      __syncthreads();
      A c[6 * R + 1];
      T *out = nullptr;
      if (active) {
        const auto &sg = tab->seg[cur.s];
        const bool odd = ((sg.parity + tick) & 1u) != 0;
        out = static_cast<T *>(odd ? sg.f0 : sg.f1);
        const kconst_ptr<A> cp = (kconst_ptr<A>)sg.coeff;
#pragma unroll
        for (int q = 0; q < 6 * R + 1; ++q) c[q] = cp[q];
        sb.template compute_planes<0, 4, 0>(c, (A)sg.theta, sg.cond_mode, out, sg.step, b,
                                            tile[w], lane, nullptr, 0u);
      }
      __syncthreads();
      if (active) sb.template stage_half<1>(tile[w], f, lane);
      __syncthreads();
      if (active) {
        const auto &sg = tab->seg[cur.s];
        sb.template compute_planes<4, 4, 4>(c, (A)sg.theta, sg.cond_mode, out, sg.step, b,
                                            tile[w], lane, nullptr, 0u);
      }
    }
    __syncthreads();
  }
}

template <class T, int R, bool COND>
hipError_t launch_batched_t(const BatchTable *tab, uint32_t tick, hipStream_t s) {
  static int nblk = 0;
  if (nblk == 0) {
    int dev = 0, cus = 256, per = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, star_sweep_batched<T, R, COND>, 256, 0);
    nblk = cus * (per > 0 ? per : 1);
  }
  hipLaunchKernelGGL((star_sweep_batched<T, R, COND>), dim3(nblk), dim3(256), 0, s, tab, tick);
  return hipGetLastError();
}

template <class T>
hipError_t launch_batched_dt(int stencil, const BatchTable *tab, uint32_t tick, hipStream_t s) {
  switch (stencil) {
    case BK_STAR7: return launch_batched_t<T, 1, false>(tab, tick, s);
    case BK_STAR13: return launch_batched_t<T, 2, false>(tab, tick, s);
    case BK_STAR25: return launch_batched_t<T, 4, false>(tab, tick, s);
    case BK_COND7: return launch_batched_t<T, 1, true>(tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T, int R, bool COND>
hipError_t launch_cols_batched_t(const BatchTable *tab, uint32_t tick, hipStream_t s) {
  static int nblk = 0;
  if (nblk == 0) {
    int dev = 0, cus = 256, per = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per, star_sweep_columns_batched<T, R, COND>, 64, 0);
    nblk = cus * (per > 0 ? per : 1);
  }
  hipLaunchKernelGGL((star_sweep_columns_batched<T, R, COND>), dim3(nblk), dim3(64), 0, s, tab,
                     tick);
  return hipGetLastError();
}

template <class T, int R, bool COND>
int quads_batched_blocks() {
  static int nblk = 0;
  if (nblk == 0) {
    int dev = 0, cus = 256, per = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per, star_sweep_quads_batched<T, R, COND>, 256, 0);
    nblk = cus * (per > 0 ? per : 1);
  }
  return nblk;
}

template <class T, int R, bool COND>
hipError_t launch_quads_batched_t(const BatchTable *tab, uint32_t tick, hipStream_t s) {
  hipLaunchKernelGGL((star_sweep_quads_batched<T, R, COND>),
                     dim3(quads_batched_blocks<T, R, COND>()), dim3(256), 0, s, tab, tick);
  return hipGetLastError();
}

template <class T>
uint32_t quads_batched_capacity_dt(int stencil) {
  switch (stencil) {
    case BK_STAR7: return (uint32_t)quads_batched_blocks<T, 1, false>();
    case BK_STAR13: return (uint32_t)quads_batched_blocks<T, 2, false>();
    case BK_STAR25: return (uint32_t)quads_batched_blocks<T, 4, false>();
    case BK_COND7: return (uint32_t)quads_batched_blocks<T, 1, true>();
    default: return 0;
  }
}

template <class T>
hipError_t launch_quads_batched_dt(int stencil, const BatchTable *tab, uint32_t tick,
                                   hipStream_t s) {
  switch (stencil) {
    case BK_STAR7: return launch_quads_batched_t<T, 1, false>(tab, tick, s);
    case BK_STAR13: return launch_quads_batched_t<T, 2, false>(tab, tick, s);
    case BK_STAR25: return launch_quads_batched_t<T, 4, false>(tab, tick, s);
    case BK_COND7: return launch_quads_batched_t<T, 1, true>(tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t launch_cols_batched_dt(int stencil, const BatchTable *tab, uint32_t tick,
                                  hipStream_t s) {
  switch (stencil) {
    case BK_STAR7: return launch_cols_batched_t<T, 1, false>(tab, tick, s);
    case BK_STAR13: return launch_cols_batched_t<T, 2, false>(tab, tick, s);
    case BK_STAR25: return launch_cols_batched_t<T, 4, false>(tab, tick, s);
    case BK_COND7: return launch_cols_batched_t<T, 1, true>(tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

template <int R, bool COND>
hipError_t launch_rows(const StarArgs<bf16, R, COND> &a, uint32_t nlist, int wg_bricks,
                       hipStream_t s) {
  constexpr uint32_t B = BK_ROWS_BPW;
  switch (wg_bricks) {
    case 1:
      hipLaunchKernelGGL((star_sweep_rows<R, COND, 1>), dim3((nlist + B - 1) / B), dim3(64), 0,
                         s, a);
      break;
    case 2:
      hipLaunchKernelGGL((star_sweep_rows<R, COND, 2>), dim3((nlist + 2 * B - 1) / (2 * B)),
                         dim3(128), 0, s, a);
      break;
    case 8:
      hipLaunchKernelGGL((star_sweep_rows<R, COND, 8>), dim3((nlist + 8 * B - 1) / (8 * B)),
                         dim3(512), 0, s, a);
      break;
    case 0:
    case 4:
      hipLaunchKernelGGL((star_sweep_rows<R, COND>), dim3((nlist + 4 * B - 1) / (4 * B)),
                         dim3(256), 0, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class T, int R, bool COND>
hipError_t launch_t(const SweepLaunch &L, hipStream_t s) {
  using A = typename acc_of<T>::type;
  StarArgs<T, R, COND> a;
  a.adj = L.adj;
  a.list = L.list;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = L.nlist;
  a.cond_mode = L.cond_mode;
  a.anchor = a.len = a.mask = a.flat = a.start = nullptr;
  a.img_off = L.img_off;
  a.img = L.img;
  a.nunits = a.kcol = 0;
  a.theta = (A)L.theta;
  for (int q = 0; q < 6 * R + 1; ++q) a.c[q] = (A)L.coeff[q];
  if constexpr (std::is_same<T, bf16>::value) {
    // the row form where fields and brick steps keep 16-byte rows aligned
    if (rows_aligned(L)) return launch_rows<R, COND>(a, L.nlist, L.wg_bricks, s);
  }
  switch (L.wg_bricks) {
    case 1:
      hipLaunchKernelGGL((star_sweep_lds<T, R, COND, 1>), dim3(L.nlist), dim3(64), 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((star_sweep_lds<T, R, COND, 2>), dim3((L.nlist + 1u) / 2u), dim3(128),
                         0, s, a);
      break;
    case 8:
      hipLaunchKernelGGL((star_sweep_lds<T, R, COND, 8>), dim3((L.nlist + 7u) / 8u), dim3(512),
                         0, s, a);
      break;
    case 0:
    case 4:
      hipLaunchKernelGGL((star_sweep_lds<T, R, COND>), dim3((L.nlist + 3u) / 4u), dim3(256), 0,
                         s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_dt(const SweepLaunch &L, hipStream_t s) {
  switch (L.stencil) {
    case BK_STAR7: return launch_t<T, 1, false>(L, s);
    case BK_STAR13: return launch_t<T, 2, false>(L, s);
    case BK_STAR25: return launch_t<T, 4, false>(L, s);
    case BK_COND7: return launch_t<T, 1, true>(L, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_star(const SweepLaunch &L, hipStream_t s) {
  switch (L.dtype) {
    case BK_F64: return launch_dt<double>(L, s);
    case BK_F32: return launch_dt<float>(L, s);
    case BK_BF16: return launch_dt<bf16>(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_star_batched(int stencil, int dtype, const BatchTable *tab, uint32_t tick,
                               hipStream_t s) {
  switch (dtype) {
    case BK_F64: return launch_batched_dt<double>(stencil, tab, tick, s);
    case BK_F32: return launch_batched_dt<float>(stencil, tab, tick, s);
    case BK_BF16: return launch_batched_dt<bf16>(stencil, tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_star_columns_batched(int stencil, int dtype, const BatchTable *tab,
                                       uint32_t tick, hipStream_t s) {
  switch (dtype) {
    case BK_F64: return launch_cols_batched_dt<double>(stencil, tab, tick, s);
    case BK_F32: return launch_cols_batched_dt<float>(stencil, tab, tick, s);
    case BK_BF16: return launch_cols_batched_dt<bf16>(stencil, tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

uint32_t star_quads_batched_capacity(int stencil, int dtype) {
  switch (dtype) {
    case BK_F64: return quads_batched_capacity_dt<double>(stencil);
    case BK_F32: return quads_batched_capacity_dt<float>(stencil);
    case BK_BF16: return quads_batched_capacity_dt<bf16>(stencil);
    default: return 0;
  }
}

hipError_t launch_star_quads_batched(int stencil, int dtype, const BatchTable *tab,
                                     uint32_t tick, hipStream_t s) {
  switch (dtype) {
    case BK_F64: return launch_quads_batched_dt<double>(stencil, tab, tick, s);
    case BK_F32: return launch_quads_batched_dt<float>(stencil, tab, tick, s);
    case BK_BF16: return launch_quads_batched_dt<bf16>(stencil, tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_star_quads(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  switch (L.dtype) {
    case BK_F64: return launch_quads_dt<double>(L, P, s);
    case BK_F32: return launch_quads_dt<float>(L, P, s);
    case BK_BF16: return launch_quads_dt<bf16>(L, P, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_star_columns(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  switch (L.dtype) {
    case BK_F64: return launch_cols_dt<double>(L, P, s);
    case BK_F32: return launch_cols_dt<float>(L, P, s);
    case BK_BF16: return launch_cols_dt<bf16>(L, P, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bk
