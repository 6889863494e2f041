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

// One quad unit: the four waves of a workgroup (member mm = wave index)
// march L levels from the unit's anchor; mk[l] holds level l's store mask
// (bit mm: store member mm).  Every wave of the workgroup must call it with
// the same unit (it synchronises the workgroup once per level).  Shared by
// the single-grid and the batched kernels.
template <class T, int R, bool COND>
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

  A cur[8], nxt[8], km[R];
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
    // the two outer faces first (waited for before the tile is filled),
    // then the prefetch of the next level, which stays in flight through
    // the compute below
    A fi[R], fj[R];
    {
      const T *pi = in + (size_t)uniform(ab[si]) * st;
      const T *pj = in + (size_t)uniform(ab[sj]) * st;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int k = q * (8 / R) + fkq;
        fi[q] = widen(pi[k * 64 + fc * 8 + (mi ? frr : 8 - R + frr)]);
        fj[q] = widen(pj[k * 64 + (mj ? frr : 8 - R + frr) * 8 + fc]);
      }
    }
    const T *pu = in + (size_t)up * st;
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) nxt[k] = widen(pu[k * 64 + lane]);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) nxt[r] = widen(pu[r * 64 + lane]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) tile[QT::at(k, ty, tx)] = cur[k];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int k = q * (8 / R) + fkq;
      tile[QT::at(k, R + 8 * mj + fc, fix)] = fi[q];
      tile[QT::at(k, fjy, R + 8 * mi + fc)] = fj[q];
    }
    __syncthreads();
    if ((mk[l] >> mm) & 1u) {
      T *po = out + (size_t)b * ost;
      // fused refresh: b's ghost images receive its output too
      const uint32_t *imgs = nullptr;
      uint32_t nimg = 0;
      if (img_off) {
        const uint32_t i0 = uniform(img_off[b]);
        nimg = uniform(img_off[b + 1]) - i0;
        imgs = img + i0;
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        auto X = [&](int kk) -> A {
          return kk < 0 ? km[kk + R] : (kk >= 8 ? nxt[kk - 8] : cur[kk]);
        };
        A acc = c[0] * tapv<A, COND>(cur[k]);
#pragma unroll
        for (int r = 1; r <= R; ++r) {
          acc = fmadd(c[6 * r - 5], tapv<A, COND>(tile[QT::at(k, ty, tx + r)]), acc);
          acc = fmadd(c[6 * r - 4], tapv<A, COND>(tile[QT::at(k, ty, tx - r)]), acc);
          acc = fmadd(c[6 * r - 3], tapv<A, COND>(tile[QT::at(k, ty + r, tx)]), acc);
          acc = fmadd(c[6 * r - 2], tapv<A, COND>(tile[QT::at(k, ty - r, tx)]), acc);
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
#pragma unroll
    for (int r = 0; r < R; ++r) km[r] = cur[8 - R + r];
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
    }
    b = up;
    __syncthreads();  // every wave's reads of this level's tile are done
  }
}

template <class T, int R, bool COND>
__global__ __launch_bounds__(256) void star_sweep_quads(const StarArgs<T, R, COND> a) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[8 * QuadTile<R, sizeof(A)>::PL];
  const uint32_t u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= a.nunits) return;  // uniform over the workgroup
  A c[6 * R + 1];
#pragma unroll
  for (int q = 0; q < 6 * R + 1; ++q) c[q] = a.c[q];
  quad_unit<T, R, COND>(a.adj, a.in, a.out, a.in_step, a.out_step, a.anchor[u], (int)a.len[u],
                        a.mask + (size_t)u * a.kcol, c, a.theta, a.cond_mode, a.img_off, a.img,
                        tile);
}

// Batched quad schedule: one persistent launch serves every request of a
// group table whose bricks come as a quad plan; work item = one quad unit of
// one request, taken by a whole workgroup.
template <class T, int R, bool COND>
__global__ __launch_bounds__(256) void star_sweep_quads_batched(const BatchTable *tab_,
                                                                uint32_t tick) {
  using A = typename acc_of<T>::type;
  __shared__ A tile[8 * QuadTile<R, sizeof(A)>::PL];
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
  hipLaunchKernelGGL((star_sweep_quads<T, R, COND>), dim3(P.nunits), dim3(256), 0, s, a);
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
hipError_t launch_quads_batched_t(const BatchTable *tab, uint32_t tick, hipStream_t s) {
  static int nblk = 0;
  if (nblk == 0) {
    int dev = 0, cus = 256, per = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per, star_sweep_quads_batched<T, R, COND>, 256, 0);
    nblk = cus * (per > 0 ? per : 1);
  }
  hipLaunchKernelGGL((star_sweep_quads_batched<T, R, COND>), dim3(nblk), dim3(256), 0, s, tab,
                     tick);
  return hipGetLastError();
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
