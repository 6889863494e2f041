// bk_shape.hip -- per-brick sweeps for bricks of other shapes and vector
// folds: Brick<Dim<4,8,16>,Dim<4,16>> and Brick<Dim<4,8,16>,Dim<8,8>>,
// plus the 8x8x8 brick through the same code (a candidate the tuner times
// against the dedicated 8x8x8 kernels).
//
// Element (k, j, i) of a (BK, BJ, BI) brick in fold (FJ, FI) sits at
//   vec * 64 + (j % FJ) * FI + i % FI,
//   vec = (k * (BJ / FJ) + j / FJ) * (BI / FI) + i / FI
// -- the reference accessor's decomposition (brick.h:235-246, 276-284):
// vectors over the brick divided by the fold, outermost first, lanes over
// the fold.  Every brick still holds 512 elements = 8 wave64 vectors, so a
// brick is 8 coalesced loads whatever the shape.
//
// One wave per listed brick, WPB waves (bricks) per workgroup.  Lane L of
// vector n holds point (k, j, i) of the fold decomposition:
//   - k-direction taps: the same lane of vector n +- r*VPP (VPP vectors per
//     k-plane) or of the k-1 / k+1 neighbour's planes, all in registers;
//   - i/j taps: an LDS tile [BK][BJ+2R][BI+2R] per wave holding the brick
//     and the R-deep faces of its four in-plane neighbours;
//   - the cube: an LDS box [BK+4][BJ+4][BI+4] with all 26 neighbours' cells.
// Each point is evaluated in the same fixed order as the 8x8x8 kernels
// (stars: centre, then i+r, i-r, j+r, j-r, k+r, k-r for r = 1..R; cube:
// (dk, dj, di) lexicographic), so a cell's value does not depend on the
// brick shape, fold or schedule.
#include "bk_device.h"
#include "bk_internal.h"

namespace bk {
namespace {

template <int LAY> struct Shape;
template <> struct Shape<BK_LAYOUT_8x8x8> {
  static constexpr int BK = 8, BJ = 8, BI = 8, FJ = 8, FI = 8;
};
template <> struct Shape<BK_LAYOUT_4x8x16> {
  static constexpr int BK = 4, BJ = 8, BI = 16, FJ = 4, FI = 16;
};
template <> struct Shape<BK_LAYOUT_4x8x16_F88> {
  static constexpr int BK = 4, BJ = 8, BI = 16, FJ = 8, FI = 8;
};

template <int LAY>
struct Fold : Shape<LAY> {
  using S = Shape<LAY>;
  static constexpr int NJ = S::BJ / S::FJ, NI = S::BI / S::FI;
  static constexpr int VPP = NJ * NI;  // vectors per k-plane
  static_assert(S::FJ * S::FI == 64 && S::BK * S::BJ * S::BI == 512, "wave64 fold of 512");
  __device__ static constexpr int off(int k, int j, int i) {
    return ((k * NJ + j / S::FJ) * NI + i / S::FI) * 64 + (j % S::FJ) * S::FI + i % S::FI;
  }
  // coordinates of vector n, lane L
  __device__ static constexpr int kof(int n) { return n / VPP; }
  __device__ static constexpr int jof(int n, int L) { return ((n / NI) % NJ) * S::FJ + L / S::FI; }
  __device__ static constexpr int iof(int n, int L) { return (n % NI) * S::FI + L % S::FI; }
};

template <class T, int NC>
struct ShapeArgs {
  using A = typename acc_of<T>::type;
  const uint32_t *__restrict__ adj;
  const uint32_t *__restrict__ list;
  const T *__restrict__ in;
  T *__restrict__ out;
  uint64_t in_step;
  uint64_t out_step;
  uint32_t nlist;
  int cond_mode;
  A theta;
  A c[NC];  // coefficients in the accumulate type
};

template <class A, bool COND>
__device__ __forceinline__ A tapv(A x) {
  if constexpr (COND) return relu_ref(x);
  else return x;
}

// ---------------------------------------------------------------------------
// stars and cond
// ---------------------------------------------------------------------------
template <class T, int R, bool COND, int LAY, int WPB>
__global__ __launch_bounds__(64 * WPB) void star_sweep_shape(const ShapeArgs<T, 6 * R + 1> a) {
  using A = typename acc_of<T>::type;
  using F = Fold<LAY>;
  constexpr int BK = F::BK, BJ = F::BJ, BI = F::BI, VPP = F::VPP;
  constexpr int PJ = BJ + 2 * R, PI = BI + 2 * R, PL = PJ * PI;
  static_assert(R <= BK, "radius beyond the k-neighbour brick");
  __shared__ A tiles[WPB][BK * PL];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * WPB + (uint32_t)w;
  const bool active = slot < a.nlist;
  A *tile = tiles[w];
  A v[8], km[R * VPP], kp[R * VPP];
  uint32_t b = 0;
  if (active) {
    b = uniform(a.list[slot]);
    const uint32_t *ab = a.adj + (size_t)b * 27;
    const size_t st = a.in_step;
    const T *pc = a.in + (size_t)b * st;
    const T *pkm = a.in + (size_t)uniform(ab[4]) * st;
    const T *pkp = a.in + (size_t)uniform(ab[22]) * st;
#pragma unroll
    for (int n = 0; n < 8; ++n) v[n] = widen(pc[n * 64 + lane]);
#pragma unroll
    for (int q = 0; q < R * VPP; ++q) {
      km[q] = widen(pkm[((BK - R) * VPP + q) * 64 + lane]);
      kp[q] = widen(pkp[q * 64 + lane]);
    }
    // the brick into the tile
#pragma unroll
    for (int n = 0; n < 8; ++n)
      tile[F::kof(n) * PL + (R + F::jof(n, lane)) * PI + R + F::iof(n, lane)] = v[n];
    // j faces: rows [BJ-R, BJ) of j-1 and [0, R) of j+1, all k and i
    const T *pjm = a.in + (size_t)uniform(ab[10]) * st;
    const T *pjp = a.in + (size_t)uniform(ab[16]) * st;
    constexpr int NJF = BK * R * BI;
#pragma unroll
    for (int q = lane; q < NJF; q += 64) {
      const int i = q % BI, r = (q / BI) % R, k = q / (BI * R);
      tile[k * PL + r * PI + R + i] = widen(pjm[F::off(k, BJ - R + r, i)]);
      tile[k * PL + (BJ + R + r) * PI + R + i] = widen(pjp[F::off(k, r, i)]);
    }
    // i faces: columns [BI-R, BI) of i-1 and [0, R) of i+1, all k and j
    const T *pim = a.in + (size_t)uniform(ab[12]) * st;
    const T *pip = a.in + (size_t)uniform(ab[14]) * st;
    constexpr int NIF = BK * BJ * R;
#pragma unroll
    for (int q = lane; q < NIF; q += 64) {
      const int r = q % R, j = (q / R) % BJ, k = q / (R * BJ);
      tile[k * PL + (R + j) * PI + r] = widen(pim[F::off(k, j, BI - R + r)]);
      tile[k * PL + (R + j) * PI + BI + R + r] = widen(pip[F::off(k, j, r)]);
    }
  }
  __syncthreads();
  if (!active) return;
  A c[6 * R + 1];
#pragma unroll
  for (int q = 0; q < 6 * R + 1; ++q) c[q] = a.c[q];
  T *__restrict__ po = a.out + (size_t)b * a.out_step;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int k = F::kof(n), j = F::jof(n, lane), i = F::iof(n, lane);
    const A *t = tile + k * PL + (R + j) * PI + R + i;
    // the same lane's value r planes up or down
    auto X = [&](int r) -> A {
      const int kk = k + r;
      return kk < 0 ? km[(kk + R) * VPP + n % VPP]
                    : (kk >= BK ? kp[(kk - BK) * VPP + n % VPP] : v[n + r * VPP]);
    };
    A acc = c[0] * tapv<A, COND>(v[n]);
#pragma unroll
    for (int r = 1; r <= R; ++r) {
      acc = fmadd(c[6 * r - 5], tapv<A, COND>(t[r]), acc);
      acc = fmadd(c[6 * r - 4], tapv<A, COND>(t[-r]), acc);
      acc = fmadd(c[6 * r - 3], tapv<A, COND>(t[r * PI]), acc);
      acc = fmadd(c[6 * r - 2], tapv<A, COND>(t[-r * PI]), acc);
      acc = fmadd(c[6 * r - 1], tapv<A, COND>(X(r)), acc);
      acc = fmadd(c[6 * r], tapv<A, COND>(X(-r)), acc);
    }
    if constexpr (COND) {
      const bool sel = a.cond_mode == 0 ? (acc > A(0)) : (v[n] > a.theta);
      acc = sel ? acc : -acc;
    }
    store_nt(&po[n * 64 + lane], acc);
  }
}

// ---------------------------------------------------------------------------
// the 125-point cube: the whole 5x5x5 neighbourhood from an LDS box
// ---------------------------------------------------------------------------
template <class T, int LAY, int WPB>
__global__ __launch_bounds__(64 * WPB) void cube_sweep_shape(const ShapeArgs<T, 125> a) {
  using A = typename acc_of<T>::type;
  using F = Fold<LAY>;
  constexpr int R = 2;
  constexpr int BK = F::BK, BJ = F::BJ, BI = F::BI;
  constexpr int PJ = BJ + 2 * R, PI = BI + 2 * R, PK = BK + 2 * R, PL = PJ * PI;
  __shared__ A boxes[WPB][PK * PL];
  __shared__ uint32_t nbr[WPB][27];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * WPB + (uint32_t)w;
  const bool active = slot < a.nlist;
  A *box = boxes[w];
  uint32_t b = 0;
  if (active) {
    b = uniform(a.list[slot]);
    if (lane < 27) nbr[w][lane] = a.adj[(size_t)b * 27 + lane];
    const T *pc = a.in + (size_t)b * a.in_step;
#pragma unroll
    for (int n = 0; n < 8; ++n)
      box[(R + F::kof(n)) * PL + (R + F::jof(n, lane)) * PI + R + F::iof(n, lane)] =
          widen(pc[n * 64 + lane]);
  }
  __syncthreads();
  if (active) {
    // the cells of the box outside the brick, each from its neighbour
    for (int q = lane; q < PK * PL; q += 64) {
      const int x = q % PI, y = (q / PI) % PJ, z = q / PL;
      const int k = z - R, j = y - R, i = x - R;
      const int kd = (k + BK) / BK, jd = (j + BJ) / BJ, id = (i + BI) / BI;
      if (kd == 1 && jd == 1 && id == 1) continue;
      const uint32_t nb = nbr[w][kd * 9 + jd * 3 + id];
      box[q] = widen(a.in[(size_t)nb * a.in_step +
                          F::off((k + BK) % BK, (j + BJ) % BJ, (i + BI) % BI)]);
    }
  }
  __syncthreads();
  if (!active) return;
  T *__restrict__ po = a.out + (size_t)b * a.out_step;
#pragma unroll 1
  for (int n = 0; n < 8; ++n) {
    const int k = F::kof(n), j = F::jof(n, lane), i = F::iof(n, lane);
    const A *t = box + k * PL + j * PI + i;  // the (-2, -2, -2) corner
    A acc = a.c[0] * t[0];
#pragma unroll
    for (int q = 1; q < 125; ++q) {
      const int dk = q / 25, dj = (q / 5) % 5, di = q % 5;
      acc = fmadd(a.c[q], t[dk * PL + dj * PI + di], acc);
    }
    store_nt(&po[n * 64 + lane], acc);
  }
}

template <class T, int NC>
ShapeArgs<T, NC> make_args(const SweepLaunch &L) {
  using A = typename acc_of<T>::type;
  ShapeArgs<T, NC> a;
  a.adj = L.adj;
  a.list = L.list;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = L.nlist;
  a.cond_mode = L.cond_mode;
  a.theta = (A)L.theta;
  for (int q = 0; q < NC; ++q) a.c[q] = (A)L.coeff[q];
  return a;
}

template <class T, int R, bool COND, int LAY, int WPB>
hipError_t launch_star_w(const SweepLaunch &L, hipStream_t s) {
  const auto a = make_args<T, 6 * R + 1>(L);
  const uint32_t nwg = (L.nlist + WPB - 1) / WPB;
  hipLaunchKernelGGL((star_sweep_shape<T, R, COND, LAY, WPB>), dim3(nwg), dim3(64 * WPB), 0, s,
                     a);
  return hipGetLastError();
}

template <class T, int LAY, int WPB>
hipError_t launch_cube_w(const SweepLaunch &L, hipStream_t s) {
  const auto a = make_args<T, 125>(L);
  const uint32_t nwg = (L.nlist + WPB - 1) / WPB;
  hipLaunchKernelGGL((cube_sweep_shape<T, LAY, WPB>), dim3(nwg), dim3(64 * WPB), 0, s, a);
  return hipGetLastError();
}

template <class T, int LAY, int WPB>
hipError_t launch_w(const SweepLaunch &L, hipStream_t s) {
  switch (L.stencil) {
    case BK_STAR7: return launch_star_w<T, 1, false, LAY, WPB>(L, s);
    case BK_STAR13: return launch_star_w<T, 2, false, LAY, WPB>(L, s);
    case BK_STAR25: return launch_star_w<T, 4, false, LAY, WPB>(L, s);
    case BK_COND7: return launch_star_w<T, 1, true, LAY, WPB>(L, s);
    case BK_CUBE125: return launch_cube_w<T, LAY, WPB>(L, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T, int LAY>
hipError_t launch_lay(const SweepLaunch &L, hipStream_t s) {
  switch (L.wg_bricks) {
    case 1: return launch_w<T, LAY, 1>(L, s);
    case 2: return launch_w<T, LAY, 2>(L, s);
    case 0:
    case 4: return launch_w<T, LAY, 4>(L, s);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t launch_dt(const SweepLaunch &L, hipStream_t s) {
  switch (L.layout) {
    case BK_LAYOUT_8x8x8: return launch_lay<T, BK_LAYOUT_8x8x8>(L, s);
    case BK_LAYOUT_4x8x16: return launch_lay<T, BK_LAYOUT_4x8x16>(L, s);
    case BK_LAYOUT_4x8x16_F88: return launch_lay<T, BK_LAYOUT_4x8x16_F88>(L, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_shape(const SweepLaunch &L, hipStream_t s) {
  if (L.nlist == 0) return hipSuccess;
  switch (L.dtype) {
    case BK_F64: return launch_dt<double>(L, s);
    case BK_F32: return launch_dt<float>(L, s);
    case BK_BF16: return launch_dt<bf16>(L, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bk
