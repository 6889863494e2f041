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
  A theta;
  A c[6 * R + 1];
};

template <class A, bool COND>
__device__ __forceinline__ A tapv(A x) {
  if constexpr (COND) return relu_ref(x);
  else return x;
}

template <class T, int R, bool COND>
__global__ __launch_bounds__(256) void star_sweep_lds(const StarArgs<T, R, COND> a) {
  using A = typename acc_of<T>::type;
  constexpr int P = 8 + 2 * R;
  __shared__ A tile[4][8][P][P];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int j = lane >> 3, i = lane & 7;
  const int fk = lane >> 3, fl = lane & 7;  // face-load lane mapping
  const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * 4u + (uint32_t)w;
  const bool active = slot < a.nlist;

  A v[8], vkm[R], vkp[R];
  uint32_t b = 0;
  if (active) {
    b = uniform(a.list[slot]);
    const uint32_t *ab = a.adj + (size_t)b * 27;
    const size_t st = a.in_step;
    const T *pc = a.in + (size_t)b * st;
    const T *pkm = a.in + (size_t)uniform(ab[4]) * st;
    const T *pkp = a.in + (size_t)uniform(ab[22]) * st;
    const T *pjm = a.in + (size_t)uniform(ab[10]) * st;
    const T *pjp = a.in + (size_t)uniform(ab[16]) * st;
    const T *pim = a.in + (size_t)uniform(ab[12]) * st;
    const T *pip = a.in + (size_t)uniform(ab[14]) * st;
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = widen(pc[k * 64 + lane]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      vkm[r] = widen(pkm[(8 - R + r) * 64 + lane]);
      vkp[r] = widen(pkp[r * 64 + lane]);
    }
    A fjm[R], fjp[R], fim[R], fip[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      fjm[r] = widen(pjm[fk * 64 + (8 - R + r) * 8 + fl]);
      fjp[r] = widen(pjp[fk * 64 + r * 8 + fl]);
      fim[r] = widen(pim[fk * 64 + fl * 8 + (8 - R + r)]);
      fip[r] = widen(pip[fk * 64 + fl * 8 + r]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) tile[w][k][R + j][R + i] = v[k];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      tile[w][fk][r][R + fl] = fjm[r];
      tile[w][fk][8 + R + r][R + fl] = fjp[r];
      tile[w][fk][R + fl][r] = fim[r];
      tile[w][fk][R + fl][8 + R + r] = fip[r];
    }
  }
  __syncthreads();
  if (!active) return;

  T *po = a.out + (size_t)b * a.out_step;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    auto X = [&](int kk) -> A {
      return kk < 0 ? vkm[kk + R] : (kk >= 8 ? vkp[kk - 8] : v[kk]);
    };
    A acc = a.c[0] * tapv<A, COND>(v[k]);
#pragma unroll
    for (int r = 1; r <= R; ++r) {
      acc = fmadd(a.c[6 * r - 5], tapv<A, COND>(tile[w][k][R + j][R + i + r]), acc);
      acc = fmadd(a.c[6 * r - 4], tapv<A, COND>(tile[w][k][R + j][R + i - r]), acc);
      acc = fmadd(a.c[6 * r - 3], tapv<A, COND>(tile[w][k][R + j + r][R + i]), acc);
      acc = fmadd(a.c[6 * r - 2], tapv<A, COND>(tile[w][k][R + j - r][R + i]), acc);
      acc = fmadd(a.c[6 * r - 1], tapv<A, COND>(X(k + r)), acc);
      acc = fmadd(a.c[6 * r], tapv<A, COND>(X(k - r)), acc);
    }
    if constexpr (COND) {
      const bool sel = a.cond_mode == 0 ? (acc > A(0)) : (v[k] > a.theta);
      acc = sel ? acc : -acc;
    }
    narrow(po[k * 64 + lane], acc);
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
  a.theta = (A)L.theta;
  for (int q = 0; q < 6 * R + 1; ++q) a.c[q] = (A)L.coeff[q];
  const uint32_t nwg = (L.nlist + 3u) / 4u;
  hipLaunchKernelGGL((star_sweep_lds<T, R, COND>), dim3(nwg), dim3(256), 0, s, a);
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

}  // namespace bk
