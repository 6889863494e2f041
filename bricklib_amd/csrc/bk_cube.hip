// bk_cube.hip -- 125-point (5x5x5, radius 2) cube stencil sweep over 8x8x8
// bricks in the 8x8 fold, for gfx950.
//
// Replaces the codegen brick kernel of weak/main.hip.cpp:32-40 built from
// stencils/mpi125pt.py:12-162 (here one coefficient per offset).
//
// One wave64 per brick, four bricks per workgroup.  The wave gathers the
// brick's 12x12x12 neighbourhood (its own 512 points plus the 2-deep shell
// from all 26 neighbours, found through the adjacency table) into LDS, then
// lane L = j*8 + i accumulates the eight outputs of column (j, i): for each
// input plane kk it reads the 5x5 window once and feeds it to every output
// plane within reach, so each LDS value serves up to five FMAs.  Per point
// the taps accumulate in (dk, dj, di) lexicographic order, starting with a
// multiply -- the coefficient order of the API.
#include "bk_device.h"
#include "bk_internal.h"

namespace bk {
namespace {

template <class T>
struct CubeArgs {
  using A = typename acc_of<T>::type;
  const uint32_t *__restrict__ adj;
  const uint32_t *__restrict__ list;
  const T *__restrict__ in;
  T *__restrict__ out;
  uint64_t in_step;
  uint64_t out_step;
  uint32_t nlist;
  A c[125];
};

template <class T>
__global__ __launch_bounds__(256) void cube_sweep_lds(const CubeArgs<T> a) {
  using A = typename acc_of<T>::type;
  constexpr int R = 2, P = 8 + 2 * R;  // 12
  __shared__ A box[4][P][P][P];
  __shared__ uint32_t sadj[4][32];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int j = lane >> 3, i = lane & 7;
  const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * 4u + (uint32_t)w;
  const bool active = slot < a.nlist;

  uint32_t b = 0;
  if (active) {
    b = uniform(a.list[slot]);
    if (lane < 27) sadj[w][lane] = a.adj[(size_t)b * 27 + lane];
  }
  __syncthreads();
  if (active) {
    const size_t st = a.in_step;
#pragma unroll
    for (int t = 0; t < 27; ++t) {  // 27 * 64 == 12^3
      const int q = t * 64 + lane;
      const int kk = q / (P * P), rem = q - kk * P * P;
      const int jj = rem / P, ii = rem - jj * P;
      const int k = kk - R + 8, jn = jj - R + 8, in_ = ii - R + 8;  // shifted by +8
      const int s = (k >> 3) * 9 + (jn >> 3) * 3 + (in_ >> 3);
      const int e = (k & 7) * 64 + (jn & 7) * 8 + (in_ & 7);
      box[w][kk][jj][ii] = widen(a.in[(size_t)sadj[w][s] * st + e]);
    }
  }
  __syncthreads();
  if (!active) return;

  A acc[8];
#pragma unroll
  for (int kk = 0; kk < P; ++kk) {
#pragma unroll
    for (int dj = 0; dj < 5; ++dj) {
#pragma unroll
      for (int di = 0; di < 5; ++di) {
        const A x = box[w][kk][j + dj][i + di];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int dk = kk - k;  // 0..4 when plane kk reaches output k
          if (dk < 0 || dk > 4) continue;
          const A c = a.c[dk * 25 + dj * 5 + di];
          if (dk == 0 && dj == 0 && di == 0) acc[k] = c * x;
          else acc[k] = fmadd(c, x, acc[k]);
        }
      }
    }
  }
  T *po = a.out + (size_t)b * a.out_step;
#pragma unroll
  for (int k = 0; k < 8; ++k) narrow(po[k * 64 + lane], acc[k]);
}

template <class T>
hipError_t launch_t(const SweepLaunch &L, hipStream_t s) {
  using A = typename acc_of<T>::type;
  CubeArgs<T> a;
  a.adj = L.adj;
  a.list = L.list;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = L.nlist;
  for (int q = 0; q < 125; ++q) a.c[q] = (A)L.coeff[q];
  const uint32_t nwg = (L.nlist + 3u) / 4u;
  hipLaunchKernelGGL((cube_sweep_lds<T>), dim3(nwg), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_cube(const SweepLaunch &L, hipStream_t s) {
  switch (L.dtype) {
    case BK_F64: return launch_t<double>(L, s);
    case BK_F32: return launch_t<float>(L, s);
    case BK_BF16: return launch_t<bf16>(L, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bk
