// brick_hook_7pt.cpp -- bricklib user code that keeps its per-brick
// brick("7pt.py", ...) call site, on host and on device-resident bricks.
//
// The flow of the reference's external example (examples/external/main.cpp:
// 64-142): init_grid, randomArray/zeroArray, copyToBrick into interleaved
// in/out bricks, the array loop, the tiled array loop through
// tile("7pt.py", ...), and the brick loop whose body is
// brick("7pt.py", VSVEC, (BDIM), (VFOLD), b) -- then compareBrick.  The same
// brick loop runs a second time over device-resident copies of the fields
// (movBrickInfo / movBrickStorage): the brick(...) hook queues the bricks
// and runs them as one GPU sweep when the result is moved back.  Finally
// `sweeps` ping-pong sweeps through the hook on the device and on the host
// must agree bit for bit (same kernel order of the multiply-adds).
// Prints "result match" and exits 0 on success.
//
// Build (host compiler only), from the repository root:
//   g++ -O2 -std=c++17 -fopenmp -I include examples/brick_hook_7pt.cpp
//       -L bricklib_amd -lbricks -Wl,-rpath,$PWD/bricklib_amd -o brick_hook_7pt
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "brick-hip.h"
#include "brickcompare.h"
#include "bricksetup.h"
#include "multiarray.h"

#define VSVEC "HIP"
#define VFOLD 8, 8
#define TILE 8
#define BDIM TILE, TILE, TILE
#define GZ TILE
#define PADDING 8

bElem *coeff;

namespace {

using BrickT = Brick<Dim<BDIM>, Dim<VFOLD>>;

// every brick of the grid but its outer layer, through the hook
void sweep_bricks(unsigned *grid_ptr, long SB, BrickT &bIn, BrickT &bOut) {
#pragma omp parallel for collapse(2)
  for (long tk = 1; tk < SB - 1; ++tk)
    for (long tj = 1; tj < SB - 1; ++tj)
      for (long ti = 1; ti < SB - 1; ++ti) {
        unsigned b = grid_ptr[(tk * SB + tj) * SB + ti];
        brick("7pt.py", VSVEC, (BDIM), (VFOLD), b);
      }
}

bool same_bricks(const BrickStorage &a, const BrickStorage &b, unsigned *grid_ptr, long SB,
                 unsigned off) {
  for (long tk = 1; tk < SB - 1; ++tk)
    for (long tj = 1; tj < SB - 1; ++tj)
      for (long ti = 1; ti < SB - 1; ++ti) {
        const unsigned q = grid_ptr[(tk * SB + tj) * SB + ti];
        if (std::memcmp(a.dat.get() + q * a.step + off, b.dat.get() + q * b.step + off,
                        512 * sizeof(bElem)) != 0)
          return false;
      }
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  const long N = argc > 1 ? std::atol(argv[1]) : 64;
  const int sweeps = argc > 2 ? std::atoi(argv[2]) : 4;
  if (N % TILE) {
    std::fprintf(stderr, "N must be a multiple of %d\n", TILE);
    return 2;
  }
  const long STRIDE = N + 2 * (GZ + PADDING), STRIDEG = N + 2 * GZ, SB = STRIDEG / TILE;

  coeff = (bElem *)std::malloc(129 * sizeof(bElem));
  std::mt19937_64 mt(77);
  std::uniform_real_distribution<bElem> u(0, 1);
  bElem sum = 0;
  for (int i = 0; i < 7; ++i) sum += (coeff[i] = u(mt));
  for (int i = 0; i < 7; ++i) coeff[i] /= sum;

  unsigned *grid_ptr;
  auto bInfo = init_grid<3>(grid_ptr, {SB, SB, SB});
  bElem *in_ptr = randomArray({STRIDE, STRIDE, STRIDE});
  bElem *out_ptr = zeroArray({STRIDE, STRIDE, STRIDE});
  bElem *tile_ptr = zeroArray({STRIDE, STRIDE, STRIDE});
  auto A = [STRIDE](long k, long j, long i) { return (k * STRIDE + j) * STRIDE + i; };

  const unsigned bSize = cal_size<BDIM>::value;
  auto bStorage = BrickStorage::allocate(bInfo.nbricks, bSize * 2);
  BrickT bIn(&bInfo, bStorage, 0);
  BrickT bOut(&bInfo, bStorage, bSize);
  copyToBrick<3>({STRIDEG, STRIDEG, STRIDEG}, {PADDING, PADDING, PADDING}, {0, 0, 0}, in_ptr,
                 grid_ptr, bIn);

  // the array loop (stencils/3axis.cpp:32-37 form) and the tiled array loop
#pragma omp parallel for collapse(2)
  for (long k = GZ + PADDING; k < GZ + PADDING + N; ++k)
    for (long j = GZ + PADDING; j < GZ + PADDING + N; ++j)
      for (long i = GZ + PADDING; i < GZ + PADDING + N; ++i)
        out_ptr[A(k, j, i)] =
            coeff[5] * in_ptr[A(k + 1, j, i)] + coeff[6] * in_ptr[A(k - 1, j, i)] +
            coeff[3] * in_ptr[A(k, j + 1, i)] + coeff[4] * in_ptr[A(k, j - 1, i)] +
            coeff[1] * in_ptr[A(k, j, i + 1)] + coeff[2] * in_ptr[A(k, j, i - 1)] +
            coeff[0] * in_ptr[A(k, j, i)];
#define bIn(i, j, k) in_ptr[A(k, j, i)]
#define bOut(i, j, k) tile_ptr[A(k, j, i)]
#pragma omp parallel for collapse(2)
  for (long tk = GZ + PADDING; tk < GZ + PADDING + N; tk += TILE)
    for (long tj = GZ + PADDING; tj < GZ + PADDING + N; tj += TILE)
      for (long ti = GZ + PADDING; ti < GZ + PADDING + N; ti += TILE)
        tile("7pt.py", VSVEC, (BDIM), ("tk", "tj", "ti"));
#undef bIn
#undef bOut
  const bool tile_ok = compareArray({STRIDE, STRIDE, STRIDE}, out_ptr, tile_ptr);

  // one sweep through the hook on the host
  sweep_bricks(grid_ptr, SB, bIn, bOut);
  const bool host_ok = compareBrick<3>({N, N, N}, {PADDING, PADDING, PADDING}, {GZ, GZ, GZ},
                                       out_ptr, grid_ptr, bOut);

  // the same loop over device-resident copies
  auto dInfo = movBrickInfo(bInfo, gpuMemcpyHostToDevice);
  auto dStorage = movBrickStorage(bStorage, gpuMemcpyHostToDevice);
  {
    BrickT dIn(&dInfo, dStorage, 0), dOut(&dInfo, dStorage, bSize);
    sweep_bricks(grid_ptr, SB, dIn, dOut);
  }
  auto back = movBrickStorage(dStorage, gpuMemcpyDeviceToHost);  // runs the queued bricks
  BrickT hOut(&bInfo, back, bSize);
  const bool dev_ok = compareBrick<3>({N, N, N}, {PADDING, PADDING, PADDING}, {GZ, GZ, GZ},
                                      out_ptr, grid_ptr, hOut);
  const bool dev_same = same_bricks(back, bStorage, grid_ptr, SB, bSize);

  // ping-pong sweeps through the hook: host and device, the same bits
  auto hs = BrickStorage::allocate(bInfo.nbricks, bSize * 2);
  std::memcpy(hs.dat.get(), bStorage.dat.get(), bInfo.nbricks * bSize * 2 * sizeof(bElem));
  auto ds = movBrickStorage(hs, gpuMemcpyHostToDevice);
  BrickT h0(&bInfo, hs, 0), h1(&bInfo, hs, bSize);
  BrickT d0(&dInfo, ds, 0), d1(&dInfo, ds, bSize);
  for (int t = 0; t < sweeps; ++t) {
    sweep_bricks(grid_ptr, SB, t % 2 ? h1 : h0, t % 2 ? h0 : h1);
    sweep_bricks(grid_ptr, SB, t % 2 ? d1 : d0, t % 2 ? d0 : d1);
  }
  brick_flush();
  auto dback = movBrickStorage(ds, gpuMemcpyDeviceToHost);
  const bool pp_same = same_bricks(dback, hs, grid_ptr, SB, sweeps % 2 ? bSize : 0);

  std::printf("tile: %s, host brick(): %s, device brick(): %s (%s host bits), %d ping-pong "
              "sweeps: device %s host\n",
              tile_ok ? "ok" : "FAILED", host_ok ? "ok" : "FAILED", dev_ok ? "ok" : "FAILED",
              dev_same ? "same as" : "differs from", sweeps, pp_same ? "==" : "!=");
  bk_free(dInfo.adj);
  std::free(in_ptr);
  std::free(out_ptr);
  std::free(tile_ptr);
  std::free(grid_ptr);
  std::free(bInfo.adj);
  std::free(coeff);
  const bool ok = tile_ok && host_ok && dev_ok && dev_same && pp_same;
  std::printf(ok ? "result match\n" : "result mismatch!\n");
  return ok ? 0 : 1;
}
