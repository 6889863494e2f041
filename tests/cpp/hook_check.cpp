// hook_check.cpp -- the brick(...) / tile(...) hooks of vecscatter.h on host
// bricks, for the parity test (tests/test_cpp_dropin.py).
//
//   hook_check <nb> <stencil.py> <out.bin>
//
// Builds the (nb+2)^3 init_grid, copies a random array into interleaved
// in/out bricks, sweeps the interior bricks through brick(stencil, ...),
// and writes the storage (nbricks x 1024 doubles: input at 0, output at
// 512) to out.bin.  The same sweep through tile(...) over the array must
// match compareArray-wise; exits 1 otherwise.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "brickcompare.h"
#include "bricksetup.h"
#include "multiarray.h"

bElem *coeff;

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  const long nb = std::atol(argv[1]), SB = nb + 2, N = nb * 8, GZ = 8, PAD = 8;
  const long STRIDE = N + 2 * (GZ + PAD), STRIDEG = N + 2 * GZ;
  const char *file = argv[2];
  coeff = randomArray({129});
  for (int q = 0; q < 7; ++q) coeff[q] = coeff[q] * 2 - 0.9;  // mixed signs
  unsigned *grid_ptr;
  auto bInfo = init_grid<3>(grid_ptr, {SB, SB, SB});
  bElem *in_ptr = randomArray({STRIDE, STRIDE, STRIDE});
  for (long l = 0; l < STRIDE * STRIDE * STRIDE; ++l) in_ptr[l] = 2 * in_ptr[l] - 1;
  bElem *tile_ptr = zeroArray({STRIDE, STRIDE, STRIDE});
  bElem *arr_ptr = zeroArray({STRIDE, STRIDE, STRIDE});
  auto bStorage = BrickStorage::allocate(bInfo.nbricks, 1024);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bIn(&bInfo, bStorage, 0), bOut(&bInfo, bStorage, 512);
  copyToBrick<3>({STRIDEG, STRIDEG, STRIDEG}, {PAD, PAD, PAD}, {0, 0, 0}, in_ptr, grid_ptr, bIn);
#pragma omp parallel for collapse(2)
  for (long tk = 1; tk < SB - 1; ++tk)
    for (long tj = 1; tj < SB - 1; ++tj)
      for (long ti = 1; ti < SB - 1; ++ti) {
        unsigned b = grid_ptr[(tk * SB + tj) * SB + ti];
        brick(file, "HIP", (8, 8, 8), (8, 8), b);
      }
  auto A = [STRIDE](long k, long j, long i) { return (k * STRIDE + j) * STRIDE + i; };
#define bIn(i, j, k) in_ptr[A(k, j, i)]
#define bOut(i, j, k) tile_ptr[A(k, j, i)]
#pragma omp parallel for collapse(2)
  for (long tk = GZ + PAD; tk < GZ + PAD + N; tk += 8)
    for (long tj = GZ + PAD; tj < GZ + PAD + N; tj += 8)
      for (long ti = GZ + PAD; ti < GZ + PAD + N; ti += 8)
        tile(file, "FLEX", (8, 8, 8), ("tk", "tj", "ti"), (1, 1, 4));
#undef bIn
#undef bOut
  // the bricks' output back to an array: the tile result must be equal
  copyFromBrick<3>({N, N, N}, {PAD, PAD, PAD}, {GZ, GZ, GZ}, arr_ptr, grid_ptr, bOut);
  const bool same = compareArray({STRIDE, STRIDE, STRIDE}, arr_ptr, tile_ptr);
  FILE *f = std::fopen(argv[3], "wb");
  if (!f) return 3;
  std::fwrite(bStorage.dat.get(), sizeof(bElem), (size_t)bInfo.nbricks * 1024, f);
  std::fwrite(coeff, sizeof(bElem), 7, f);
  std::fclose(f);
  std::printf("%s\n", same ? "tile == brick" : "tile != brick");
  std::free(in_ptr);
  std::free(tile_ptr);
  std::free(arr_ptr);
  std::free(coeff);
  std::free(grid_ptr);
  std::free(bInfo.adj);
  return same ? 0 : 1;
}
