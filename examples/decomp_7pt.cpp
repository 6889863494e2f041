// decomp_7pt.cpp -- BrickDecomp user code on the MI355X library.
//
// The flow of the reference's single-node decomposition driver
// (single/mpi.cpp:58-120) on one GPU: BrickDecomp<3,8,8,8> with a ghost zone
// of one brick, initialize(skin list), the grid read back through
// bDecomp[k][j][i], copyToBrick into two interleaved fields, then
//   (a) `sweeps` steps of exchange() + brick_sweep on device storage -- the
//       reference's exchange-then-compute loop (weak/main.cpp:120-150), the
//       exchange being the periodic wrap of the one subdomain;
//   (b) the same steps as one brick_sweeps call over the decomposition's
//       device_pairs() (the refresh inside the sweep loop);
// each checked against `sweeps` host array sweeps on a periodic array.
// Prints "result match" and exits 0 on success.
//
// Usage: decomp_7pt [n=32] [sweeps=4] [good|normal|bad]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "brick-decomp.h"
#include "brick-hip.h"
#include "brickcompare.h"
#include "bricksetup.h"

namespace {

constexpr long GZ = 8, PAD = 8;

void periodic_fill(std::vector<bElem> &a, long n) {
  const long S = n + 2 * (GZ + PAD);
  for (long z = PAD; z < S - PAD; ++z)
    for (long y = PAD; y < S - PAD; ++y)
      for (long x = PAD; x < S - PAD; ++x) {
        auto wrap = [n](long v) { return ((v - PAD - GZ) % n + n) % n + PAD + GZ; };
        a[(z * S + y) * S + x] = a[(wrap(z) * S + wrap(y)) * S + wrap(x)];
      }
}

void array_7pt(const std::vector<bElem> &in, std::vector<bElem> &out, long n, const double *c) {
  const long S = n + 2 * (GZ + PAD), o = PAD + GZ;
#pragma omp parallel for collapse(2)
  for (long k = o; k < o + n; ++k)
    for (long j = o; j < o + n; ++j)
      for (long i = o; i < o + n; ++i) {
        auto A = [&](long kk, long jj, long ii) { return in[(kk * S + jj) * S + ii]; };
        out[(k * S + j) * S + i] = c[5] * A(k + 1, j, i) + c[6] * A(k - 1, j, i) +
                                   c[3] * A(k, j + 1, i) + c[4] * A(k, j - 1, i) +
                                   c[1] * A(k, j, i + 1) + c[2] * A(k, j, i - 1) +
                                   c[0] * A(k, j, i);
      }
}

}  // namespace

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 32;
  const int sweeps = argc > 2 ? std::atoi(argv[2]) : 4;
  const char *skin = argc > 3 ? argv[3] : "good";
  if (n % 8 || n < 16 || sweeps < 1) {
    std::fprintf(stderr, "n must be a multiple of 8, at least 16; sweeps >= 1\n");
    return 2;
  }
  const std::vector<BitSet> &skinlist = !std::strcmp(skin, "normal") ? skin3d_normal
                                        : !std::strcmp(skin, "bad")  ? skin3d_bad
                                                                     : skin3d_good;
  const long nb = n / 8, SB = nb + 2, S = n + 2 * (GZ + PAD);

  double coeff[7];
  std::mt19937_64 mt(77);
  std::uniform_real_distribution<double> u(0, 1);
  double sum = 0;
  for (double &c : coeff) sum += (c = u(mt));
  for (double &c : coeff) c /= sum;
  std::vector<bElem> in((size_t)S * S * S, 0.0);
  std::uniform_real_distribution<double> v(-1, 1);
  for (long z = PAD + GZ; z < PAD + GZ + n; ++z)
    for (long y = PAD + GZ; y < PAD + GZ + n; ++y)
      for (long x = PAD + GZ; x < PAD + GZ + n; ++x) in[(z * S + y) * S + x] = v(mt);
  periodic_fill(in, n);

  // the decomposition and its grid (ghost included), as the reference reads it
  BrickDecomp<3, 8, 8, 8> bDecomp({(unsigned)n, (unsigned)n, (unsigned)n}, 8, 2);
  bDecomp.initialize(skinlist);
  BrickInfo<3> bInfo = bDecomp.getBrickInfo();
  std::vector<unsigned> grid((size_t)SB * SB * SB);
  for (long k = 0; k < SB; ++k)
    for (long j = 0; j < SB; ++j)
      for (long i = 0; i < SB; ++i) grid[(k * SB + j) * SB + i] = bDecomp[k][j][i];
  auto bStorage = bInfo.allocate(512 * 2);
  std::memset(bStorage.dat.get(), 0, (size_t)bInfo.nbricks * 1024 * sizeof(bElem));
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bIn(&bInfo, bStorage, 0);
  copyToBrick<3>({n + 2 * GZ, n + 2 * GZ, n + 2 * GZ}, {PAD, PAD, PAD}, {0, 0, 0}, in.data(),
                 grid.data(), bIn);

  std::vector<bElem> a = in, b((size_t)S * S * S, 0.0);
  for (int t = 0; t < sweeps; ++t) {
    array_7pt(a, b, n, coeff);
    periodic_fill(b, n);
    a.swap(b);
  }

  auto dInfo = movBrickInfo(bInfo, gpuMemcpyHostToDevice);
  const std::vector<unsigned> interior = bDecomp.interior();
  unsigned *d_list = bk_upload_u32(interior);
  const unsigned nlist = (unsigned)interior.size();
  const BrickStencil st = brick_stencil("7pt.py", coeff);

  // (a) exchange() + one sweep per step, ping-pong between the two fields
  auto dA = movBrickStorage(bStorage, gpuMemcpyHostToDevice);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> f0(&dInfo, dA, 0), f1(&dInfo, dA, 512);
  for (int t = 0; t < sweeps; ++t) {
    bDecomp.exchange(dA);
    if (t % 2 == 0)
      brick_sweep(st, f1, f0, d_list, nlist);
    else
      brick_sweep(st, f0, f1, d_list, nlist);
  }
  auto backA = movBrickStorage(dA, gpuMemcpyDeviceToHost);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> resA(&bInfo, backA, sweeps % 2 ? 512 : 0);
  auto eA = compareBrickStats<3>({n, n, n}, {PAD, PAD, PAD}, {GZ, GZ, GZ}, a.data(), grid.data(),
                                 resA);
  std::printf("exchange + brick_sweep, %d steps (%s): max abs %.3e, max rel %.3e\n", sweeps,
              skin, eA.max_abs, eA.max_rel);

  // (b) brick_sweeps over the decomposition's exchange pairs
  auto dB = movBrickStorage(bStorage, gpuMemcpyHostToDevice);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> g0(&dInfo, dB, 0), g1(&dInfo, dB, 512);
  unsigned npairs = 0;
  const unsigned *d_pairs = bDecomp.device_pairs(npairs);
  brick_sweeps(st, g0, g1, d_list, nlist, d_pairs, npairs, sweeps);
  auto backB = movBrickStorage(dB, gpuMemcpyDeviceToHost);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> resB(&bInfo, backB, sweeps % 2 ? 512 : 0);
  auto eB = compareBrickStats<3>({n, n, n}, {PAD, PAD, PAD}, {GZ, GZ, GZ}, a.data(), grid.data(),
                                 resB);
  std::printf("brick_sweeps over device_pairs (%u pairs): max abs %.3e, max rel %.3e\n", npairs,
              eB.max_abs, eB.max_rel);
  // the two device paths run the same kernels on the same data: bit-equal
  const bool same = std::memcmp(backA.dat.get(), backB.dat.get(),
                                (size_t)bInfo.nbricks * 1024 * sizeof(bElem)) == 0;
  std::printf("paths (a) and (b) %s\n", same ? "bit-identical" : "DIFFER");

  bk_free(d_list);
  bk_free(dInfo.adj);
  const bool ok = same && eA.max_rel <= 1e-10 && eB.max_rel <= 1e-10;
  std::printf(ok ? "result match\n" : "result mismatch!\n");
  return ok ? 0 : 1;
}
