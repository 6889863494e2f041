// drop_in_7pt.cpp -- bricklib user code on the MI355X library.
//
// The flow of the reference's d3pt7 test (stencils/3axis.cpp:14-92) and its
// GPU twin (stencils/3axis.hip.cpp:81-197): build the brick grid with
// init_grid, copy a random array into interleaved in/out bricks with
// copyToBrick, run the array loop on the host, run the brick sweep on the
// GPU, copy back and check with compareBrick.  The only GPU-specific lines
// are movBrickInfo / movBrickStorage / brick_sweep(s); the per-brick loop
// with brick("7pt.py", ...) becomes one brick_sweep call.
//
// It then runs `sweeps` ping-pong sweeps with the on-device periodic ghost
// refresh and checks them against the same number of host array sweeps on
// a periodic array.  Prints "result match" and exits 0 on success.
//
// Build (host compiler only), from the repository root:
//   g++ -O2 -std=c++17 -fopenmp -I include examples/drop_in_7pt.cpp
//       -L bricklib_amd -lbricks -Wl,-rpath,$PWD/bricklib_amd -o drop_in_7pt
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "brick-hip.h"
#include "brickcompare.h"
#include "bricksetup.h"

namespace {

constexpr long GZ = 8, PAD = 8;

// periodic ghost fill of an array with a GZ ghost zone inside PAD padding
void periodic_fill(std::vector<bElem> &a, long n) {
  const long S = n + 2 * (GZ + PAD);
  for (long z = PAD; z < S - PAD; ++z)
    for (long y = PAD; y < S - PAD; ++y)
      for (long x = PAD; x < S - PAD; ++x) {
        auto wrap = [n](long v) { return ((v - PAD - GZ) % n + n) % n + PAD + GZ; };
        a[(z * S + y) * S + x] = a[(wrap(z) * S + wrap(y)) * S + wrap(x)];
      }
}

// the reference's array loop (3axis.cpp:32-37) over the interior
void array_7pt(const std::vector<bElem> &in, std::vector<bElem> &out, long n,
               const double *c) {
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
  const long n = argc > 1 ? std::atol(argv[1]) : 64;  // interior cells per axis
  const int sweeps = argc > 2 ? std::atoi(argv[2]) : 8;
  if (n % 8) {
    std::fprintf(stderr, "n must be a multiple of 8\n");
    return 2;
  }
  const long nb = n / 8, SB = nb + 2, S = n + 2 * (GZ + PAD);

  double coeff[7];
  std::mt19937_64 mt(2024);
  std::uniform_real_distribution<double> u(0, 1);
  double sum = 0;
  for (double &c : coeff) sum += (c = u(mt));
  for (double &c : coeff) c /= sum;

  std::vector<bElem> in((size_t)S * S * S, 0.0), out((size_t)S * S * S, 0.0);
  std::uniform_real_distribution<double> v(-1, 1);
  for (long z = PAD + GZ; z < PAD + GZ + n; ++z)
    for (long y = PAD + GZ; y < PAD + GZ + n; ++y)
      for (long x = PAD + GZ; x < PAD + GZ + n; ++x) in[(z * S + y) * S + x] = v(mt);
  periodic_fill(in, n);

  // bricks: the reference's setup, bIn and bOut interleaved in one storage
  unsigned *grid_ptr = nullptr;
  auto bInfo = init_grid<3>(grid_ptr, {SB, SB, SB});
  auto bStorage = BrickStorage::allocate(bInfo.nbricks, 512 * 2);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bIn(&bInfo, bStorage, 0);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bOut(&bInfo, bStorage, 512);
  copyToBrick<3>({n + 2 * GZ, n + 2 * GZ, n + 2 * GZ}, {PAD, PAD, PAD}, {0, 0, 0}, in.data(),
                 grid_ptr, bIn);

  array_7pt(in, out, n, coeff);

  // device: adjacency, storage, interior list, ghost pairs
  auto dInfo = movBrickInfo(bInfo, gpuMemcpyHostToDevice);
  auto dStorage = movBrickStorage(bStorage, gpuMemcpyHostToDevice);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> dIn(&dInfo, dStorage, 0);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> dOut(&dInfo, dStorage, 512);
  const long nbv[3] = {nb, nb, nb};
  uint32_t nlist = 0, npairs = 0;
  bkCheck(bk_interior_list(nbv, nullptr, &nlist), "bk_interior_list");
  std::vector<unsigned> list(nlist);
  bkCheck(bk_interior_list(nbv, list.data(), &nlist), "bk_interior_list");
  bkCheck(bk_ghost_pairs(nbv, nullptr, &npairs), "bk_ghost_pairs");
  std::vector<unsigned> pairs(2 * (size_t)npairs);
  bkCheck(bk_ghost_pairs(nbv, pairs.data(), &npairs), "bk_ghost_pairs");
  unsigned *d_list = bk_upload_u32(list), *d_pairs = bk_upload_u32(pairs);

  const BrickStencil st = brick_stencil("7pt.py", coeff);
  brick_sweep(st, dOut, dIn, d_list, nlist);
  auto back = movBrickStorage(dStorage, gpuMemcpyDeviceToHost);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> hOut(&bInfo, back, 512);
  const bool one = compareBrick<3>({n, n, n}, {PAD, PAD, PAD}, {GZ, GZ, GZ}, out.data(), grid_ptr,
                                   hOut);
  auto e1 = compareBrickStats<3>({n, n, n}, {PAD, PAD, PAD}, {GZ, GZ, GZ}, out.data(), grid_ptr,
                                 hOut);
  std::printf("1 sweep: compareBrick %s, max abs %.3e, max rel %.3e\n", one ? "ok" : "FAILED",
              e1.max_abs, e1.max_rel);

  // ping-pong sweeps with the on-device periodic ghost refresh, against the
  // host array loop on a periodic array
  std::vector<bElem> a = in, b((size_t)S * S * S, 0.0);
  for (int t = 0; t < sweeps; ++t) {
    array_7pt(a, b, n, coeff);
    periodic_fill(b, n);
    a.swap(b);
  }
  auto dStorage2 = movBrickStorage(bStorage, gpuMemcpyHostToDevice);  // fresh input
  Brick<Dim<8, 8, 8>, Dim<8, 8>> d2In(&dInfo, dStorage2, 0), d2Out(&dInfo, dStorage2, 512);
  brick_sweeps(st, d2In, d2Out, d_list, nlist, d_pairs, npairs, sweeps, /*use_graph=*/true);
  auto back2 = movBrickStorage(dStorage2, gpuMemcpyDeviceToHost);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> res(&bInfo, back2, sweeps % 2 ? 512 : 0);
  auto eN = compareBrickStats<3>({n, n, n}, {PAD, PAD, PAD}, {GZ, GZ, GZ}, a.data(), grid_ptr, res);
  std::printf("%d sweeps: max abs %.3e, max rel %.3e, rms rel %.3e\n", sweeps, eN.max_abs,
              eN.max_rel, eN.rms_rel());

  // the same sweeps with the optional speed knobs: the bricks in 2x2 blocks
  // of one level (a BK_PLAN_QUAD plan over the interior list) and the
  // i-direction taps from DPP lane shifts -- the same bits
  bk_plan plan = nullptr;
  bkCheck(bk_plan_create(&bInfo.adj[0][0], bInfo.nbricks, list.data(), nlist, BK_PLAN_QUAD, 1,
                         &plan),
          "bk_plan_create");
  auto dStorage3 = movBrickStorage(bStorage, gpuMemcpyHostToDevice);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> d3In(&dInfo, dStorage3, 0), d3Out(&dInfo, dStorage3, 512);
  brick_sweeps(st, d3In, d3Out, d_list, nlist, d_pairs, npairs, sweeps, /*use_graph=*/true,
               nullptr, plan, /*lane_shift=*/1);
  auto back3 = movBrickStorage(dStorage3, gpuMemcpyDeviceToHost);
  const bool same = std::memcmp(back3.dat.get(), back2.dat.get(),
                                (size_t)bStorage.chunks * bStorage.step * sizeof(bElem)) == 0;
  std::printf("%d sweeps in 2x2 blocks with lane shifts: %s\n", sweeps,
              same ? "identical" : "DIFFERENT");
  bk_plan_destroy(plan);

  bk_free(d_list);
  bk_free(d_pairs);
  bk_free(dInfo.adj);
  std::free(grid_ptr);
  std::free(bInfo.adj);
  const bool ok = one && e1.max_rel <= 1e-12 && eN.max_rel <= 1e-10 && same;
  std::printf(ok ? "result match\n" : "result mismatch!\n");
  return ok ? 0 : 1;
}
