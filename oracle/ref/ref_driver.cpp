// ref_driver.cpp -- C entry points over CtopCsUtahEdu/bricklib's own headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/ref/Makefile) against the
// reference's include/ directory and src/brickcompare.cpp where they lie,
// into oracle/_ref/libref.so.  The parity tests use it to check the oracle:
// brick numbering and adjacency (init_grid, bricksetup.h:73-89), the
// accessor's element addressing for several vector folds (brick.h:213-395),
// copyToBrick / copyFromBrick (bricksetup.h:172-221), compareBrick
// (brickcompare.h:30-57), and the brick-accessor sweep loops of
// stencils/3axis.cpp:51-66 (7pt) and :123-140 (cond), whose loop bodies are
// restated here because they live inside the reference's benchmark
// functions.  Nothing here is part of the product.
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "brick.h"
#include "bricksetup.h"
#include "brickcompare.h"

namespace {

BrickStorage wrap_storage(double *dat, long chunks, size_t step) {
  BrickStorage s;
  s.dat = std::shared_ptr<bElem>(dat, [](bElem *) {});
  s.chunks = chunks;
  s.step = step;
  return s;
}

struct InfoView {
  BrickInfo<3> info;
  InfoView(unsigned *adj, unsigned nbricks) : info(1) {
    free(info.adj);
    info.adj = (BrickInfo<3>::adjlist)adj;
    info.nbricks = nbricks;
  }
};

template <typename B>
void offsets_of(B &br, unsigned b, long *out, int BK = 8, int BJ = 8, int BI = 8) {
  const bElem *base = br.dat;
  long n = 0;
  for (int k = 0; k < BK; ++k)
    for (int j = 0; j < BJ; ++j)
      for (int i = 0; i < BI; ++i) out[n++] = (long)(&br[b][k][j][i] - base);
}

}  // namespace

extern "C" {

int ref_init_grid(const long dims[3], unsigned *grid_out, unsigned *adj_out) {
  unsigned *grid_ptr = nullptr;
  auto bInfo = init_grid<3>(grid_ptr, {dims[0], dims[1], dims[2]});
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  std::memcpy(grid_out, grid_ptr, n * sizeof(unsigned));
  std::memcpy(adj_out, bInfo.adj, n * 27 * sizeof(unsigned));
  free(grid_ptr);
  free(bInfo.adj);
  return (int)bInfo.nbricks;
}

// Element offsets (relative to brick 0 of the field) of every (k,j,i) of
// brick b, for the 8x8x8 brick in the given fold:
//   0: Dim<8,8> (HIP/AVX-512)   1: Dim<8> (SSE-style)   2: Dim<2,4>
//   3: Dim<4,8> (CUDA)          4: Dim<2,2> (AVX2)
// and for the 4x8x16 brick (k, j, i extents 4, 8, 16):
//   5: Dim<4,16>                6: Dim<8,8>
int ref_fold_offsets(int fold, unsigned *adj, unsigned nbricks, unsigned b,
                     size_t step, long *out) {
  InfoView iv(adj, nbricks);
  std::vector<double> dummy(1);
  auto st = wrap_storage(dummy.data(), nbricks, step);
  switch (fold) {
    case 0: { Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&iv.info, st, 0); offsets_of(br, b, out); break; }
    case 1: { Brick<Dim<8, 8, 8>, Dim<8>> br(&iv.info, st, 0); offsets_of(br, b, out); break; }
    case 2: { Brick<Dim<8, 8, 8>, Dim<2, 4>> br(&iv.info, st, 0); offsets_of(br, b, out); break; }
    case 3: { Brick<Dim<8, 8, 8>, Dim<4, 8>> br(&iv.info, st, 0); offsets_of(br, b, out); break; }
    case 4: { Brick<Dim<8, 8, 8>, Dim<2, 2>> br(&iv.info, st, 0); offsets_of(br, b, out); break; }
    case 5: { Brick<Dim<4, 8, 16>, Dim<4, 16>> br(&iv.info, st, 0); offsets_of(br, b, out, 4, 8, 16); break; }
    case 6: { Brick<Dim<4, 8, 16>, Dim<8, 8>> br(&iv.info, st, 0); offsets_of(br, b, out, 4, 8, 16); break; }
    default: return -1;
  }
  return 0;
}

// Offset of element (k, j, i), each in [-8, 16), of brick b through the
// neighbour redirection of the accessor (Dim<8,8> fold).
long ref_neighbor_offset(unsigned *adj, unsigned nbricks, unsigned b,
                         size_t step, int k, int j, int i) {
  InfoView iv(adj, nbricks);
  std::vector<double> dummy(1);
  auto st = wrap_storage(dummy.data(), nbricks, step);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&iv.info, st, 0);
  return (long)(&br[b][k][j][i] - br.dat);
}

// The same through the 4x8x16 brick: layout 1 = Dim<4,16> fold,
// layout 2 = Dim<8,8> fold; (k, j, i) anywhere in the 3x3x3 neighbourhood.
long ref_neighbor_offset_layout(int lay, unsigned *adj, unsigned nbricks, unsigned b,
                                size_t step, int k, int j, int i) {
  InfoView iv(adj, nbricks);
  std::vector<double> dummy(1);
  auto st = wrap_storage(dummy.data(), nbricks, step);
  if (lay == 1) {
    Brick<Dim<4, 8, 16>, Dim<4, 16>> br(&iv.info, st, 0);
    return (long)(&br[b][k][j][i] - br.dat);
  }
  Brick<Dim<4, 8, 16>, Dim<8, 8>> br(&iv.info, st, 0);
  return (long)(&br[b][k][j][i] - br.dat);
}

void ref_copy_to_brick(const long dims[3], const long pad[3],
                       const long ghost[3], double *arr, unsigned *grid,
                       unsigned *adj, unsigned nbricks, double *dat,
                       size_t step) {
  InfoView iv(adj, nbricks);
  auto st = wrap_storage(dat, nbricks, step);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&iv.info, st, 0);
  copyToBrick<3>({dims[0], dims[1], dims[2]}, {pad[0], pad[1], pad[2]},
                 {ghost[0], ghost[1], ghost[2]}, arr, grid, br);
}

void ref_copy_from_brick(const long dims[3], const long pad[3],
                         const long ghost[3], double *arr, unsigned *grid,
                         unsigned *adj, unsigned nbricks, double *dat,
                         size_t step) {
  InfoView iv(adj, nbricks);
  auto st = wrap_storage(dat, nbricks, step);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&iv.info, st, 0);
  copyFromBrick<3>({dims[0], dims[1], dims[2]}, {pad[0], pad[1], pad[2]},
                   {ghost[0], ghost[1], ghost[2]}, arr, grid, br);
}

int ref_compare_brick(const long dims[3], const long pad[3],
                      const long ghost[3], double *arr, unsigned *grid,
                      unsigned *adj, unsigned nbricks, double *dat,
                      size_t step) {
  InfoView iv(adj, nbricks);
  auto st = wrap_storage(dat, nbricks, step);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&iv.info, st, 0);
  return compareBrick<3>({dims[0], dims[1], dims[2]}, {pad[0], pad[1], pad[2]},
                         {ghost[0], ghost[1], ghost[2]}, arr, grid, br)
             ? 1
             : 0;
}

// stencils/3axis.cpp:51-66 over a brick list, through the reference accessor.
void ref_sweep_7pt(unsigned *adj, unsigned nbricks, const unsigned *list,
                   unsigned nlist, double *in, size_t in_step, double *out,
                   size_t out_step, const double *coeff) {
  InfoView iv(adj, nbricks);
  auto si = wrap_storage(in, nbricks, in_step);
  auto so = wrap_storage(out, nbricks, out_step);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bIn(&iv.info, si, 0);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bOut(&iv.info, so, 0);
#pragma omp parallel for
  for (long l = 0; l < (long)nlist; ++l) {
    unsigned b = list[l];
    for (long k = 0; k < 8; ++k)
      for (long j = 0; j < 8; ++j)
        for (long i = 0; i < 8; ++i)
          bOut[b][k][j][i] = coeff[5] * bIn[b][k + 1][j][i] + coeff[6] * bIn[b][k - 1][j][i] +
                             coeff[3] * bIn[b][k][j + 1][i] + coeff[4] * bIn[b][k][j - 1][i] +
                             coeff[1] * bIn[b][k][j][i + 1] + coeff[2] * bIn[b][k][j][i - 1] +
                             coeff[0] * bIn[b][k][j][i];
  }
}

// stencils/3axis.cpp:123-140 (cond) over a brick list.
void ref_sweep_cond(unsigned *adj, unsigned nbricks, const unsigned *list,
                    unsigned nlist, double *in, size_t in_step, double *out,
                    size_t out_step, const double *coeff) {
  using std::max;
  InfoView iv(adj, nbricks);
  auto si = wrap_storage(in, nbricks, in_step);
  auto so = wrap_storage(out, nbricks, out_step);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bIn(&iv.info, si, 0);
  Brick<Dim<8, 8, 8>, Dim<8, 8>> bOut(&iv.info, so, 0);
#pragma omp parallel for
  for (long l = 0; l < (long)nlist; ++l) {
    unsigned b = list[l];
    for (long k = 0; k < 8; ++k)
      for (long j = 0; j < 8; ++j)
        for (long i = 0; i < 8; ++i) {
          bElem partial =
              coeff[5] * max(bIn[b][k + 1][j][i], 0.0) + coeff[6] * max(bIn[b][k - 1][j][i], 0.0) +
              coeff[3] * max(bIn[b][k][j + 1][i], 0.0) + coeff[4] * max(bIn[b][k][j - 1][i], 0.0) +
              coeff[1] * max(bIn[b][k][j][i + 1], 0.0) + coeff[2] * max(bIn[b][k][j][i - 1], 0.0) +
              coeff[0] * max(bIn[b][k][j][i], 0.0);
          bOut[b][k][j][i] = partial > 0 ? partial : -partial;
        }
  }
}

}  // extern "C"
