// header_check.cpp -- host-only checks of the C++ brick headers.
//
// Prints, for the CPU tests to compare with the oracle and oracle/_ref:
//   grid <d0> <d1> <d2> <periodic>: the grid numbering and adjacency;
//   fold <id>: element offsets of every (k,j,i) of brick 13 of a 3x3x3 grid
//              in fold 0: Dim<8,8>  1: Dim<8>  2: Dim<2,4>  3: Dim<4,8>  4: Dim<2,2>,
//              and of the 4x8x16 brick in fold 5: Dim<4,16>  6: Dim<8,8>;
//   neighbor: offsets of (k,j,i) in [-8,16)^3 on a stride of 5, brick 13;
//   copy: copyToBrick then copyFromBrick round trip and compareBrick.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "brickcompare.h"
#include "bricksetup.h"

template <typename B>
void print_offsets(B &br, unsigned b, int BK = 8, int BJ = 8, int BI = 8) {
  for (int k = 0; k < BK; ++k)
    for (int j = 0; j < BJ; ++j)
      for (int i = 0; i < BI; ++i) std::printf("%ld\n", (long)(&br[b][k][j][i] - br.dat));
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  if (!std::strcmp(argv[1], "grid")) {
    const long d0 = std::atol(argv[2]), d1 = std::atol(argv[3]), d2 = std::atol(argv[4]);
    unsigned *g = nullptr;
    auto info = std::atoi(argv[5]) ? init_grid_periodic<3>(g, {d0, d1, d2})
                                   : init_grid<3>(g, {d0, d1, d2});
    for (long p = 0; p < d0 * d1 * d2; ++p) std::printf("%u\n", g[p]);
    for (long p = 0; p < d0 * d1 * d2; ++p)
      for (int s = 0; s < 27; ++s) std::printf("%u\n", info.adj[p][s]);
    return 0;
  }
  unsigned *g = nullptr;
  auto info = init_grid<3>(g, {3, 3, 3});
  auto st = BrickStorage::allocate(info.nbricks, 512);
  if (!std::strcmp(argv[1], "fold")) {
    switch (std::atoi(argv[2])) {
      case 0: { Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&info, st, 0); print_offsets(br, 13); break; }
      case 1: { Brick<Dim<8, 8, 8>, Dim<8>> br(&info, st, 0); print_offsets(br, 13); break; }
      case 2: { Brick<Dim<8, 8, 8>, Dim<2, 4>> br(&info, st, 0); print_offsets(br, 13); break; }
      case 3: { Brick<Dim<8, 8, 8>, Dim<4, 8>> br(&info, st, 0); print_offsets(br, 13); break; }
      case 4: { Brick<Dim<8, 8, 8>, Dim<2, 2>> br(&info, st, 0); print_offsets(br, 13); break; }
      case 5: { Brick<Dim<4, 8, 16>, Dim<4, 16>> br(&info, st, 0); print_offsets(br, 13, 4, 8, 16); break; }
      case 6: { Brick<Dim<4, 8, 16>, Dim<8, 8>> br(&info, st, 0); print_offsets(br, 13, 4, 8, 16); break; }
      default: return 2;
    }
    return 0;
  }
  if (!std::strcmp(argv[1], "neighbor")) {
    Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&info, st, 0);
    for (int k = -8; k < 16; k += 5)
      for (int j = -8; j < 16; j += 5)
        for (int i = -8; i < 16; i += 5)
          std::printf("%d %d %d %ld\n", k, j, i, (long)(&br[13][k][j][i] - br.dat));
    std::printf("nb %ld\n", (long)(br.neighbor<2, 1, 0>(13) - br.dat));
    return 0;
  }
  if (!std::strcmp(argv[1], "copy")) {
    // 16x24x8 region, padding 8, ghost 8 -> bricks (16+16)/8 x (24+16)/8 x (8+16)/8
    unsigned *g2 = nullptr;
    auto info2 = init_grid<3>(g2, {4, 5, 3});
    auto st2 = BrickStorage::allocate(info2.nbricks, 512);
    Brick<Dim<8, 8, 8>, Dim<8, 8>> br(&info2, st2, 0);
    const long e0 = 16 + 32, e1 = 24 + 32, e2 = 8 + 32;
    std::vector<bElem> arr(e0 * e1 * e2), back(e0 * e1 * e2, 0.0);
    for (size_t q = 0; q < arr.size(); ++q) arr[q] = (double)((q * 2654435761u) % 1000) / 7.0;
    copyToBrick<3>({16, 24, 8}, {8, 8, 8}, {8, 8, 8}, arr.data(), g2, br);
    copyFromBrick<3>({16, 24, 8}, {8, 8, 8}, {8, 8, 8}, back.data(), g2, br);
    long diff = 0, touched = 0;
    for (long z = 16; z < 16 + 8; ++z)  // the region: padding + ghost .. + dims
      for (long y = 16; y < 16 + 24; ++y)
        for (long x = 16; x < 16 + 16; ++x) {
          const size_t q = (z * e1 + y) * e0 + x;
          diff += back[q] != arr[q];
          ++touched;
        }
    const bool same = compareBrick<3>({16, 24, 8}, {8, 8, 8}, {8, 8, 8}, arr.data(), g2, br);
    arr[(20 * e1 + 30) * e0 + 25] += 1.0;  // a planted difference inside the region
    const bool caught = !compareBrick<3>({16, 24, 8}, {8, 8, 8}, {8, 8, 8}, arr.data(), g2, br);
    std::printf("touched %ld diff %ld same %d caught %d\n", touched, diff, (int)same, (int)caught);
    return 0;
  }
  return 2;
}
