// decomp_check.cpp -- host checks of include/brick-decomp.h (BrickDecomp,
// BitSet, the skin lists) against libbricks.so, driven by
// tests/test_cpp_decomp.py.
//
//   decomp_check bitset
//       BitSet algebra on known cases; prints "bitset ok".
//   decomp_check layout <good|normal|bad> <ni> <nj> <nk> <numfield> <B>
//       the numbering of BrickDecomp<3,B,B,B>({Bni, Bnj, Bnk}, B, numfield)
//       for B = 8 or 4 (4: 512-byte bricks, grouped 8 to a 4 KiB page):
//       sep_pos, skin_size, the ghost/skin runs and grid[k][j][i], one
//       record per line for the Python side to compare.
//   decomp_check exchange <good|normal|bad> <ni> <nj> <nk>
//       host exchange on two interleaved fields: every ghost element must
//       equal its periodic image afterwards; prints "exchange ok".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "brick-decomp.h"

namespace {

const std::vector<BitSet> &pick(const char *name) {
  if (!std::strcmp(name, "good")) return skin3d_good;
  if (!std::strcmp(name, "normal")) return skin3d_normal;
  if (!std::strcmp(name, "bad")) return skin3d_bad;
  std::fprintf(stderr, "unknown skin list %s\n", name);
  std::exit(2);
}

int fail(const char *what) {
  std::printf("FAILED: %s\n", what);
  return 1;
}

int bitset() {
  BitSet a = {1, -2, 3};
  BitSet na = !a;
  if (na.set != BitSet({-1, 2, -3}).set) return fail("negation");
  if ((!na).set != a.set) return fail("double negation");
  if (a.size() != 3 || !a.get(-2) || a.get(2)) return fail("size/get");
  if (!(BitSet({1, -3}) <= BitSet({1, 2, -3})) || BitSet({1, 3}) <= BitSet({1, 2, -3}))
    return fail("subset");
  if (!(BitSet({1, 2, -3}) >= BitSet({2}))) return fail("superset");
  if (BitSet({1}).set != 2ull || BitSet({-1}).set != (1ull << 32)) return fail("bit layout");
  if ((BitSet({1, 2}) & BitSet({2, 3})).set != BitSet({2}).set) return fail("intersection");
  if ((BitSet({1}) | BitSet({-3})).set != BitSet({1, -3}).set) return fail("union");
  if ((BitSet({1, 2}) ^ BitSet({2, 3})).set != BitSet({1, 3}).set) return fail("symmetric diff");
  // (! is element negation, not logical not: test emptiness through bool)
  if (static_cast<bool>(BitSet()) || !static_cast<bool>(BitSet({2}))) return fail("emptiness");
  BitSet f;
  f.flip(3).flip(-1).flip(3);
  if (f.set != BitSet({-1}).set) return fail("flip");
  if (skin3d_good.size() != 26 || skin3d_normal.size() != 27 || skin3d_bad.size() != 51)
    return fail("skin list sizes");
  std::printf("bitset ok\n");
  return 0;
}

// signed axes of a BitSet as "i j k" in -1/0/1
void put_dirs(const BitSet &b) {
  for (int a = 1; a <= 3; ++a) std::printf(" %d", b.get(a) ? 1 : (b.get(-a) ? -1 : 0));
}

template <unsigned B>
int layout(const char *name, unsigned ni, unsigned nj, unsigned nk, unsigned numfield) {
  BrickDecomp<3, B, B, B> d({B * ni, B * nj, B * nk}, B, numfield);
  d.initialize(pick(name));
  std::printf("sep %u %u %u\n", d.sep_pos[0], d.sep_pos[1], d.sep_pos[2]);
  std::printf("skin_size");
  for (long s : d.skin_size) std::printf(" %ld", s);
  std::printf("\n");
  if (d.ghost.size() != d.skin.size()) return fail("ghost/skin count");
  for (size_t q = 0; q < d.ghost.size(); ++q) {
    const auto &g = d.ghost[q], &s = d.skin[q];
    if ((!g.neighbor).set != s.neighbor.set) return fail("skin neighbour is !ghost neighbour");
    std::printf("run");
    put_dirs(g.neighbor);
    std::printf(" %u %u %u %u %u %u %u\n", g.pos, s.pos, g.len, g.skin_st, g.skin_ed,
                g.first_pad, g.last_pad);
  }
  BrickInfo<3> info = d.getBrickInfo();
  if (info.nbricks != d.sep_pos[2]) return fail("nbricks");
  std::printf("grid");
  for (unsigned k = 0; k < nk + 2; ++k)
    for (unsigned j = 0; j < nj + 2; ++j)
      for (unsigned i = 0; i < ni + 2; ++i) std::printf(" %u", d[k][j][i]);
  std::printf("\n");
  // the reference's adjacency check (single/mpi.cpp:81-88) over the interior
  for (unsigned k = 1; k <= nk; ++k)
    for (unsigned j = 1; j <= nj; ++j)
      for (unsigned i = 1; i <= ni; ++i) {
        const unsigned l = d[k][j][i];
        for (int s = 0; s < 27; ++s)
          if (info.adj[info.adj[l][s]][26 - s] != l) return fail("adjacency symmetry");
      }
  return 0;
}

int exchange(const char *name, long ni, long nj, long nk) {
  BrickDecomp<3, 8, 8, 8> d({(unsigned)(8 * ni), (unsigned)(8 * nj), (unsigned)(8 * nk)}, 8, 2);
  d.initialize(pick(name));
  BrickInfo<3> info = d.getBrickInfo();
  const size_t step = 1024;  // two interleaved 512-element fields
  BrickStorage st = info.allocate(step);
  bElem *dat = st.dat.get();
  for (size_t e = 0; e < (size_t)info.nbricks * step; ++e) dat[e] = -1.0;
  auto value = [](long gi, long gj, long gk, long e, int field) {
    return (bElem)(((gk * 1000 + gj) * 1000 + gi) * 1024 + e + field * 512);
  };
  const long n[3] = {ni, nj, nk};
  for (long k = 1; k <= nk; ++k)
    for (long j = 1; j <= nj; ++j)
      for (long i = 1; i <= ni; ++i)
        for (int f = 0; f < 2; ++f)
          for (long e = 0; e < 512; ++e)
            dat[(size_t)d[k][j][i] * step + f * 512 + e] = value(i - 1, j - 1, k - 1, e, f);
  d.exchange(st);
  long checked = 0;
  for (long k = 0; k < nk + 2; ++k)
    for (long j = 0; j < nj + 2; ++j)
      for (long i = 0; i < ni + 2; ++i) {
        const long c[3] = {i, j, k};
        long w[3];
        bool ghost = false;
        for (int a = 0; a < 3; ++a) {
          ghost |= c[a] == 0 || c[a] == n[a] + 1;
          w[a] = ((c[a] - 1) % n[a] + n[a]) % n[a];
        }
        if (!ghost) continue;
        const bElem *b = dat + (size_t)d[k][j][i] * step;
        for (int f = 0; f < 2; ++f)
          for (long e = 0; e < 512; ++e)
            if (b[f * 512 + e] != value(w[0], w[1], w[2], e, f)) {
              std::printf("ghost brick (%ld %ld %ld) field %d elem %ld wrong\n", i, j, k, f, e);
              return 1;
            }
        ++checked;
      }
  std::printf("checked %ld ghost bricks\nexchange ok\n", checked);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    if (argc >= 2 && !std::strcmp(argv[1], "bitset")) return bitset();
    if (argc == 8 && !std::strcmp(argv[1], "layout")) {
      auto f = std::atoi(argv[7]) == 4 ? layout<4> : layout<8>;
      return f(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
               std::atoi(argv[6]));
    }
    if (argc == 6 && !std::strcmp(argv[1], "exchange"))
      return exchange(argv[2], std::atol(argv[3]), std::atol(argv[4]), std::atol(argv[5]));
  } catch (const std::exception &e) {
    std::printf("error: %s\n", e.what());
    return 3;
  }
  std::fprintf(stderr, "usage: decomp_check bitset | layout ... | exchange ...\n");
  return 2;
}
