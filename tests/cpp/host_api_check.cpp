// host_api_check.cpp -- the host-only C-ABI calls of libbricks.so (grid
// numbering, interior and tiled lists, ghost pairs and images, execution
// plans of every shape and length over grid-ordered and shuffled lists,
// BrickDecomp layouts with empty skin entries, argument errors), built and
// run against a library whose host code carries AddressSanitizer and
// UndefinedBehaviorSanitizer checks.  No GPU is needed.  Prints "host api
// ok" and exits 0 when every call succeeds and every plan covers its list.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include <algorithm>
#include "bk_api.h"
#define CK(x) do { int rc_ = (x); if (rc_) { std::printf("FAIL %s -> %d: %s\n", #x, rc_, bk_last_error()); std::exit(1); } } while (0)
int main() {
  std::mt19937 rng(7);
  long dims_list[][3] = {{1,1,1},{2,3,4},{5,5,5},{3,1,2},{8,4,6},{1,7,1},{16,16,16}};
  for (auto &nb : dims_list) {
    for (int ghost = 0; ghost <= 1; ++ghost) {
      long padded[3] = {nb[0] + 2*ghost, nb[1] + 2*ghost, nb[2] + 2*ghost};
      const uint32_t N = (uint32_t)(padded[0]*padded[1]*padded[2]);
      std::vector<uint32_t> grid(N), adj((size_t)N*27);
      CK(bk_init_grid(padded, ghost ? 0 : 1, grid.data(), adj.data()));
      uint32_t n = 0;
      if (ghost) {
        CK(bk_interior_list(nb, nullptr, &n));
        std::vector<uint32_t> list(n); CK(bk_interior_list(nb, list.data(), &n));
        uint32_t np = 0; CK(bk_ghost_pairs(nb, nullptr, &np));
        std::vector<uint32_t> pairs(2*(size_t)np); CK(bk_ghost_pairs(nb, pairs.data(), &np));
        uint32_t ni = 0; CK(bk_ghost_images(nb, nullptr, nullptr, &ni));
        std::vector<uint32_t> off(N + 1), img(std::max(1u, ni));
        CK(bk_ghost_images(nb, off.data(), img.data(), &ni));
      }
      for (uint32_t ti : {0u, 1u, 3u, 32u}) for (uint32_t tj : {0u, 2u, 16u}) {
        CK(bk_tiled_list(nb, ghost, ti, tj, nullptr, &n));
        std::vector<uint32_t> list(n); CK(bk_tiled_list(nb, ghost, ti, tj, list.data(), &n));
        for (int shuffle = 0; shuffle < 2; ++shuffle) {
          if (shuffle) std::shuffle(list.begin(), list.end(), rng);
          for (int shape : {BK_PLAN_COLUMN, BK_PLAN_QUAD}) for (uint32_t kcol : {1u, 2u, 3u, 8u}) {
            uint32_t nu = 0, nc = 0;
            int rc = bk_plan_build_host(adj.data(), N, list.data(), n, shape, kcol, nullptr, nullptr, nullptr, nullptr, &nu, &nc);
            if (rc == BK_ERR_UNSUPPORTED) continue;  // quads on a non-periodic outer layer
            CK(rc);
            std::vector<uint32_t> an(nu), le(nu), mk((size_t)nu*kcol), fl(nc);
            CK(bk_plan_build_host(adj.data(), N, list.data(), n, shape, kcol, an.data(), le.data(), mk.data(), fl.data(), &nu, &nc));
            if (nc != n) { std::printf("FAIL covered %u != %u\n", nc, n); return 1; }
          }
        }
      }
    }
    if (nb[0] >= 2 && nb[1] >= 2 && nb[2] >= 2)
    for (size_t page : {(size_t)4096, (size_t)16384, (size_t)65536}) {
      uint32_t nbr = 0, nreg = 0, sep[3];
      CK(bk_decomp_layout(nb, 1, 4096, page, nullptr, nullptr, &nbr, sep, nullptr, &nreg));
      long S[3] = {nb[0]+2, nb[1]+2, nb[2]+2};
      std::vector<uint32_t> g((size_t)S[0]*S[1]*S[2]), a((size_t)nbr*27);
      std::vector<bk_decomp_region> regs(nreg);
      CK(bk_decomp_layout(nb, 1, 4096, page, g.data(), a.data(), &nbr, sep, regs.data(), &nreg));
      // a skin list with empty entries mixed in
      std::vector<int32_t> skin;
      for (int k = -1; k <= 1; ++k) for (int j = -1; j <= 1; ++j) for (int i = -1; i <= 1; ++i) {
        if (!i && !j && !k) continue;
        skin.insert(skin.end(), {i, j, k});
        if ((i + 2*j + 3*k) % 5 == 0) skin.insert(skin.end(), {0, 0, 0});
      }
      const int ns = (int)skin.size() / 3;
      std::vector<uint32_t> ss(ns);
      CK(bk_decomp_layout_skin(nb, 1, 4096, page, skin.data(), ns, nullptr, nullptr, &nbr, sep, nullptr, &nreg, nullptr));
      std::vector<uint32_t> g2((size_t)S[0]*S[1]*S[2]), a2((size_t)nbr*27);
      std::vector<bk_decomp_region> r2(nreg);
      CK(bk_decomp_layout_skin(nb, 1, 4096, page, skin.data(), ns, g2.data(), a2.data(), &nbr, sep, r2.data(), &nreg, ss.data()));
    }
  }
  // argument errors
  long bad[3] = {0, 1, 1};
  uint32_t n = 0;
  if (bk_tiled_list(bad, 1, 32, 16, nullptr, &n) == 0) { std::printf("FAIL bad dims accepted\n"); return 1; }
  if (bk_plan_build_host(nullptr, 0, nullptr, 0, BK_PLAN_QUAD, 1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr) == 0) return 1;
  int ext[3];
  for (int l = 0; l < 4; ++l) (void)bk_layout_extents(l, ext);
  for (int st = -1; st < 7; ++st) { (void)bk_stencil_ncoeff(st); (void)bk_stencil_radius(st); }
  std::printf("host api ok\n");
  return 0;
}
