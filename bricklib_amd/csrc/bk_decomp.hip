// bk_decomp.hip -- the BrickDecomp brick numbering (host code).
//
// Restates the layout of BrickDecomp::initialize (include/brick-mpi.h:322-460,
// with _populate :200-230, populate :232-246, get_region_size :248-256,
// _adj_populate :258-281 and allneighbors, src/brick-mpi.cpp:9-23) for one
// subdomain of nb bricks per axis with a ghost zone of g bricks:
//
//   ids from `factor` up (factor = page / brick bytes when bricks are smaller
//   than a page, else 1), each region's end rounded up to a multiple of
//   factor:
//     1. the inner bricks (2g .. nb-1 on every axis, padded coordinates);
//     2. the skin, region by region in the skin list's order (any order of
//        the 26 regions, as BrickDecomp::initialize takes; bk_decomp_layout
//        uses skin3d_good, src/brick-mpi.cpp:25-52), a region padded in
//        front when it shares fewer axes with its predecessor than with its
//        successor;
//     3. per neighbour n (26, allneighbors order), the ghost bricks owned by
//        n, again region by region in skin-list order over the regions that
//        contain the opposite direction !n -- each run of consecutive such
//        regions is one (ghost, skin) pair of equally long contiguous
//        ranges, the unit of the exchange.
//
// Adjacency: out-of-range neighbours are brick 0, as in _adj_populate.
// Regions are encoded as signed-axis sets: bit a (a = 0, 1, 2) for +axis,
// bit 3 + a for -axis (axis 0 = i, contiguous).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bk_api.h"

namespace bk {
int plan_fail(int code, const char *msg);
}

namespace {

using Set = unsigned;  // bit a: +a, bit 3+a: -a

inline bool has(Set s, int a, int sign) { return (s >> (sign > 0 ? a : 3 + a)) & 1u; }
inline Set flip_dirs(Set s) { return ((s & 7u) << 3) | (s >> 3); }
inline bool subset(Set a, Set b) { return (a & b) == a; }
inline int popc(Set s) { return __builtin_popcount(s); }

constexpr Set P(int a) { return 1u << a; }
constexpr Set M(int a) { return 1u << (3 + a); }

// The recommended surface order (skin3d_good): consecutive regions share
// axes so that every neighbour's skin is one or two contiguous runs.
const Set kSkinGood[26] = {
    P(0),               P(0) | M(2),        P(0) | P(1) | M(2), P(0) | P(1),
    P(0) | P(1) | P(2), P(1) | P(2),        P(1),               P(1) | M(2),
    M(0) | P(1) | M(2), M(0) | P(1),        M(0) | P(1) | P(2), M(0) | P(2),
    M(0),               M(2),               M(0) | M(2),        M(0) | M(1) | M(2),
    M(0) | M(1),        M(0) | M(1) | P(2), M(1) | P(2),        M(1),
    M(1) | M(2),        P(0) | M(1) | M(2), P(0) | M(1),        P(0) | M(1) | P(2),
    P(0) | P(2),        P(2)};

constexpr int kMaxSkin = 1024;  // bound on skin-list entries (empty ones included)

// the 27 direction sets in allneighbors order: per axis 1..3 (i, j, k),
// "+", "none", "-", the first axis varying slowest
void all_neighbors(std::vector<Set> &out) {
  for (int a0 = 0; a0 < 3; ++a0)
    for (int a1 = 0; a1 < 3; ++a1)
      for (int a2 = 0; a2 < 3; ++a2) {
        const int c[3] = {a0, a1, a2};
        Set s = 0;
        for (int a = 0; a < 3; ++a) s |= c[a] == 0 ? P(a) : (c[a] == 2 ? M(a) : 0u);
        out.push_back(s);
      }
}

struct Layout {
  long dims[3], g, t[3], stride[3];
  uint32_t factor;
  std::vector<uint32_t> grid;
  uint32_t pos = 0;

  long region_size(Set r) const {
    long n = 1;
    for (int a = 0; a < 3; ++a)
      n *= (has(r, a, 1) || has(r, a, -1)) ? g : dims[a] - 2 * g;
    return n;
  }
  uint32_t pad(Set r) const {
    const long n = region_size(r);
    return (uint32_t)((factor - n % factor) % factor);
  }
  void fill(Set r, long ref, int a) {
    if (a < 0) {
      grid[(size_t)ref] = pos++;
      return;
    }
    long st = -1;
    if (has(r, a, 1)) st = dims[a];
    if (has(r, a, -1)) st = g;
    if (st >= 0)
      for (long q = 0; q < g; ++q) fill(r, ref + (st + q) * stride[a], a - 1);
    else
      for (long q = 2 * g; q < dims[a]; ++q) fill(r, ref + q * stride[a], a - 1);
  }
  // the bricks of region r of the subdomain in direction `owner`
  void populate(Set owner, Set r) {
    long ref = 0;
    for (int a = 0; a < 3; ++a) {
      if (has(owner, a, 1)) ref += dims[a] * stride[a];
      if (has(owner, a, -1)) ref -= dims[a] * stride[a];
    }
    fill(r, ref, 2);
    pos = (pos + factor - 1) / factor * factor;
  }
};

int build(const long nb[3], long g, size_t brick_bytes, size_t page_bytes,
          const std::vector<Set> &skin, std::vector<uint32_t> *grid_out,
          std::vector<uint32_t> *adj_out, uint32_t *nbricks, uint32_t sep[3],
          std::vector<bk_decomp_region> *regions, std::vector<uint32_t> *skin_size) {
  Layout L;
  L.g = g;
  long n = 1;
  for (int a = 0; a < 3; ++a) {
    L.dims[a] = nb[a];
    L.t[a] = nb[a] + 2 * g;
    L.stride[a] = n;
    n *= L.t[a];
  }
  L.factor = (brick_bytes < page_bytes) ? (uint32_t)(page_bytes / brick_bytes) : 1u;
  L.grid.assign((size_t)n, 0u);
  L.pos = L.factor;

  const int NS = (int)skin.size();
  std::vector<bool> pad_first(NS);
  Set last = 0;
  for (int q = 0; q < NS; ++q) {
    const Set next = q + 1 < NS ? skin[q + 1] : 0u;
    pad_first[q] = popc(last & skin[q]) < popc(skin[q] & next);
    last = skin[q];
  }
  L.populate(0, 0);  // inner bricks
  std::vector<uint32_t> st_pos{L.pos};
  sep[0] = L.pos;
  for (int q = 0; q < NS; ++q) {
    const uint32_t before = L.pos;
    if (pad_first[q]) L.pos += L.pad(skin[q]);
    if (skin[q]) L.populate(0, skin[q]);  // an empty entry takes no ids
    st_pos.push_back(L.pos);
    skin_size->push_back(L.pos - before);  // with the region's padding
  }
  sep[1] = L.pos;

  std::vector<Set> nbrs;
  all_neighbors(nbrs);
  for (Set nset : nbrs) {
    if (nset == 0) continue;
    const Set in = flip_dirs(nset);
    int run = -1;
    bk_decomp_region r{};
    bool ok = true;
    auto close = [&](int end) {
      r.skin_ed = (uint32_t)end;
      r.last_pad = pad_first[end - 1] ? 0u : L.pad(skin[end - 1]);
      r.len = L.pos - r.ghost_pos;
      // the ghost run and its skin run are laid out alike: equal lengths
      ok = ok && st_pos[end] - r.skin_pos == r.len;
      if (regions) regions->push_back(r);
    };
    for (int q = 0; q < NS; ++q) {
      if (subset(in, skin[q])) {
        if (run < 0) {
          run = q;
          for (int a = 0; a < 3; ++a)
            r.neighbor[a] = has(nset, a, 1) ? 1 : (has(nset, a, -1) ? -1 : 0);
          r.skin_st = (uint32_t)q;
          r.first_pad = pad_first[q] ? L.pad(skin[q]) : 0u;
          r.ghost_pos = L.pos;
          r.skin_pos = st_pos[q];
        }
        if (pad_first[q]) L.pos += L.pad(skin[q]);
        L.populate(nset, skin[q]);
      } else if (run >= 0) {
        close(q);
        run = -1;
      }
    }
    if (run >= 0) close(NS);
    if (!ok)
      return bk::plan_fail(BK_ERR_UNSUPPORTED, "bk_decomp_layout: ghost and skin runs differ");
  }
  sep[2] = L.pos;
  *nbricks = L.pos;

  if (adj_out) {
    adj_out->assign((size_t)L.pos * 27, 0u);
    for (long k = 0; k < L.t[2]; ++k)
      for (long j = 0; j < L.t[1]; ++j)
        for (long i = 0; i < L.t[0]; ++i) {
          const long ref = i + L.stride[1] * j + L.stride[2] * k;
          uint32_t *row = adj_out->data() + (size_t)L.grid[(size_t)ref] * 27;
          for (int kd = 0; kd < 3; ++kd)
            for (int jd = 0; jd < 3; ++jd)
              for (int id = 0; id < 3; ++id) {
                const long ii = i + id - 1, jj = j + jd - 1, kk = k + kd - 1;
                const bool in_range = ii >= 0 && ii < L.t[0] && jj >= 0 && jj < L.t[1] &&
                                      kk >= 0 && kk < L.t[2];
                row[kd * 9 + jd * 3 + id] =
                    in_range ? L.grid[(size_t)(ii + L.stride[1] * jj + L.stride[2] * kk)] : 0u;
              }
        }
  }
  if (grid_out) grid_out->swap(L.grid);
  return BK_OK;
}

}  // namespace

extern "C" {

int bk_decomp_layout_skin(const long nb[3], int ghost_bricks, size_t brick_bytes,
                          size_t page_bytes, const int32_t *h_skin, int nskin, uint32_t *h_grid,
                          uint32_t *h_adj, uint32_t *nbricks, uint32_t sep_pos[3],
                          bk_decomp_region *h_regions, uint32_t *nregions,
                          uint32_t *h_skin_size) {
  if (!nb || !nbricks || !nregions || !sep_pos || !h_skin)
    return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: null pointer");
  if (ghost_bricks < 1) return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: ghost_bricks >= 1");
  for (int a = 0; a < 3; ++a)
    if (nb[a] < 2L * ghost_bricks)
      return bk::plan_fail(BK_ERR_ARG,
                           "bk_decomp_layout: every axis needs at least 2 * ghost_bricks bricks");
  if (brick_bytes == 0 || page_bytes == 0 ||
      (std::max(brick_bytes, page_bytes) % std::min(brick_bytes, page_bytes)) != 0)
    return bk::plan_fail(BK_ERR_ARG,
                         "bk_decomp_layout: brick size must be a factor or multiple of the page");
  const double cells = (double)(nb[0] + 2 * ghost_bricks) * (nb[1] + 2 * ghost_bricks) *
                       (nb[2] + 2 * ghost_bricks);
  if (cells > 1.0e9) return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: grid too large");
  if (nskin < 26 || nskin > kMaxSkin)
    return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: the skin list needs 26 to 1024 entries");
  // the surface order: the 26 non-empty direction sets once each, empty
  // entries anywhere (they take no ids but break runs and move padding)
  std::vector<Set> skin(nskin);
  uint64_t seen = 0;  // bit r: direction set r already listed
  int nonempty = 0;
  for (int q = 0; q < nskin; ++q) {
    Set r = 0;
    for (int a = 0; a < 3; ++a) {
      const int32_t v = h_skin[3 * q + a];
      if (v == 1) r |= P(a);
      else if (v == -1) r |= M(a);
      else if (v != 0)
        return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: skin components must be -1, 0 or 1");
    }
    if (r != 0) {
      if (seen >> r & 1u)
        return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: a skin region is listed twice");
      seen |= uint64_t(1) << r;
      ++nonempty;
    }
    skin[q] = r;
  }
  if (nonempty != 26)
    return bk::plan_fail(BK_ERR_ARG, "bk_decomp_layout: the skin list must name all 26 regions");
  std::vector<uint32_t> grid, adj, sizes;
  std::vector<bk_decomp_region> regions;
  uint32_t nbr = 0, sep[3];
  int rc = build(nb, ghost_bricks, brick_bytes, page_bytes, skin, h_grid ? &grid : nullptr,
                 h_adj ? &adj : nullptr, &nbr, sep, &regions, &sizes);
  if (rc != BK_OK) return rc;
  if (h_grid) std::memcpy(h_grid, grid.data(), grid.size() * sizeof(uint32_t));
  if (h_adj) std::memcpy(h_adj, adj.data(), adj.size() * sizeof(uint32_t));
  if (h_regions) std::memcpy(h_regions, regions.data(), regions.size() * sizeof(regions[0]));
  if (h_skin_size) std::memcpy(h_skin_size, sizes.data(), sizes.size() * sizeof(uint32_t));
  *nbricks = nbr;
  *nregions = (uint32_t)regions.size();
  for (int q = 0; q < 3; ++q) sep_pos[q] = sep[q];
  return BK_OK;
}

int bk_decomp_layout(const long nb[3], int ghost_bricks, size_t brick_bytes, size_t page_bytes,
                     uint32_t *h_grid, uint32_t *h_adj, uint32_t *nbricks, uint32_t sep_pos[3],
                     bk_decomp_region *h_regions, uint32_t *nregions) {
  int32_t skin[26 * 3];
  for (int q = 0; q < 26; ++q)
    for (int a = 0; a < 3; ++a)
      skin[3 * q + a] = has(kSkinGood[q], a, 1) ? 1 : (has(kSkinGood[q], a, -1) ? -1 : 0);
  return bk_decomp_layout_skin(nb, ghost_bricks, brick_bytes, page_bytes, skin, 26, h_grid,
                               h_adj, nbricks, sep_pos, h_regions, nregions, nullptr);
}

}  // extern "C"
