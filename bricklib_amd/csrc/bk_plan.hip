// bk_plan.hip -- execution plans: a schedule for sweeping a set of bricks.
//
// A plan groups the bricks of a list into units a single wave can march
// through along k:
//   BK_PLAN_COLUMN  unit = up to kcol bricks b, adj[b][22], adj[adj[b][22]][22] ...
//   BK_PLAN_QUAD    unit = up to kcol levels of a 2x2 block of columns:
//                   members m0 = anchor, m1 = adj[m0][14] (i+1),
//                   m2 = adj[m0][16] (j+1), m3 = adj[m1][16]; the next level's
//                   anchor is adj[anchor][22] (k+1).
// Each (unit, level) carries a mask of the members to store; a member not
// in the list, or already covered, is computed but not stored.  Every
// listed brick is stored exactly once.  Units are emitted in the order of
// their first brick in the list, so a grid-ordered list gives spatially
// adjacent units adjacent indices (and so the same XCD).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "bk_api.h"
#include "bk_internal.h"

namespace bk {
int plan_fail(int code, const char *msg);
int plan_build(const uint32_t *h_adj, uint32_t nbricks, const uint32_t *h_list, uint32_t nlist,
               int shape, uint32_t kcol, uint32_t *h_anchor, uint32_t *h_len, uint32_t *h_mask,
               uint32_t *h_flat, uint32_t *h_start, uint32_t *nunits, uint32_t *ncovered);
void graphs_forget_plan(uint64_t plan_id);
}

struct bk_plan_s {
  int shape = 1;
  uint32_t kcol = 1;
  uint32_t nunits = 0;
  uint32_t ncovered = 0;
  uint32_t *d_anchor = nullptr;  // [nunits] level-0 anchor
  uint32_t *d_len = nullptr;     // [nunits] levels in the unit
  uint32_t *d_mask = nullptr;    // [nunits * kcol] member store mask per level
  uint32_t *d_flat = nullptr;    // [ncovered] covered bricks in plan order
  uint32_t *d_start = nullptr;   // [nunits] offset of each unit's bricks in d_flat
  uint32_t nbricks = 0;          // bricks of the adjacency the plan was built on
  int device = 0;
  uint64_t id = 0;               // unique per plan for the process's lifetime (never 0)
};

namespace {
std::atomic<uint64_t> g_next_plan_id{1};
}

namespace bk {

const PlanView plan_view(const bk_plan_s *p) {
  PlanView v;
  v.shape = p->shape;
  v.kcol = p->kcol;
  v.nunits = p->nunits;
  v.ncovered = p->ncovered;
  v.anchor = p->d_anchor;
  v.len = p->d_len;
  v.mask = p->d_mask;
  v.flat = p->d_flat;
  v.start = p->d_start;
  v.nbricks = p->nbricks;
  v.id = p->id;
  return v;
}

}  // namespace bk

namespace {

struct Unit {
  uint32_t anchor;
  std::vector<uint32_t> masks;
};

// Build k-chains of "cells" (a cell = one brick for columns, one quad level
// for quads) keyed by their anchor brick, then cut them into units.
void chain_cells(const uint32_t *adj, const std::vector<uint32_t> &cell_anchor,
                 const std::vector<uint32_t> &cell_mask, uint32_t kcol,
                 std::vector<Unit> &units) {
  std::unordered_map<uint32_t, uint32_t> by_anchor;
  by_anchor.reserve(cell_anchor.size() * 2);
  for (uint32_t c = 0; c < cell_anchor.size(); ++c) by_anchor.emplace(cell_anchor[c], c);
  std::vector<char> used(cell_anchor.size(), 0);
  for (uint32_t c = 0; c < cell_anchor.size(); ++c) {
    if (used[c]) continue;
    // walk up from c while the k+1 anchor is an unused cell
    uint32_t cur = c;
    Unit u;
    u.anchor = cell_anchor[cur];
    for (;;) {
      used[cur] = 1;
      u.masks.push_back(cell_mask[cur]);
      if (u.masks.size() == kcol) {
        units.push_back(u);
        u.masks.clear();
      }
      const uint32_t up = adj[(size_t)cell_anchor[cur] * 27 + 22];
      auto it = by_anchor.find(up);
      if (it == by_anchor.end() || used[it->second]) break;
      // the kernels hand a level's top planes to the next level as its k-1
      // face: link only where the adjacency agrees both ways (it does not
      // on the outer layer of a non-periodic grid, whose out-of-range
      // entries stand for no brick)
      if (adj[(size_t)up * 27 + 4] != cell_anchor[cur]) break;
      cur = it->second;
      if (u.masks.empty()) u.anchor = cell_anchor[cur];
    }
    if (!u.masks.empty()) units.push_back(u);
  }
}

// The quad kernel derives every input brick of a unit from its level-0
// anchor: members m1 = adj[m0][i+1], m2 = adj[m0][j+1], m3 = adj[m1][j+1];
// a 4x4 in-plane table of columns (entry (r, c) from m0 for r, c < 3, m1
// for column 3, m2 for row 3, m3 for the corner), level 0 from the members'
// in-plane neighbours, level -1 through adj[.][k-1], higher levels walked
// up through adj[.][k+1].  True when every stored member's 27 neighbours at
// every level are what those tables give -- always on a periodic grid and
// on the interior of a padded one, not on the outer layer of a
// non-periodic grid.
bool quad_unit_consistent(const uint32_t *adj, const Unit &u) {
  auto A = [adj](uint32_t b, int s) { return adj[(size_t)b * 27 + s]; };
  const uint32_t m0 = u.anchor, m1 = A(m0, 14), m2 = A(m0, 16), m3 = A(m1, 16);
  const int L = (int)u.masks.size();
  std::vector<std::array<uint32_t, 16>> T((size_t)L + 2);  // T[l + 1]: level l
  for (int e = 0; e < 16; ++e) {
    const int r = e >> 2, c = e & 3;
    const uint32_t m = r < 3 ? (c < 3 ? m0 : m1) : (c < 3 ? m2 : m3);
    const uint32_t n0 = A(m, 9 + 3 * (r < 3 ? r : 2) + (c < 3 ? c : 2));
    T[1][e] = n0;
    T[0][e] = A(n0, 4);
  }
  for (int l = 1; l <= L; ++l)
    for (int e = 0; e < 16; ++e) T[(size_t)l + 1][e] = A(T[l][e], 22);
  uint32_t om[4] = {m0, m1, m2, m3};
  for (int l = 0; l < L; ++l) {
    if (l > 0)
      for (auto &b : om) b = A(b, 22);
    for (int mm = 0; mm < 4; ++mm) {
      if (!((u.masks[l] >> mm) & 1u)) continue;
      for (int kd = 0; kd < 3; ++kd)
        for (int jd = 0; jd < 3; ++jd)
          for (int id = 0; id < 3; ++id)
            if (A(om[mm], kd * 9 + jd * 3 + id) !=
                T[(size_t)l + kd][((mm >> 1) + jd) * 4 + (mm & 1) + id])
              return false;
    }
  }
  return true;
}

}  // namespace

extern "C" {

int bk_plan_build_host(const uint32_t *h_adj, uint32_t nbricks, const uint32_t *h_list,
                       uint32_t nlist, int shape, uint32_t kcol, uint32_t *h_anchor,
                       uint32_t *h_len, uint32_t *h_mask, uint32_t *h_flat, uint32_t *nunits,
                       uint32_t *ncovered) {
  return bk::plan_build(h_adj, nbricks, h_list, nlist, shape, kcol, h_anchor, h_len, h_mask,
                        h_flat, nullptr, nunits, ncovered);
}

}  // extern "C"

namespace bk {
int plan_build(const uint32_t *h_adj, uint32_t nbricks, const uint32_t *h_list, uint32_t nlist,
               int shape, uint32_t kcol, uint32_t *h_anchor, uint32_t *h_len, uint32_t *h_mask,
               uint32_t *h_flat, uint32_t *h_start, uint32_t *nunits, uint32_t *ncovered) {
  if (!nunits || !ncovered) return bk::plan_fail(BK_ERR_ARG, "bk_plan_build_host: null count pointer");
  if (!h_adj || (nlist && !h_list))
    return bk::plan_fail(BK_ERR_ARG, "bk_plan_build_host: null pointer");
  if (shape != BK_PLAN_COLUMN && shape != BK_PLAN_QUAD)
    return bk::plan_fail(BK_ERR_ARG, "bk_plan: shape must be BK_PLAN_COLUMN or BK_PLAN_QUAD");
  if (kcol == 0 || kcol > 4096) return bk::plan_fail(BK_ERR_ARG, "bk_plan: kcol in [1, 4096]");
  for (uint32_t l = 0; l < nlist; ++l)
    if (h_list[l] >= nbricks) return bk::plan_fail(BK_ERR_ARG, "bk_plan: list entry out of range");
  for (size_t q = 0; q < (size_t)nbricks * 27; ++q)
    if (h_adj[q] >= nbricks) return bk::plan_fail(BK_ERR_ARG, "bk_plan: adjacency entry out of range");

  std::vector<uint8_t> listed(nbricks, 0), covered(nbricks, 0);
  for (uint32_t l = 0; l < nlist; ++l) listed[h_list[l]] = 1;

  std::vector<uint32_t> cell_anchor, cell_mask;
  for (uint32_t l = 0; l < nlist; ++l) {
    const uint32_t b = h_list[l];
    if (covered[b]) continue;
    if (shape == BK_PLAN_COLUMN) {
      covered[b] = 1;
      cell_anchor.push_back(b);
      cell_mask.push_back(1u);
    } else {
      const uint32_t m0 = b, m1 = h_adj[(size_t)m0 * 27 + 14], m2 = h_adj[(size_t)m0 * 27 + 16],
                     m3 = h_adj[(size_t)m1 * 27 + 16];
      const uint32_t mem[4] = {m0, m1, m2, m3};
      uint32_t mask = 0;
      for (int q = 0; q < 4; ++q) {
        bool dup = false;
        for (int p = 0; p < q; ++p) dup = dup || mem[p] == mem[q];
        if (!dup && listed[mem[q]] && !covered[mem[q]]) {
          covered[mem[q]] = 1;
          mask |= 1u << q;
        }
      }
      cell_anchor.push_back(m0);
      cell_mask.push_back(mask);
    }
  }
  std::vector<Unit> units;
  chain_cells(h_adj, cell_anchor, cell_mask, kcol, units);
  if (shape == BK_PLAN_QUAD)
    for (const Unit &u : units)
      if (!quad_unit_consistent(h_adj, u))
        return bk::plan_fail(BK_ERR_UNSUPPORTED,
                             "bk_plan: a quad of the list has neighbours the adjacency does not "
                             "place consistently (the outer layer of a non-periodic grid); use "
                             "a column plan or no plan for such lists");
  // emit units in the list order of their anchors (for a grid-ordered list:
  // k-segment by k-segment, so units that run together are neighbours)
  {
    std::unordered_map<uint32_t, uint32_t> pos;
    pos.reserve((size_t)nlist * 2);
    for (uint32_t l = 0; l < nlist; ++l) pos.emplace(h_list[l], l);
    std::vector<std::pair<uint32_t, uint32_t>> key(units.size());
    for (uint32_t u = 0; u < units.size(); ++u) {
      auto it = pos.find(units[u].anchor);
      key[u] = {it == pos.end() ? UINT32_MAX : it->second, u};
    }
    std::stable_sort(key.begin(), key.end());
    std::vector<Unit> sorted;
    sorted.reserve(units.size());
    for (auto &k : key) sorted.push_back(std::move(units[k.second]));
    units.swap(sorted);
  }

  uint32_t nf = 0;
  for (size_t u = 0; u < units.size(); ++u) {
    if (h_start) h_start[u] = nf;
    if (h_anchor) h_anchor[u] = units[u].anchor;
    if (h_len) h_len[u] = (uint32_t)units[u].masks.size();
    uint32_t a = units[u].anchor;
    for (size_t l = 0; l < kcol; ++l) {
      const uint32_t m = l < units[u].masks.size() ? units[u].masks[l] : 0u;
      if (h_mask) h_mask[u * kcol + l] = m;
      if (l >= units[u].masks.size()) continue;
      if (shape == BK_PLAN_COLUMN) {
        if (m & 1u) {
          if (h_flat) h_flat[nf] = a;
          ++nf;
        }
      } else {
        const uint32_t m1 = h_adj[(size_t)a * 27 + 14];
        const uint32_t mem[4] = {a, m1, h_adj[(size_t)a * 27 + 16], h_adj[(size_t)m1 * 27 + 16]};
        for (int q = 0; q < 4; ++q)
          if (m & (1u << q)) {
            if (h_flat) h_flat[nf] = mem[q];
            ++nf;
          }
      }
      a = h_adj[(size_t)a * 27 + 22];
    }
  }
  *nunits = (uint32_t)units.size();
  *ncovered = nf;
  return BK_OK;
}
}  // namespace bk

extern "C" {

int bk_plan_create(const uint32_t *h_adj, uint32_t nbricks, const uint32_t *h_list,
                   uint32_t nlist, int shape, uint32_t kcol, bk_plan *out) {
  if (!out) return bk::plan_fail(BK_ERR_ARG, "bk_plan_create: null out pointer");
  *out = nullptr;
  uint32_t nu = 0, nc = 0;
  // sizes first (nunits <= nlist), then the arrays
  int rc = bk::plan_build(h_adj, nbricks, h_list, nlist, shape, kcol, nullptr, nullptr, nullptr,
                          nullptr, nullptr, &nu, &nc);
  if (rc != BK_OK) return rc;
  std::vector<uint32_t> anchor(nu), len(nu), mask((size_t)nu * kcol), flat(nc), start(nu);
  rc = bk::plan_build(h_adj, nbricks, h_list, nlist, shape, kcol, anchor.data(), len.data(),
                      mask.data(), flat.data(), start.data(), &nu, &nc);
  if (rc != BK_OK) return rc;

  bk_plan_s *p = new bk_plan_s();
  p->shape = shape;
  p->kcol = kcol;
  p->nunits = nu;
  p->ncovered = nc;
  p->nbricks = nbricks;
  p->id = g_next_plan_id.fetch_add(1);
  (void)hipGetDevice(&p->device);  // the current device owns the arrays below
  auto up = [](uint32_t **dst, const std::vector<uint32_t> &v) -> hipError_t {
    if (v.empty()) return hipSuccess;
    hipError_t e = hipMalloc(dst, v.size() * sizeof(uint32_t));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  };
  hipError_t e = up(&p->d_anchor, anchor);
  if (e == hipSuccess) e = up(&p->d_len, len);
  if (e == hipSuccess) e = up(&p->d_mask, mask);
  if (e == hipSuccess) e = up(&p->d_flat, flat);
  if (e == hipSuccess) e = up(&p->d_start, start);
  if (e != hipSuccess) {
    bk_plan_destroy(p);
    std::string m = std::string("bk_plan_create: ") + hipGetErrorString(e);
    return bk::plan_fail(e == hipErrorOutOfMemory ? BK_ERR_NOMEM : BK_ERR_HIP, m.c_str());
  }
  *out = p;
  return BK_OK;
}

int bk_plan_destroy(bk_plan p) {
  if (!p) return BK_OK;
  // cached graphs captured the plan's device arrays: drop them first
  bk::graphs_forget_plan(p->id);
  (void)hipFree(p->d_anchor);
  (void)hipFree(p->d_len);
  (void)hipFree(p->d_mask);
  (void)hipFree(p->d_flat);
  (void)hipFree(p->d_start);
  delete p;
  return BK_OK;
}

int bk_plan_query(bk_plan p, uint32_t *nunits, uint32_t *kcol, int *shape,
                  uint32_t *ncovered) {
  if (!p) return bk::plan_fail(BK_ERR_ARG, "bk_plan_query: null plan");
  if (nunits) *nunits = p->nunits;
  if (kcol) *kcol = p->kcol;
  if (shape) *shape = p->shape;
  if (ncovered) *ncovered = p->ncovered;
  return BK_OK;
}

int bk_plan_download(bk_plan p, uint32_t *h_anchor, uint32_t *h_len, uint32_t *h_mask,
                     uint32_t *h_flat) {
  if (!p) return bk::plan_fail(BK_ERR_ARG, "bk_plan_download: null plan");
  auto down = [](uint32_t *dst, const uint32_t *src, size_t n) -> hipError_t {
    if (!dst || !src || n == 0) return hipSuccess;
    return hipMemcpy(dst, src, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
  };
  hipError_t e = down(h_anchor, p->d_anchor, p->nunits);
  if (e == hipSuccess) e = down(h_len, p->d_len, p->nunits);
  if (e == hipSuccess) e = down(h_mask, p->d_mask, (size_t)p->nunits * p->kcol);
  if (e == hipSuccess) e = down(h_flat, p->d_flat, p->ncovered);
  if (e != hipSuccess) return bk::plan_fail(BK_ERR_HIP, hipGetErrorString(e));
  return BK_OK;
}

}  // extern "C"
