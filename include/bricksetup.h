/**
 * @file bricksetup.h
 * @brief Brick grid numbering/adjacency and array <-> brick copies (host).
 *
 * Same entry points and semantics as bricklib's include/bricksetup.h:
 *  - init_grid<dims>(grid_ptr, dimlist) (bricksetup.h:73-89): bricks of a
 *    grid of dimlist bricks (contiguous first) numbered grid[pos] = pos;
 *    adjacency of every brick from that numbering, a neighbour whose linear
 *    position falls outside the grid becoming brick 0 (so edge rows wrap,
 *    bricksetup.h:42-47 -- only interior bricks get a true neighbourhood);
 *  - init_grid_periodic<dims>: the same numbering with every neighbour
 *    coordinate taken modulo the grid (no ghost bricks needed);
 *  - iter_grid / copyToBrick / copyFromBrick (bricksetup.h:139-221): walk
 *    the bricks s_d in [ghost_d/T_d, (dimlist_d+ghost_d)/T_d) and their
 *    elements side by side with an array of extents
 *    dimlist_d + 2*(padding_d + ghost_d).
 * For 3-D grids the device twins are bk_init_grid / bk_copy_to_brick /
 * bk_copy_from_brick in bk_api.h.
 */
#ifndef BK_BRICKSETUP_H
#define BK_BRICKSETUP_H

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "brick.h"

namespace bk_detail {

template <unsigned dims>
inline BrickInfo<dims> make_grid(unsigned *&grid_ptr, const std::vector<long> &dimlist,
                                 bool periodic) {
  long size = 1;
  std::vector<long> stride;
  for (long a : dimlist) {
    stride.push_back(size);
    size *= a;
  }
  grid_ptr = static_cast<unsigned *>(std::malloc((size_t)size * sizeof(unsigned)));
  for (long p = 0; p < size; ++p) grid_ptr[p] = (unsigned)p;
  BrickInfo<dims> info((unsigned)size);
  constexpr unsigned nslot = static_power<3, dims>::value;
#pragma omp parallel for
  for (long pos = 0; pos < size; ++pos) {
    long c[dims];
    for (unsigned d = 0; d < dims; ++d) c[d] = (pos / stride[d]) % dimlist[d];
    for (unsigned s = 0; s < nslot; ++s) {
      // digit of dimension d (contiguous-first index) has weight 3^d
      long q = 0, lin = pos;
      unsigned rest = s;
      bool inside = true;
      for (unsigned d = 0; d < dims; ++d) {
        const long off = (long)(rest % 3) - 1;
        rest /= 3;
        if (periodic) {
          q += ((c[d] + off + dimlist[d]) % dimlist[d]) * stride[d];
        } else {
          lin += off * stride[d];
        }
      }
      if (!periodic) {
        inside = lin >= 0 && lin < size;
        q = inside ? lin : 0;
      }
      info.adj[grid_ptr[pos]][s] = grid_ptr[q];
    }
  }
  return info;
}

}  // namespace bk_detail

template <unsigned dims>
BrickInfo<dims> init_grid(unsigned *&grid_ptr, const std::vector<long> &dimlist) {
  return bk_detail::make_grid<dims>(grid_ptr, dimlist, false);
}

template <unsigned dims>
BrickInfo<dims> init_grid_periodic(unsigned *&grid_ptr, const std::vector<long> &dimlist) {
  return bk_detail::make_grid<dims>(grid_ptr, dimlist, true);
}

/**
 * Iterate elements side by side in bricks and an array.
 *   dimlist: region iterated (contiguous first); padding: array-only margin
 *   (skipped); ghost: margin of both (skipped); f(bElem &brick, bElem *arr).
 */
template <unsigned dims, typename F, typename T, unsigned... BDims>
inline void iter_grid(const std::vector<long> &dimlist, const std::vector<long> &padding,
                      const std::vector<long> &ghost, bElem *arr, unsigned *grid_ptr,
                      Brick<Dim<BDims...>, T> &brick, F f) {
  std::vector<long> tile = {BDims...};
  std::reverse(tile.begin(), tile.end());  // contiguous first, like the arrays
  std::vector<long> sa(dims), sb(dims), lo(dims), nbr(dims);
  long za = 1, zb = 1, total = 1;
  for (unsigned a = 0; a < dims; ++a) {
    sa[a] = za;
    sb[a] = zb;
    za *= dimlist[a] + 2 * (padding[a] + ghost[a]);
    zb *= (dimlist[a] + 2 * ghost[a]) / tile[a];
    lo[a] = ghost[a] / tile[a];
    nbr[a] = (dimlist[a] + ghost[a]) / tile[a] - lo[a];
    total *= nbr[a];
  }
  long belems = 1;
  for (unsigned a = 0; a < dims; ++a) belems *= tile[a];
#pragma omp parallel for
  for (long t = 0; t < total; ++t) {
    long s[dims], rem = t;
    for (unsigned a = 0; a < dims; ++a) {
      s[a] = lo[a] + rem % nbr[a];
      rem /= nbr[a];
    }
    long gb = 0;
    for (unsigned a = 0; a < dims; ++a) gb += s[a] * sb[a];
    const unsigned b = grid_ptr[gb];
    for (long e = 0; e < belems; ++e) {
      long r = e, ai = 0;
      int x[dims];  // outermost first for the accessor
      for (unsigned a = 0; a < dims; ++a) {
        const long l = r % tile[a];
        r /= tile[a];
        ai += (padding[a] + s[a] * tile[a] + l) * sa[a];
        x[dims - 1 - a] = (int)l;
      }
      f(brick.elem(b, x), arr + ai);
    }
  }
}

/// copyToBrick (bricksetup.h:172-200)
template <unsigned dims, typename T>
inline void copyToBrick(const std::vector<long> &dimlist, const std::vector<long> &padding,
                        const std::vector<long> &ghost, bElem *arr, unsigned *grid_ptr, T &brick) {
  iter_grid<dims>(dimlist, padding, ghost, arr, grid_ptr, brick,
                  [](bElem &br, bElem *a) { br = *a; });
}

template <unsigned dims, typename T>
inline void copyToBrick(const std::vector<long> &dimlist, bElem *arr, unsigned *grid_ptr,
                        T &brick) {
  std::vector<long> zero(dimlist.size(), 0);
  copyToBrick<dims>(dimlist, zero, zero, arr, grid_ptr, brick);
}

/// copyFromBrick (bricksetup.h:213-221)
template <unsigned dims, typename T>
inline void copyFromBrick(const std::vector<long> &dimlist, const std::vector<long> &padding,
                          const std::vector<long> &ghost, bElem *arr, unsigned *grid_ptr,
                          T &brick) {
  iter_grid<dims>(dimlist, padding, ghost, arr, grid_ptr, brick,
                  [](bElem &br, bElem *a) { *a = br; });
}

#endif  // BK_BRICKSETUP_H
