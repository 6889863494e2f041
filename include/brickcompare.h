/**
 * @file brickcompare.h
 * @brief Compare bricks with an array (host).
 *
 * compareBrick (bricklib include/brickcompare.h:30-79): true when every
 * element of the brick region passes
 *     diff < BRICK_TOLERANCE || diff < (|brick| + |array|) * BRICK_TOLERANCE.
 * compareBrickStats returns the error figures the GPU parity checks use
 * (max abs, max rel with a floor, RMS rel) over the same region.  For
 * device-resident fields, bk_field_compare (bk_api.h) computes them in FP64
 * on the GPU.
 */
#ifndef BK_BRICKCOMPARE_H
#define BK_BRICKCOMPARE_H

#include <cmath>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "bricksetup.h"
#include "cmpconst.h"

template <unsigned dims, typename T>
inline bool compareBrick(const std::vector<long> &dimlist, const std::vector<long> &padding,
                         const std::vector<long> &ghost, bElem *arr, unsigned *grid_ptr,
                         T &brick) {
  long bad = 0;
  iter_grid<dims>(dimlist, padding, ghost, arr, grid_ptr, brick, [&bad](bElem &b, bElem *a) {
    const double diff = std::abs((double)b - (double)*a);
    const bool ok = (diff < BRICK_TOLERANCE) ||
                    (diff < (std::abs((double)b) + std::abs((double)*a)) * BRICK_TOLERANCE);
    if (!ok) {
#pragma omp atomic
      ++bad;
    }
  });
  return bad == 0;
}

template <unsigned dims, typename T>
inline bool compareBrick(const std::vector<long> &dimlist, bElem *arr, unsigned *grid_ptr,
                         T &brick) {
  std::vector<long> zero(dimlist.size(), 0);
  return compareBrick<dims, T>(dimlist, zero, zero, arr, grid_ptr, brick);
}

struct BrickErrors {
  double max_abs = 0, max_rel = 0, sum_sq_diff = 0, sum_sq_ref = 0;
  long points = 0;
  double rms_rel() const { return sum_sq_ref > 0 ? std::sqrt(sum_sq_diff / sum_sq_ref) : 0.0; }
};

/// Error of the bricks against the array taken as the reference:
/// max |b - a|, max |b - a| / max(|a|, floor), RMS relative error.  Each
/// OpenMP thread accumulates its own partial figures (no lock per element);
/// they are combined at the end.
template <unsigned dims, typename T>
inline BrickErrors compareBrickStats(const std::vector<long> &dimlist,
                                     const std::vector<long> &padding,
                                     const std::vector<long> &ghost, bElem *arr,
                                     unsigned *grid_ptr, T &brick, double floor_ = 1.0) {
  struct alignas(64) Part {
    BrickErrors e;
  };
#ifdef _OPENMP
  std::vector<Part> part((size_t)omp_get_max_threads());
#else
  std::vector<Part> part(1);
#endif
  iter_grid<dims>(dimlist, padding, ghost, arr, grid_ptr, brick, [&](bElem &b, bElem *a) {
#ifdef _OPENMP
    BrickErrors &e = part[(size_t)omp_get_thread_num()].e;
#else
    BrickErrors &e = part[0].e;
#endif
    const double d = std::abs((double)b - (double)*a);
    const double den = std::abs((double)*a) > floor_ ? std::abs((double)*a) : floor_;
    e.max_abs = d > e.max_abs ? d : e.max_abs;
    e.max_rel = d / den > e.max_rel ? d / den : e.max_rel;
    e.sum_sq_diff += d * d;
    e.sum_sq_ref += (double)*a * (double)*a;
    ++e.points;
  });
  BrickErrors e;
  for (const Part &p : part) {
    e.max_abs = p.e.max_abs > e.max_abs ? p.e.max_abs : e.max_abs;
    e.max_rel = p.e.max_rel > e.max_rel ? p.e.max_rel : e.max_rel;
    e.sum_sq_diff += p.e.sum_sq_diff;
    e.sum_sq_ref += p.e.sum_sq_ref;
    e.points += p.e.points;
  }
  return e;
}

#endif  // BK_BRICKCOMPARE_H
