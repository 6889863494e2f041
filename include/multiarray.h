/**
 * @file multiarray.h
 * @brief Multidimensional array helpers (host).
 *
 * The helpers of bricklib's include/multiarray.h:18-41 with the behaviour of
 * src/multiarray.cpp:26-63, header-only here (nothing to link):
 *  - uninitArray: an ALIGN-aligned array of prod(list) bElem, its element
 *    count returned in `size`; release with free();
 *  - randomArray: values uniform in [0, 1), each thread of the OpenMP loop
 *    drawing from its own 64-bit Mersenne twister seeded from
 *    std::random_device (so the values are not reproducible, as in the
 *    reference);
 *  - zeroArray: all zeros;
 *  - compareArray: true when every element pair passes compareBrick's rule
 *    diff < BRICK_TOLERANCE || diff < (|a| + |b|) * BRICK_TOLERANCE.
 */
#ifndef BK_MULTIARRAY_H
#define BK_MULTIARRAY_H

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <random>
#include <vector>

#include "brick.h"
#include "cmpconst.h"

inline bElem *uninitArray(const std::vector<long> &list, long &size) {
  size = 1;
  for (long d : list) size *= d;
  size_t bytes = (size_t)size * sizeof(bElem);
  bytes = (bytes + ALIGN - 1) / ALIGN * ALIGN;  // aligned_alloc wants a multiple
  return static_cast<bElem *>(::aligned_alloc(ALIGN, bytes ? bytes : ALIGN));
}

inline bElem *randomArray(const std::vector<long> &list) {
  long size = 0;
  bElem *arr = uninitArray(list, size);
#pragma omp parallel
  {
    std::random_device rd;
    std::mt19937_64 gen(((uint64_t)rd() << 32) ^ rd());
    std::uniform_real_distribution<bElem> u(0, 1);
#pragma omp for
    for (long l = 0; l < size; ++l) arr[l] = u(gen);
  }
  return arr;
}

inline bElem *zeroArray(const std::vector<long> &list) {
  long size = 0;
  bElem *arr = uninitArray(list, size);
#pragma omp parallel for
  for (long l = 0; l < size; ++l) arr[l] = 0;
  return arr;
}

inline bool compareArray(const std::vector<long> &list, bElem *arrA, bElem *arrB) {
  long size = 1;
  for (long d : list) size *= d;
  long bad = 0;
#pragma omp parallel for reduction(+ : bad)
  for (long l = 0; l < size; ++l) {
    const double a = arrA[l], b = arrB[l];
    const double diff = std::abs(a - b);
    const bool ok = diff < BRICK_TOLERANCE || diff < (std::abs(a) + std::abs(b)) * BRICK_TOLERANCE;
    bad += ok ? 0 : 1;
  }
  return bad == 0;
}

#endif  // BK_MULTIARRAY_H
