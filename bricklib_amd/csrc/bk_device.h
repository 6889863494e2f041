// bk_device.h -- device-side helpers shared by the brick kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bk {

// BF16 storage word.  Widened to FP32 exactly; narrowed with round to
// nearest even (NaN stays quiet).
struct bf16 {
  uint16_t u;
};

template <class T> struct acc_of;
template <> struct acc_of<double> { using type = double; };
template <> struct acc_of<float> { using type = float; };
template <> struct acc_of<bf16> { using type = float; };

__device__ __forceinline__ double widen(double x) { return x; }
__device__ __forceinline__ float widen(float x) { return x; }
__device__ __forceinline__ float widen(bf16 x) {
  return __uint_as_float((uint32_t)x.u << 16);
}

__device__ __forceinline__ void narrow(double &d, double v) { d = v; }
__device__ __forceinline__ void narrow(float &d, float v) { d = v; }
__device__ __forceinline__ void narrow(bf16 &d, float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) {
    d.u = (uint16_t)((u >> 16) | 0x40u);
  } else {
    u += 0x7fffu + ((u >> 16) & 1u);
    d.u = (uint16_t)(u >> 16);
  }
}

// Output stores bypass L2 allocation: a sweep never reads what it writes,
// and keeping outputs out of L2 leaves it to the neighbour faces (plain
// stores ran the 512^3 block kernels 5-7% slower: FP64 7-point 420 against
// 393 us, 25-point 474 against 452, FP32 7-point 216 against 203).
__device__ __forceinline__ void store_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(bf16 *p, float v) {
  bf16 h;
  narrow(h, v);
  __builtin_nontemporal_store(h.u, &p->u);
}

__device__ __forceinline__ double fmadd(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
__device__ __forceinline__ float fmadd(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// std::max(x, 0.0) of the reference's cond stencil: x < 0 ? 0 : x
// (keeps -0.0 and NaN as std::max does; fmax would not).
template <class A> __device__ __forceinline__ A relu_ref(A x) {
  return x < A(0) ? A(0) : x;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup w runs on XCD w % 8).  Map them so each XCD gets one
// contiguous run of the logical index space: neighbouring bricks then share
// that XCD's L2.  A bijection on [0, n) for any n.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t w, uint32_t n) {
  const uint32_t q = n >> 3, r = n & 7u, x = w & 7u;
  return x * q + (x < r ? x : r) + (w >> 3);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Pointer into the constant address space: uniform loads through it are
// scalar loads into SGPRs.  Used for coefficient tables and batch tables.
template <class A>
using kconst_ptr = const A __attribute__((address_space(4))) *;

// Segment of a persistent batched launch holding `item`: the s with
// prefix[s] <= item < prefix[s+1] (prefix has nsegs+1 entries).
__device__ __forceinline__ uint32_t find_segment(kconst_ptr<uint32_t> prefix, uint32_t nsegs,
                                                 uint32_t item) {
  uint32_t lo = 0, hi = nsegs;  // invariant: prefix[lo] <= item < prefix[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Segment cursor for a persistent batched loop in which each wave visits
// its items in increasing order: the current and the next segment are
// checked first, and the binary search runs only on a longer jump (or the
// first item).  All members are wave-uniform.  A found segment is never
// empty, so lo == hi marks the cursor as unset.
struct SegCursor {
  uint32_t s = 0, lo = 0, hi = 0;  // segment s holds items [lo, hi)

  __device__ __forceinline__ void seek(kconst_ptr<uint32_t> prefix, uint32_t nsegs,
                                       uint32_t item) {
    if (item >= lo && item < hi) return;
    if (hi > lo && item >= hi && s + 2 <= nsegs && item < prefix[s + 2]) {
      s += 1;
    } else {
      s = find_segment(prefix, nsegs, item);
    }
    lo = prefix[s];
    hi = prefix[s + 1];
  }
};

}  // namespace bk
