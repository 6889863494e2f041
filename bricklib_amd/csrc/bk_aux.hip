// bk_aux.hip -- ghost refresh, fills, FP64 reductions and array<->brick
// conversion kernels for gfx950.
#include "bk_device.h"
#include "bk_internal.h"

namespace bk {
namespace {

// ---------------------------------------------------------------------------
// Periodic ghost refresh: one wave per (dst, src) pair, whole-brick copy in
// 16-byte vectors.  Stands in for the one-rank periodic MPI exchange of
// weak/main.hip.cpp:249-269 (ghost bricks receive their periodic image).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ghost_copy16(char *__restrict__ base,
                                                    uint64_t step_bytes,
                                                    uint32_t vec_per_brick,
                                                    const uint32_t *__restrict__ pairs,
                                                    uint32_t npairs) {
  const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (p >= npairs) return;
  const int lane = threadIdx.x & 63;
  const uint32_t dst = pairs[2 * p], src = pairs[2 * p + 1];
  const uint4 *s = reinterpret_cast<const uint4 *>(base + (size_t)src * step_bytes);
  uint4 *d = reinterpret_cast<uint4 *>(base + (size_t)dst * step_bytes);
  for (uint32_t v = lane; v < vec_per_brick; v += 64) d[v] = s[v];
}

// Whole-brick copy with NV 16-byte vectors per lane, every load issued
// before the first store (the generic loop above alternates them, since a
// store may alias the next load).
template <int NV>
__global__ __launch_bounds__(256) void ghost_copy16_n(char *__restrict__ base,
                                                      uint64_t step_bytes,
                                                      const uint32_t *__restrict__ pairs,
                                                      uint32_t npairs) {
  const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (p >= npairs) return;
  const int lane = threadIdx.x & 63;
  const uint32_t dst = pairs[2 * p], src = pairs[2 * p + 1];
  const uint4 *s = reinterpret_cast<const uint4 *>(base + (size_t)src * step_bytes);
  uint4 *d = reinterpret_cast<uint4 *>(base + (size_t)dst * step_bytes);
  uint4 t[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = s[lane + 64 * q];
#pragma unroll
  for (int q = 0; q < NV; ++q) d[lane + 64 * q] = t[q];
}

__global__ __launch_bounds__(256) void ghost_copy2(char *__restrict__ base,
                                                   uint64_t step_bytes,
                                                   uint32_t half_per_brick,
                                                   const uint32_t *__restrict__ pairs,
                                                   uint32_t npairs) {
  const uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (p >= npairs) return;
  const int lane = threadIdx.x & 63;
  const uint32_t dst = pairs[2 * p], src = pairs[2 * p + 1];
  const uint16_t *s = reinterpret_cast<const uint16_t *>(base + (size_t)src * step_bytes);
  uint16_t *d = reinterpret_cast<uint16_t *>(base + (size_t)dst * step_bytes);
  for (uint32_t v = lane; v < half_per_brick; v += 64) d[v] = s[v];
}

// ---------------------------------------------------------------------------
// streaming copy used to measure achievable HBM bandwidth: 16-byte vectors,
// four loads in flight per lane before their (non-temporal) stores
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_copy16(const u32x4 *__restrict__ s,
                                                     u32x4 *__restrict__ d, size_t n) {
  const size_t stride = (size_t)gridDim.x * 1024;
  size_t q = (size_t)blockIdx.x * 1024 + threadIdx.x;
  for (; q + 768 < n; q += stride) {
    const u32x4 t0 = s[q], t1 = s[q + 256], t2 = s[q + 512], t3 = s[q + 768];
    __builtin_nontemporal_store(t0, d + q);
    __builtin_nontemporal_store(t1, d + q + 256);
    __builtin_nontemporal_store(t2, d + q + 512);
    __builtin_nontemporal_store(t3, d + q + 768);
  }
  for (; q < n; q += 256) d[q] = s[q];
}

// ---------------------------------------------------------------------------
// register-only multiply-add loop used to measure the VALU rate the part
// sustains: C independent chains per lane, the coefficient wave-uniform (an
// SGPR operand, as in the stencil kernels), 64 FMAs per iteration
// ---------------------------------------------------------------------------
template <class T, int C>
__global__ __launch_bounds__(64) void valu_fma_probe(T *__restrict__ out, T c0, T c1,
                                                     uint32_t iters) {
  T acc[C], x[C];
#pragma unroll
  for (int i = 0; i < C; ++i) {
    acc[i] = T(threadIdx.x) * T(1e-3) + T(i);
    x[i] = T(threadIdx.x + i + 1) * T(0.5);
  }
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 32 / C; ++r)
#pragma unroll
      for (int i = 0; i < C; ++i) acc[i] = fmadd(c0, x[i], acc[i]);
#pragma unroll
    for (int r = 0; r < 32 / C; ++r)
#pragma unroll
      for (int i = 0; i < C; ++i) acc[i] = fmadd(c1, x[i], acc[i]);
  }
  T s = T(0);
#pragma unroll
  for (int i = 0; i < C; ++i) s += acc[i];
  out[(size_t)blockIdx.x * 64 + threadIdx.x] = s;
}

// Batched periodic refresh: one persistent launch refreshes the input field
// of every request of a table (all of one dtype); work item = one (dst,
// src) ghost pair of one request; NV 16-byte vectors per lane per brick.
template <int NV>
__global__ __launch_bounds__(256) void ghost_batched(const BatchTable *tab_, uint32_t tick) {
  const kconst_ptr<BatchTable> tab = (kconst_ptr<BatchTable>)tab_;
  const uint32_t nsegs = tab->nsegs;
  const uint32_t total = tab->prefix[nsegs];
  const uint32_t w = uniform(threadIdx.x >> 6);  // wave-uniform item, segment, pointers
  const int lane = threadIdx.x & 63;
  SegCursor cur;
  for (uint32_t p = blockIdx.x * 4u + w; p < total; p += gridDim.x * 4u) {
    cur.seek(tab->prefix, nsegs, p);
    const auto &sg = tab->seg[cur.s];
    if (tick >= sg.ticks) continue;
    const bool odd = ((sg.parity + tick) & 1u) != 0;
    char *base = static_cast<char *>(odd ? sg.f1 : sg.f0);
    const uint32_t q = p - cur.lo;
    const uint32_t dst = sg.items[2 * q], src = sg.items[2 * q + 1];
    const size_t step_bytes = (size_t)sg.step * (NV * 16 * 64 / 512);
    const uint4 *sp = reinterpret_cast<const uint4 *>(base + (size_t)src * step_bytes);
    uint4 *dp = reinterpret_cast<uint4 *>(base + (size_t)dst * step_bytes);
    uint4 t[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) t[v] = sp[lane + 64 * v];
#pragma unroll
    for (int v = 0; v < NV; ++v) dp[lane + 64 * v] = t[v];
  }
}

template <int NV>
hipError_t launch_ghost_batched_t(const BatchTable *tab, uint32_t tick, hipStream_t s) {
  static int nblk = 0;
  if (nblk == 0) {
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    nblk = cus * 8;
  }
  hipLaunchKernelGGL(ghost_batched<NV>, dim3(nblk), dim3(256), 0, s, tab, tick);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// zero fill (a kernel, so it can sit inside a captured graph)
// ---------------------------------------------------------------------------
__global__ void zero16(uint4 *p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    p[q] = make_uint4(0, 0, 0, 0);
}
__global__ void zero1(unsigned char *p, size_t n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0;
}

// ---------------------------------------------------------------------------
// seeded uniform fill of the interior bricks (SplitMix64 finaliser of
// seed + (n+1)*golden, 53-bit uniform), same values as the oracle
// ---------------------------------------------------------------------------
// Brick extents (k, j, i) and vector fold (j, i) of a layout.
struct FoldGeom {
  uint32_t bk, bj, bi, fj, fi;
};

__device__ __forceinline__ uint64_t hash64(uint64_t seed, uint64_t n) {
  uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

template <class T>
__global__ __launch_bounds__(256) void fill_uniform(T *__restrict__ dat, uint64_t step,
                                                    uint32_t nb0, uint32_t nb1,
                                                    uint32_t nb2, uint32_t ghost,
                                                    FoldGeom g, uint64_t seed, double lo,
                                                    double hi) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nbricks = (uint64_t)nb0 * nb1 * nb2;
  if (t >= nbricks * 512) return;
  const uint64_t bi = t >> 9;
  const uint32_t e = (uint32_t)(t & 511);
  const uint32_t ti = (uint32_t)(bi % nb0), tj = (uint32_t)((bi / nb0) % nb1),
                 tk = (uint32_t)(bi / ((uint64_t)nb0 * nb1));
  // element e -> (k, j, i) through the fold: vector e / 64, lane e % 64
  const uint32_t vec = e >> 6, lane = e & 63;
  const uint32_t nj = g.bj / g.fj, ni = g.bi / g.fi;
  const uint32_t k = vec / (nj * ni);
  const uint32_t j = ((vec / ni) % nj) * g.fj + lane / g.fi;
  const uint32_t i = (vec % ni) * g.fi + lane % g.fi;
  const uint64_t n0 = (uint64_t)nb0 * g.bi, n1 = (uint64_t)nb1 * g.bj;
  const uint64_t x = (uint64_t)ti * g.bi + i, y = (uint64_t)tj * g.bj + j,
                 z = (uint64_t)tk * g.bk + k;
  const double u = (double)(hash64(seed, x + n0 * (y + n1 * z)) >> 11) * 0x1.0p-53;
  const double v = lo + (hi - lo) * u;
  const uint64_t S0 = nb0 + 2 * ghost, S1 = nb1 + 2 * ghost;
  const uint64_t b = (ti + ghost) + S0 * ((tj + ghost) + S1 * (tk + ghost));
  if constexpr (sizeof(T) == 8) {
    dat[b * step + e] = v;
  } else if constexpr (sizeof(T) == 4) {
    dat[b * step + e] = (float)v;
  } else {
    narrow(dat[b * step + e], (float)v);
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions: pass 1 writes one partial per block, pass 2 combines them
// in a fixed order (results do not depend on scheduling).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ld64(int dt, const void *p, size_t idx) {
  if (dt == BK_F64) return static_cast<const double *>(p)[idx];
  if (dt == BK_F32) return (double)static_cast<const float *>(p)[idx];
  return (double)widen(static_cast<const bf16 *>(p)[idx]);
}

template <int NV>
__device__ void block_reduce(double (&v)[NV], const bool (&is_max)[NV], double *out) {
  __shared__ double sh[NV][256];
#pragma unroll
  for (int q = 0; q < NV; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < NV; ++q) {
        const double o = sh[q][threadIdx.x + s];
        sh[q][threadIdx.x] = is_max[q] ? fmax(sh[q][threadIdx.x], o) : sh[q][threadIdx.x] + o;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = sh[q][0];
  }
}

__global__ __launch_bounds__(256) void stats_pass1(int dt, const void *dat, uint64_t step,
                                                   const uint32_t *list, uint32_t nlist,
                                                   uint32_t brick_elems, double *part) {
  double v[3] = {0.0, 0.0, 0.0};
  const uint64_t total = (uint64_t)nlist * brick_elems;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint64_t l = t / brick_elems, e = t - l * brick_elems;
    const double x = ld64(dt, dat, (size_t)list[l] * step + e);
    v[0] += x;
    v[1] += x * x;
    v[2] = fmax(v[2], fabs(x));
  }
  const bool mx[3] = {false, false, true};
  block_reduce<3>(v, mx, part + (size_t)blockIdx.x * 3);
}

__global__ __launch_bounds__(256) void compare_pass1(int dta, const void *a, uint64_t sa,
                                                     int dtr, const void *r, uint64_t sr,
                                                     const uint32_t *list, uint32_t nlist,
                                                     uint32_t brick_elems, double floor_,
                                                     double atol, double rtol,
                                                     double *part) {
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const uint64_t total = (uint64_t)nlist * brick_elems;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const uint64_t l = t / brick_elems, e = t - l * brick_elems;
    const size_t b = list[l];
    const double xa = ld64(dta, a, b * sa + e);
    const double xr = ld64(dtr, r, b * sr + e);
    const double d = fabs(xa - xr);
    const double den = fmax(fabs(xr), floor_);
    v[0] = fmax(v[0], d);
    v[1] = fmax(v[1], d / den);
    v[2] += d * d;
    v[3] += xr * xr;
    v[4] += (d > atol + rtol * fabs(xr)) ? 1.0 : 0.0;
  }
  const bool mx[5] = {true, true, false, false, false};
  block_reduce<5>(v, mx, part + (size_t)blockIdx.x * 5);
}

template <int NV>
__global__ __launch_bounds__(256) void reduce_pass2(const double *part, int nblocks, double *out,
                                                    int m0, int m1, int m2, int m3, int m4) {
  double v[NV];
  const int ms[5] = {m0, m1, m2, m3, m4};
  bool mx[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    v[q] = 0.0;
    mx[q] = ms[q] != 0;
  }
  for (int t = threadIdx.x; t < nblocks; t += 256) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double o = part[(size_t)t * NV + q];
      v[q] = mx[q] ? fmax(v[q], o) : v[q] + o;
    }
  }
  __syncthreads();
  block_reduce<NV>(v, mx, out);  // out may be part itself: every read precedes the write
}

// ---------------------------------------------------------------------------
// copyToBrick / copyFromBrick (bricksetup.h:139-221) on the device
// ---------------------------------------------------------------------------
struct CopyGeom {
  int64_t sa1, sa2;          // array strides (elements) for j and k
  int64_t pad0, pad1, pad2;  // padding per axis
  int64_t sb1, sb2;          // grid strides (bricks)
  int64_t k0, j0, i0;        // first brick per axis
  int64_t nk, nj, ni;        // bricks per axis in the iteration
};

template <class T, bool TO_BRICK>
__global__ __launch_bounds__(256) void copy_brick(CopyGeom g, T *arr, const uint32_t *grid,
                                                  T *dat, uint64_t step) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nbr = (uint64_t)g.nk * g.nj * g.ni;
  if (t >= nbr * 512) return;
  const uint64_t bi = t >> 9;
  const int e = (int)(t & 511);
  const int64_t ti = g.i0 + (int64_t)(bi % g.ni);
  const int64_t tj = g.j0 + (int64_t)((bi / g.ni) % g.nj);
  const int64_t tk = g.k0 + (int64_t)(bi / ((uint64_t)g.ni * g.nj));
  const int k = e >> 6, j = (e >> 3) & 7, i = e & 7;
  const size_t b = grid[ti + tj * g.sb1 + tk * g.sb2];
  const size_t ai = (size_t)((g.pad0 + ti * 8 + i) + (g.pad1 + tj * 8 + j) * g.sa1 +
                             (g.pad2 + tk * 8 + k) * g.sa2);
  if constexpr (TO_BRICK) dat[b * step + e] = arr[ai];
  else arr[ai] = dat[b * step + e];
}

template <class T>
hipError_t launch_copy_t(int to_brick, const CopyGeom &g, void *arr, const uint32_t *grid,
                         void *dat, size_t step, hipStream_t s) {
  const uint64_t n = (uint64_t)g.nk * g.nj * g.ni * 512;
  if (n == 0) return hipSuccess;
  const uint64_t nblk = (n + 255) / 256;
  if (to_brick)
    hipLaunchKernelGGL((copy_brick<T, true>), dim3((uint32_t)nblk), dim3(256), 0, s, g,
                       static_cast<T *>(arr), grid, static_cast<T *>(dat), (uint64_t)step);
  else
    hipLaunchKernelGGL((copy_brick<T, false>), dim3((uint32_t)nblk), dim3(256), 0, s, g,
                       static_cast<T *>(arr), grid, static_cast<T *>(dat), (uint64_t)step);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_ghost_refresh(void *dat, size_t step_bytes, size_t brick_bytes,
                                const uint32_t *pairs, uint32_t npairs, hipStream_t s) {
  if (npairs == 0) return hipSuccess;
  const uint32_t nwg = (npairs + 3u) / 4u;
  const bool v16 = (reinterpret_cast<uintptr_t>(dat) % 16 == 0) && (step_bytes % 16 == 0) &&
                   (brick_bytes % 16 == 0);
  char *b = static_cast<char *>(dat);
  if (v16 && brick_bytes == 4096)
    hipLaunchKernelGGL(ghost_copy16_n<4>, dim3(nwg), dim3(256), 0, s, b, (uint64_t)step_bytes,
                       pairs, npairs);
  else if (v16 && brick_bytes == 2048)
    hipLaunchKernelGGL(ghost_copy16_n<2>, dim3(nwg), dim3(256), 0, s, b, (uint64_t)step_bytes,
                       pairs, npairs);
  else if (v16 && brick_bytes == 1024)
    hipLaunchKernelGGL(ghost_copy16_n<1>, dim3(nwg), dim3(256), 0, s, b, (uint64_t)step_bytes,
                       pairs, npairs);
  else if (v16)
    hipLaunchKernelGGL(ghost_copy16, dim3(nwg), dim3(256), 0, s, b, (uint64_t)step_bytes,
                       (uint32_t)(brick_bytes / 16), pairs, npairs);
  else
    hipLaunchKernelGGL(ghost_copy2, dim3(nwg), dim3(256), 0, s, static_cast<char *>(dat),
                       (uint64_t)step_bytes, (uint32_t)(brick_bytes / 2), pairs, npairs);
  return hipGetLastError();
}

hipError_t launch_ghost_batched(int dtype, const BatchTable *tab, uint32_t tick, hipStream_t s) {
  switch (dtype) {
    case BK_F64: return launch_ghost_batched_t<4>(tab, tick, s);
    case BK_F32: return launch_ghost_batched_t<2>(tab, tick, s);
    case BK_BF16: return launch_ghost_batched_t<1>(tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stream_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  size_t nblk = (n + 1023) / 1024;
  const size_t cap = (size_t)cus * 8;
  if (nblk > cap) nblk = cap;
  hipLaunchKernelGGL(stream_copy16, dim3((uint32_t)nblk), dim3(256), 0, s,
                     static_cast<const u32x4 *>(src), static_cast<u32x4 *>(dst), n);
  return hipGetLastError();
}

hipError_t launch_valu_probe(int dtype, int chains, uint32_t nblocks, uint32_t iters, void *out,
                             hipStream_t s) {
  const dim3 g(nblocks), b(64);
  if (dtype == BK_F64 && chains == 1)
    hipLaunchKernelGGL((valu_fma_probe<double, 1>), g, b, 0, s, static_cast<double *>(out),
                       0.999, 1.0001, iters);
  else if (dtype == BK_F64 && chains == 8)
    hipLaunchKernelGGL((valu_fma_probe<double, 8>), g, b, 0, s, static_cast<double *>(out),
                       0.999, 1.0001, iters);
  else if (dtype == BK_F32 && chains == 1)
    hipLaunchKernelGGL((valu_fma_probe<float, 1>), g, b, 0, s, static_cast<float *>(out),
                       0.999f, 1.0001f, iters);
  else if (dtype == BK_F32 && chains == 8)
    hipLaunchKernelGGL((valu_fma_probe<float, 8>), g, b, 0, s, static_cast<float *>(out),
                       0.999f, 1.0001f, iters);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_zero(void *dst, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  unsigned char *p = static_cast<unsigned char *>(dst);
  size_t head = (16 - reinterpret_cast<uintptr_t>(p) % 16) % 16;
  if (head > bytes) head = bytes;
  if (head) hipLaunchKernelGGL(zero1, dim3(1), dim3(64), 0, s, p, head);
  const size_t nvec = (bytes - head) / 16;
  if (nvec) {
    size_t nblk = (nvec + 255) / 256;
    if (nblk > 65536) nblk = 65536;
    hipLaunchKernelGGL(zero16, dim3((uint32_t)nblk), dim3(256), 0, s,
                       reinterpret_cast<uint4 *>(p + head), nvec);
  }
  const size_t tail = bytes - head - nvec * 16;
  if (tail)
    hipLaunchKernelGGL(zero1, dim3(1), dim3(64), 0, s, p + head + nvec * 16, tail);
  return hipGetLastError();
}

hipError_t launch_fill_uniform(int layout, int dtype, void *dat, size_t step, const long nb[3],
                               int ghost, uint64_t seed, double lo, double hi, hipStream_t s) {
  const uint64_t n = (uint64_t)nb[0] * nb[1] * nb[2] * 512;
  if (n == 0) return hipSuccess;
  FoldGeom g;
  switch (layout) {
    case BK_LAYOUT_8x8x8: g = {8, 8, 8, 8, 8}; break;
    case BK_LAYOUT_4x8x16: g = {4, 8, 16, 4, 16}; break;
    case BK_LAYOUT_4x8x16_F88: g = {4, 8, 16, 8, 8}; break;
    default: return hipErrorInvalidValue;
  }
  const uint32_t nblk = (uint32_t)((n + 255) / 256);
  const uint32_t n0 = (uint32_t)nb[0], n1 = (uint32_t)nb[1], n2 = (uint32_t)nb[2];
  switch (dtype) {
    case BK_F64:
      hipLaunchKernelGGL(fill_uniform<double>, dim3(nblk), dim3(256), 0, s,
                         static_cast<double *>(dat), (uint64_t)step, n0, n1, n2,
                         (uint32_t)ghost, g, seed, lo, hi);
      break;
    case BK_F32:
      hipLaunchKernelGGL(fill_uniform<float>, dim3(nblk), dim3(256), 0, s,
                         static_cast<float *>(dat), (uint64_t)step, n0, n1, n2,
                         (uint32_t)ghost, g, seed, lo, hi);
      break;
    case BK_BF16:
      hipLaunchKernelGGL(fill_uniform<bf16>, dim3(nblk), dim3(256), 0, s,
                         static_cast<bf16 *>(dat), (uint64_t)step, n0, n1, n2,
                         (uint32_t)ghost, g, seed, lo, hi);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_field_stats(int dtype, const void *dat, size_t step, const uint32_t *list,
                              uint32_t nlist, size_t brick_elems, double *scratch, double *out,
                              hipStream_t s) {
  hipLaunchKernelGGL(stats_pass1, dim3(kReduceBlocks), dim3(256), 0, s, dtype, dat,
                     (uint64_t)step, list, nlist, (uint32_t)brick_elems, scratch);
  hipLaunchKernelGGL(reduce_pass2<3>, dim3(1), dim3(256), 0, s, scratch, kReduceBlocks,
                     out ? out : scratch, 0, 0, 1, 0, 0);
  return hipGetLastError();
}

hipError_t launch_field_compare(int dta, const void *a, size_t step_a, int dtr, const void *r,
                                size_t step_r, const uint32_t *list, uint32_t nlist,
                                size_t brick_elems, double floor_, double atol, double rtol,
                                double *scratch, hipStream_t s) {
  hipLaunchKernelGGL(compare_pass1, dim3(kReduceBlocks), dim3(256), 0, s, dta, a,
                     (uint64_t)step_a, dtr, r, (uint64_t)step_r, list, nlist,
                     (uint32_t)brick_elems, floor_, atol, rtol, scratch);
  hipLaunchKernelGGL(reduce_pass2<5>, dim3(1), dim3(256), 0, s, scratch, kReduceBlocks, scratch,
                     1, 1, 0, 0, 0);
  return hipGetLastError();
}

hipError_t launch_copy_brick(int to_brick, int dtype, const long dims[3], const long padding[3],
                             const long ghost[3], void *arr, const uint32_t *grid, void *dat,
                             size_t step, hipStream_t s) {
  CopyGeom g;
  const int64_t e0 = dims[0] + 2 * (padding[0] + ghost[0]);
  const int64_t e1 = dims[1] + 2 * (padding[1] + ghost[1]);
  g.sa1 = e0;
  g.sa2 = e0 * e1;
  g.pad0 = padding[0];
  g.pad1 = padding[1];
  g.pad2 = padding[2];
  const int64_t b0 = (dims[0] + 2 * ghost[0]) / 8, b1 = (dims[1] + 2 * ghost[1]) / 8;
  g.sb1 = b0;
  g.sb2 = b0 * b1;
  g.i0 = ghost[0] / 8;
  g.j0 = ghost[1] / 8;
  g.k0 = ghost[2] / 8;
  g.ni = (dims[0] + ghost[0]) / 8 - g.i0;
  g.nj = (dims[1] + ghost[1]) / 8 - g.j0;
  g.nk = (dims[2] + ghost[2]) / 8 - g.k0;
  switch (dtype) {
    case BK_F64: return launch_copy_t<double>(to_brick, g, arr, grid, dat, step, s);
    case BK_F32: return launch_copy_t<float>(to_brick, g, arr, grid, dat, step, s);
    case BK_BF16: return launch_copy_t<uint16_t>(to_brick, g, arr, grid, dat, step, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bk
