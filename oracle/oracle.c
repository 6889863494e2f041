/*
 * oracle.c -- CPU restatement of bricklib's brick-layout stencil path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the reference the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark compare against.
 * Nothing in the product library calls, links or loads it.
 *
 * What it restates (CtopCsUtahEdu/bricklib, file:line):
 *   - brick numbering and 27-neighbour adjacency: init_grid / init_iter /
 *     init_fill, include/bricksetup.h:30-89 (grid[pos] = pos, out-of-range
 *     neighbour -> brick 0), plus a periodic variant;
 *   - element addressing of Brick<Dim<8,8,8>,Dim<8,8>>: the accessor of
 *     include/brick.h:213-327 reduces to dat[adj[b][kd*9+jd*3+id]*step +
 *     k*64 + j*8 + i] for the 8x8 fold;
 *   - copyToBrick / copyFromBrick / iter_grid, include/bricksetup.h:139-221;
 *   - compareBrick's tolerance rule, include/brickcompare.h:35-39;
 *   - the brick sweep loop of stencils/3axis.cpp:51-66 (7pt) and :123-140
 *     (cond), the array loops of :32-37 and :115-122, and the 13/25/125
 *     point tap sets of stencils/mpi13pt.py:18-31, mpi25pt.py:20-45,
 *     mpi125pt.py:12-162 (one coefficient per offset);
 *   - the one-rank periodic ghost exchange (weak/main.hip.cpp:249-269 over
 *     the periodic communicator of weak/args.cpp:107-108) as a whole-brick
 *     copy from the periodic image.
 *
 * Two evaluation orders per point:
 *   OR_MODE_REF   products summed left to right in the reference's textual
 *                 order, separate multiply and add (g++ -std=c++14 on x86-64
 *                 without -mfma contracts nothing).  7pt/cond follow
 *                 3axis.cpp:60-63 (k+1, k-1, j+1, j-1, i+1, i-1, centre);
 *                 13/25 follow the .py files; 125 goes (dk, dj, di)
 *                 lexicographic.
 *   OR_MODE_FMA   the library's fixed order: acc = c_0*x_0 then
 *                 acc = fma(c_t, x_t, acc) over the coefficient order of
 *                 bk_api.h.  The HIP kernels must match this bit for bit.
 *
 * Built with -ffp-contract=off so the compiler adds no fused operations.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_F64 0
#define OR_F32 1
#define OR_BF16 2

#define OR_STAR7 0
#define OR_STAR13 1
#define OR_STAR25 2
#define OR_CUBE125 3
#define OR_COND7 4

#define OR_MODE_FMA 0
#define OR_MODE_REF 1

/* ---------------------------------------------------------------------- */
/* seeded counter-based RNG: SplitMix64 finaliser of seed + (n+1)*golden   */
/* ---------------------------------------------------------------------- */
uint64_t or_hash64(uint64_t seed, uint64_t n) {
  uint64_t z = seed + (n + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* 53-bit uniform in [0, 1) */
double or_u01(uint64_t seed, uint64_t n) {
  return (double)(or_hash64(seed, n) >> 11) * 0x1.0p-53;
}

/* ---------------------------------------------------------------------- */
/* BF16 <-> FP32 (round to nearest even, NaN kept quiet)                   */
/* ---------------------------------------------------------------------- */
static inline float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

uint16_t or_f32_to_bf16(float f) { return f32_to_bf16(f); }
float or_bf16_to_f32(uint16_t h) { return bf16_to_f32(h); }

/* ---------------------------------------------------------------------- */
/* grid numbering and adjacency: bricksetup.h:30-89                         */
/* ---------------------------------------------------------------------- */
/* dims contiguous-first: dims[0] = bricks along i, dims[2] along k.
 * The reference walks a pointer grid_ptr +/- stride and accepts it when it
 * lies inside [grid, grid+size) (bricksetup.h:42-47); for an edge brick
 * that makes i-1 of i == 0 land on the last brick of the previous row
 * (row wrap) and anything before the first or past the last brick map to
 * 0.  Same here: the neighbour's linear position is pos + di + dj*s1 +
 * dk*s2, accepted when 0 <= that < size. */
void or_init_grid(const long dims[3], int periodic, uint32_t *grid,
                  uint32_t *adj) {
  const long s1 = dims[0], s2 = dims[0] * dims[1];
  const long size = s2 * dims[2];
  for (long p = 0; p < size; ++p) grid[p] = (uint32_t)p;
#pragma omp parallel for
  for (long k = 0; k < dims[2]; ++k)
    for (long j = 0; j < dims[1]; ++j)
      for (long i = 0; i < dims[0]; ++i) {
        const long pos = i + j * s1 + k * s2;
        uint32_t *a = adj + (size_t)grid[pos] * 27;
        for (int kd = 0; kd < 3; ++kd)
          for (int jd = 0; jd < 3; ++jd)
            for (int id = 0; id < 3; ++id) {
              long v;
              if (periodic) {
                long ii = (i + id - 1 + dims[0]) % dims[0];
                long jj = (j + jd - 1 + dims[1]) % dims[1];
                long kk = (k + kd - 1 + dims[2]) % dims[2];
                v = grid[ii + jj * s1 + kk * s2];
              } else {
                long q = pos + (id - 1) + (jd - 1) * s1 + (kd - 1) * s2;
                v = (q >= 0 && q < size) ? grid[q] : 0;
              }
              a[kd * 9 + jd * 3 + id] = (uint32_t)v;
            }
      }
}

/* interior bricks of a (nb+2)^3 padded grid, k slowest */
uint32_t or_interior_list(const long nb[3], uint32_t *list) {
  const long S0 = nb[0] + 2, S1 = nb[1] + 2;
  uint32_t n = 0;
  for (long k = 1; k <= nb[2]; ++k)
    for (long j = 1; j <= nb[1]; ++j)
      for (long i = 1; i <= nb[0]; ++i) {
        if (list) list[n] = (uint32_t)(i + S0 * (j + S1 * k));
        ++n;
      }
  return n;
}

/* (dst, src) pairs: every ghost brick and its periodic image */
uint32_t or_ghost_pairs(const long nb[3], uint32_t *pairs) {
  const long S0 = nb[0] + 2, S1 = nb[1] + 2, S2 = nb[2] + 2;
  uint32_t n = 0;
  for (long k = 0; k < S2; ++k)
    for (long j = 0; j < S1; ++j)
      for (long i = 0; i < S0; ++i) {
        int ghost = (i == 0 || i == S0 - 1 || j == 0 || j == S1 - 1 ||
                     k == 0 || k == S2 - 1);
        if (!ghost) continue;
        long si = (i - 1 + nb[0]) % nb[0] + 1;
        long sj = (j - 1 + nb[1]) % nb[1] + 1;
        long sk = (k - 1 + nb[2]) % nb[2] + 1;
        if (pairs) {
          pairs[2 * n] = (uint32_t)(i + S0 * (j + S1 * k));
          pairs[2 * n + 1] = (uint32_t)(si + S0 * (sj + S1 * sk));
        }
        ++n;
      }
  return n;
}

/* ---------------------------------------------------------------------- */
/* typed element access                                                    */
/* ---------------------------------------------------------------------- */
static inline double ld_as_f64(int dt, const void *p, size_t idx) {
  switch (dt) {
    case OR_F64: return ((const double *)p)[idx];
    case OR_F32: return (double)((const float *)p)[idx];
    default: return (double)bf16_to_f32(((const uint16_t *)p)[idx]);
  }
}

static inline void st_from_f64(int dt, void *p, size_t idx, double v) {
  switch (dt) {
    case OR_F64: ((double *)p)[idx] = v; break;
    case OR_F32: ((float *)p)[idx] = (float)v; break;
    default: ((uint16_t *)p)[idx] = f32_to_bf16((float)v); break;
  }
}

static size_t dt_size(int dt) { return dt == OR_F64 ? 8 : dt == OR_F32 ? 4 : 2; }

/* seeded fill of the interior of a padded (nb+2)^3 brick grid */
void or_fill_uniform(int dt, void *dat, size_t step, const long nb[3],
                     uint64_t seed, double lo, double hi) {
  const long n0 = nb[0] * 8, n1 = nb[1] * 8, n2 = nb[2] * 8;
  const long S0 = nb[0] + 2, S1 = nb[1] + 2;
#pragma omp parallel for collapse(2)
  for (long z = 0; z < n2; ++z)
    for (long y = 0; y < n1; ++y)
      for (long x = 0; x < n0; ++x) {
        uint64_t lin = (uint64_t)x + (uint64_t)n0 * ((uint64_t)y + (uint64_t)n1 * (uint64_t)z);
        double v = lo + (hi - lo) * or_u01(seed, lin);
        size_t b = (size_t)((x / 8 + 1) + S0 * ((y / 8 + 1) + S1 * (z / 8 + 1)));
        size_t e = (size_t)((z % 8) * 64 + (y % 8) * 8 + (x % 8));
        st_from_f64(dt, dat, b * step + e, v);
      }
}

/* dense periodic field (n0 x n1 x n2, x fastest) with the same values */
void or_fill_uniform_array(double *arr, const long n[3], uint64_t seed,
                           double lo, double hi) {
  const long total = n[0] * n[1] * n[2];
#pragma omp parallel for
  for (long l = 0; l < total; ++l) arr[l] = lo + (hi - lo) * or_u01(seed, (uint64_t)l);
}

void or_ghost_refresh(int dt, void *dat, size_t step, size_t brick_elems,
                      const uint32_t *pairs, uint32_t npairs) {
  const size_t es = dt_size(dt);
#pragma omp parallel for
  for (long p = 0; p < (long)npairs; ++p) {
    char *base = (char *)dat;
    memcpy(base + (size_t)pairs[2 * p] * step * es,
           base + (size_t)pairs[2 * p + 1] * step * es, brick_elems * es);
  }
}

/* ---------------------------------------------------------------------- */
/* copyToBrick / copyFromBrick: bricksetup.h:139-221 (8x8x8, fold 8x8)     */
/* ---------------------------------------------------------------------- */
/* array extents dims[a] + 2*(padding[a] + ghost[a]); brick grid extents
 * (dims[a] + 2*ghost[a]) / 8; bricks s in [ghost/8, (dims+ghost)/8). */
static void copy_grid(int to_brick, int dt, const long dims[3],
                      const long padding[3], const long ghost[3], void *arr,
                      const uint32_t *grid, void *dat, size_t step) {
  long sa[3], sb[3];
  long za = 1, zb = 1;
  for (int a = 0; a < 3; ++a) {
    sa[a] = za;
    sb[a] = zb;
    za *= dims[a] + 2 * (padding[a] + ghost[a]);
    zb *= (dims[a] + 2 * ghost[a]) / 8;
  }
  const size_t es = dt_size(dt);
  const long k0 = ghost[2] / 8, k1 = (dims[2] + ghost[2]) / 8;
  const long j0 = ghost[1] / 8, j1 = (dims[1] + ghost[1]) / 8;
  const long i0 = ghost[0] / 8, i1 = (dims[0] + ghost[0]) / 8;
#pragma omp parallel for collapse(2)
  for (long tk = k0; tk < k1; ++tk)
    for (long tj = j0; tj < j1; ++tj)
      for (long ti = i0; ti < i1; ++ti) {
        const size_t b = grid[ti * sb[0] + tj * sb[1] + tk * sb[2]];
        for (long k = 0; k < 8; ++k)
          for (long j = 0; j < 8; ++j)
            for (long i = 0; i < 8; ++i) {
              const size_t ai = (size_t)((padding[0] + ti * 8 + i) * sa[0] +
                                         (padding[1] + tj * 8 + j) * sa[1] +
                                         (padding[2] + tk * 8 + k) * sa[2]);
              const size_t bi = b * step + (size_t)(k * 64 + j * 8 + i);
              if (to_brick)
                memcpy((char *)dat + bi * es, (char *)arr + ai * es, es);
              else
                memcpy((char *)arr + ai * es, (char *)dat + bi * es, es);
            }
      }
}

void or_copy_to_brick(int dt, const long dims[3], const long padding[3],
                      const long ghost[3], const void *arr,
                      const uint32_t *grid, void *dat, size_t step) {
  copy_grid(1, dt, dims, padding, ghost, (void *)arr, grid, dat, step);
}

void or_copy_from_brick(int dt, const long dims[3], const long padding[3],
                        const long ghost[3], void *arr, const uint32_t *grid,
                        const void *dat, size_t step) {
  copy_grid(0, dt, dims, padding, ghost, arr, grid, (void *)dat, step);
}

/* ---------------------------------------------------------------------- */
/* stencil tap tables                                                      */
/* ---------------------------------------------------------------------- */
typedef struct {
  int di, dj, dk;
} tap_t;

int or_stencil_radius(int st) {
  switch (st) {
    case OR_STAR7: case OR_COND7: return 1;
    case OR_STAR13: return 2;
    case OR_STAR25: return 4;
    case OR_CUBE125: return 2;
    default: return 0;
  }
}

int or_stencil_ncoeff(int st) {
  switch (st) {
    case OR_STAR7: case OR_COND7: return 7;
    case OR_STAR13: return 13;
    case OR_STAR25: return 25;
    case OR_CUBE125: return 125;
    default: return 0;
  }
}

/* taps in coefficient order: tap t uses coefficient t */
static int stencil_taps(int st, tap_t *t) {
  int n = 0;
  if (st == OR_CUBE125) {
    for (int dk = -2; dk <= 2; ++dk)
      for (int dj = -2; dj <= 2; ++dj)
        for (int di = -2; di <= 2; ++di) {
          t[n].di = di; t[n].dj = dj; t[n].dk = dk; ++n;
        }
    return n;
  }
  const int R = or_stencil_radius(st);
  t[n].di = 0; t[n].dj = 0; t[n].dk = 0; ++n;
  for (int r = 1; r <= R; ++r) {
    const int d[6][3] = {{r, 0, 0}, {-r, 0, 0}, {0, r, 0}, {0, -r, 0}, {0, 0, r}, {0, 0, -r}};
    for (int q = 0; q < 6; ++q) {
      t[n].di = d[q][0]; t[n].dj = d[q][1]; t[n].dk = d[q][2]; ++n;
    }
  }
  return n;
}

/* evaluation order (indices into the tap table) for a mode */
static int stencil_order(int st, int mode, int *ord) {
  const int n = or_stencil_ncoeff(st);
  if (mode == OR_MODE_REF && (st == OR_STAR7 || st == OR_COND7)) {
    /* 3axis.cpp:60-63 and :132-137: coeff[5] k+1, coeff[6] k-1, coeff[3]
     * j+1, coeff[4] j-1, coeff[1] i+1, coeff[2] i-1, coeff[0] centre */
    const int o[7] = {5, 6, 3, 4, 1, 2, 0};
    for (int q = 0; q < 7; ++q) ord[q] = o[q];
    return 7;
  }
  for (int q = 0; q < n; ++q) ord[q] = q;
  return n;
}

/* ---------------------------------------------------------------------- */
/* per-point evaluation from a gathered (8+2R)^3 box                       */
/* ---------------------------------------------------------------------- */
#define OR_BOXMAX 16

static inline double relu_ref(double x) { return x < 0.0 ? 0.0 : x; } /* std::max(x, 0.0) */
static inline float relu_reff(float x) { return x < 0.0f ? 0.0f : x; }

typedef struct {
  int st, mode, cond_mode, ntap;
  int ord[125];
  int boff[125]; /* tap offset inside the box relative to the point */
  double c[125];
  float cf[125];
  double theta;
  float thetaf;
  int R, P; /* radius, box edge */
} plan_t;

static void make_plan(plan_t *pl, int st, int mode, const double *coeff,
                      double theta, int cond_mode) {
  tap_t taps[125];
  pl->st = st;
  pl->mode = mode;
  pl->cond_mode = cond_mode;
  pl->R = or_stencil_radius(st);
  pl->P = 8 + 2 * pl->R;
  pl->ntap = stencil_taps(st, taps);
  stencil_order(st, mode, pl->ord);
  for (int q = 0; q < pl->ntap; ++q) {
    pl->c[q] = coeff[q];
    pl->cf[q] = (float)coeff[q];
    pl->boff[q] = (taps[q].dk * pl->P + taps[q].dj) * pl->P + taps[q].di;
  }
  pl->theta = theta;
  pl->thetaf = (float)theta;
}

/* gather the (8+2R)^3 neighbourhood of brick b, as double or float */
static void gather_f64(const plan_t *pl, const uint32_t *adj, uint32_t b,
                       int dt, const void *in, size_t step, double *box) {
  const int R = pl->R, P = pl->P;
  const uint32_t *ab = adj + (size_t)b * 27;
  for (int k = -R; k < 8 + R; ++k)
    for (int j = -R; j < 8 + R; ++j)
      for (int i = -R; i < 8 + R; ++i) {
        const int kd = (k + 8) >> 3, jd = (j + 8) >> 3, id = (i + 8) >> 3;
        const size_t e = (size_t)(((k + 8) & 7) * 64 + ((j + 8) & 7) * 8 + ((i + 8) & 7));
        box[((k + R) * P + (j + R)) * P + (i + R)] =
            ld_as_f64(dt, in, (size_t)ab[kd * 9 + jd * 3 + id] * step + e);
      }
}

static inline double eval_f64(const plan_t *pl, const double *x) {
  const int *o = pl->ord;
  const int cond = pl->st == OR_COND7;
  double acc;
  if (pl->mode == OR_MODE_REF) {
    acc = pl->c[o[0]] * (cond ? relu_ref(x[pl->boff[o[0]]]) : x[pl->boff[o[0]]]);
    for (int q = 1; q < pl->ntap; ++q) {
      const double v = cond ? relu_ref(x[pl->boff[o[q]]]) : x[pl->boff[o[q]]];
      const double p = pl->c[o[q]] * v;
      acc = acc + p;
    }
  } else {
    acc = pl->c[o[0]] * (cond ? relu_ref(x[pl->boff[o[0]]]) : x[pl->boff[o[0]]]);
    for (int q = 1; q < pl->ntap; ++q) {
      const double v = cond ? relu_ref(x[pl->boff[o[q]]]) : x[pl->boff[o[q]]];
      acc = fma(pl->c[o[q]], v, acc);
    }
  }
  if (cond) {
    const int sel = pl->cond_mode == 0 ? (acc > 0.0) : (x[0] > pl->theta);
    acc = sel ? acc : -acc;
  }
  return acc;
}

static inline float eval_f32(const plan_t *pl, const float *x) {
  const int *o = pl->ord;
  const int cond = pl->st == OR_COND7;
  float acc = pl->cf[o[0]] * (cond ? relu_reff(x[pl->boff[o[0]]]) : x[pl->boff[o[0]]]);
  for (int q = 1; q < pl->ntap; ++q) {
    const float v = cond ? relu_reff(x[pl->boff[o[q]]]) : x[pl->boff[o[q]]];
    if (pl->mode == OR_MODE_REF) {
      const float p = pl->cf[o[q]] * v;
      acc = acc + p;
    } else {
      acc = fmaf(pl->cf[o[q]], v, acc);
    }
  }
  if (cond) {
    const int sel = pl->cond_mode == 0 ? (acc > 0.0f) : (x[0] > pl->thetaf);
    acc = sel ? acc : -acc;
  }
  return acc;
}

/* ---------------------------------------------------------------------- */
/* brick sweep: 3axis.cpp:51-66 over a brick list                          */
/* ---------------------------------------------------------------------- */
/* FP64 computes in double; FP32 and BF16 compute in float (BF16 values are
 * widened exactly, results rounded to nearest even on store). */
void or_brick_sweep(int st, int dt, int mode, const uint32_t *adj,
                    const uint32_t *list, uint32_t nlist, const void *in,
                    size_t in_step, void *out, size_t out_step,
                    const double *coeff, double theta, int cond_mode) {
  plan_t pl;
  make_plan(&pl, st, mode, coeff, theta, cond_mode);
  const int R = pl.R, P = pl.P;
#pragma omp parallel
  {
    double box[OR_BOXMAX * OR_BOXMAX * OR_BOXMAX];
    float boxf[OR_BOXMAX * OR_BOXMAX * OR_BOXMAX];
#pragma omp for schedule(static)
    for (long l = 0; l < (long)nlist; ++l) {
      const uint32_t b = list[l];
      gather_f64(&pl, adj, b, dt, in, in_step, box);
      if (dt != OR_F64)
        for (int q = 0; q < P * P * P; ++q) boxf[q] = (float)box[q]; /* exact */
      for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 8; ++j)
          for (int i = 0; i < 8; ++i) {
            const int c = ((k + R) * P + (j + R)) * P + (i + R);
            const size_t oi = (size_t)b * out_step + (size_t)(k * 64 + j * 8 + i);
            if (dt == OR_F64) {
              ((double *)out)[oi] = eval_f64(&pl, box + c);
            } else {
              const float v = eval_f32(&pl, boxf + c);
              if (dt == OR_F32) ((float *)out)[oi] = v;
              else ((uint16_t *)out)[oi] = f32_to_bf16(v);
            }
          }
    }
  }
}

/* ---------------------------------------------------------------------- */
/* other brick shapes and vector folds: Brick<Dim<BK,BJ,BI>,Dim<FJ,FI>>    */
/* ---------------------------------------------------------------------- */
/* Layouts (the library's bk_layout):
 *   0  Dim<8,8,8>,  Dim<8,8>   element k*64 + j*8 + i
 *   1  Dim<4,8,16>, Dim<4,16>  element k*128 + j*16 + i (row-major)
 *   2  Dim<4,8,16>, Dim<8,8>   element k*128 + (i/8)*64 + j*8 + i%8
 * The element offsets follow the reference accessor's fold decomposition
 * (brick.h:276-284 with the vector index of brick.h:235-246): the vector
 * index runs over the brick divided by the fold, outermost first, and the
 * lane over the fold itself.  For fold (FJ, FI) of a (BK, BJ, BI) brick:
 * vector ((k*(BJ/FJ) + j/FJ)*(BI/FI) + i/FI), lane (j%FJ)*FI + i%FI. */
#define OR_LAYOUTS 3
static const int lay_ext[OR_LAYOUTS][3] = {{8, 8, 8}, {4, 8, 16}, {4, 8, 16}}; /* k, j, i */
static const int lay_fold[OR_LAYOUTS][2] = {{8, 8}, {4, 16}, {8, 8}};        /* j, i */

int or_layout_extent(int lay, int axis) { /* axis 0 = k, 1 = j, 2 = i */
  if (lay < 0 || lay >= OR_LAYOUTS || axis < 0 || axis > 2) return 0;
  return lay_ext[lay][axis];
}

static inline int lay_off(int lay, int k, int j, int i) {
  const int BJ = lay_ext[lay][1], BI = lay_ext[lay][2];
  const int FJ = lay_fold[lay][0], FI = lay_fold[lay][1];
  const int vec = (k * (BJ / FJ) + j / FJ) * (BI / FI) + i / FI;
  return vec * (FJ * FI) + (j % FJ) * FI + i % FI;
}

int or_layout_offset(int lay, int k, int j, int i) {
  if (lay < 0 || lay >= OR_LAYOUTS) return -1;
  return lay_off(lay, k, j, i);
}

/* seeded fill of a layout's brick grid: nb interior bricks per axis
 * (contiguous-first), `ghost` (0 or 1) ghost layers; cell (x, y, z) of the
 * n = nb * extent interior gets or_u01(seed, x + n0*(y + n1*z)), the value
 * or_fill_uniform gives the same cell. */
void or_fill_uniform_layout(int lay, int dt, void *dat, size_t step, const long nb[3],
                            int ghost, uint64_t seed, double lo, double hi) {
  const int BK = lay_ext[lay][0], BJ = lay_ext[lay][1], BI = lay_ext[lay][2];
  const long n0 = nb[0] * BI, n1 = nb[1] * BJ, n2 = nb[2] * BK;
  const long S0 = nb[0] + 2 * ghost, S1 = nb[1] + 2 * ghost;
#pragma omp parallel for collapse(2)
  for (long z = 0; z < n2; ++z)
    for (long y = 0; y < n1; ++y)
      for (long x = 0; x < n0; ++x) {
        uint64_t lin = (uint64_t)x + (uint64_t)n0 * ((uint64_t)y + (uint64_t)n1 * (uint64_t)z);
        double v = lo + (hi - lo) * or_u01(seed, lin);
        size_t b = (size_t)((x / BI + ghost) + S0 * ((y / BJ + ghost) + S1 * (z / BK + ghost)));
        st_from_f64(dt, dat, b * step + (size_t)lay_off(lay, (int)(z % BK), (int)(y % BJ),
                                                        (int)(x % BI)), v);
      }
}

/* the interior of a layout's brick grid (as above) to a dense FP64 array
 * (n2, n1, n0), x fastest */
void or_to_dense_layout(int lay, int dt, const void *dat, size_t step, const long nb[3],
                        int ghost, double *arr) {
  const int BK = lay_ext[lay][0], BJ = lay_ext[lay][1], BI = lay_ext[lay][2];
  const long n0 = nb[0] * BI, n1 = nb[1] * BJ, n2 = nb[2] * BK;
  const long S0 = nb[0] + 2 * ghost, S1 = nb[1] + 2 * ghost;
#pragma omp parallel for collapse(2)
  for (long z = 0; z < n2; ++z)
    for (long y = 0; y < n1; ++y)
      for (long x = 0; x < n0; ++x) {
        size_t b = (size_t)((x / BI + ghost) + S0 * ((y / BJ + ghost) + S1 * (z / BK + ghost)));
        arr[(z * n1 + y) * n0 + x] =
            ld_as_f64(dt, dat, b * step + (size_t)lay_off(lay, (int)(z % BK), (int)(y % BJ),
                                                         (int)(x % BI)));
      }
}

#define OR_LBOXMAX (12 * 16 * 24) /* largest (BK+2R)(BJ+2R)(BI+2R): R = 4 on 4x8x16 */

/* or_brick_sweep for any layout: the same per-point evaluation (and so the
 * same bits per cell), the neighbourhood gathered through the adjacency
 * with the layout's extents (neighbour digit (x + B) / B per axis, as the
 * accessor's, brick.h:235-246) and element offsets. */
void or_brick_sweep_layout(int lay, int st, int dt, int mode, const uint32_t *adj,
                           const uint32_t *list, uint32_t nlist, const void *in,
                           size_t in_step, void *out, size_t out_step,
                           const double *coeff, double theta, int cond_mode) {
  plan_t pl;
  make_plan(&pl, st, mode, coeff, theta, cond_mode);
  const int R = pl.R, P = pl.P;
  const int BK = lay_ext[lay][0], BJ = lay_ext[lay][1], BI = lay_ext[lay][2];
  const int Pj = BJ + 2 * R, Pi = BI + 2 * R, Pk = BK + 2 * R;
  /* re-base the tap offsets on this box */
  for (int q = 0; q < pl.ntap; ++q) {
    const int bo = pl.boff[q] + (R * P + R) * P + R;
    const int di = bo % P - R, dj = (bo / P) % P - R, dk = bo / (P * P) - R;
    pl.boff[q] = (dk * Pj + dj) * Pi + di;
  }
#pragma omp parallel
  {
    double box[OR_LBOXMAX];
    float boxf[OR_LBOXMAX];
#pragma omp for schedule(static)
    for (long l = 0; l < (long)nlist; ++l) {
      const uint32_t b = list[l];
      const uint32_t *ab = adj + (size_t)b * 27;
      for (int k = -R; k < BK + R; ++k)
        for (int j = -R; j < BJ + R; ++j)
          for (int i = -R; i < BI + R; ++i) {
            const int kd = (k + BK) / BK, jd = (j + BJ) / BJ, id = (i + BI) / BI;
            const int e = lay_off(lay, (k + BK) % BK, (j + BJ) % BJ, (i + BI) % BI);
            box[((k + R) * Pj + (j + R)) * Pi + (i + R)] =
                ld_as_f64(dt, in, (size_t)ab[kd * 9 + jd * 3 + id] * in_step + (size_t)e);
          }
      if (dt != OR_F64)
        for (int q = 0; q < Pk * Pj * Pi; ++q) boxf[q] = (float)box[q]; /* exact */
      for (int k = 0; k < BK; ++k)
        for (int j = 0; j < BJ; ++j)
          for (int i = 0; i < BI; ++i) {
            const int c = ((k + R) * Pj + (j + R)) * Pi + (i + R);
            const size_t oi = (size_t)b * out_step + (size_t)lay_off(lay, k, j, i);
            if (dt == OR_F64) {
              ((double *)out)[oi] = eval_f64(&pl, box + c);
            } else {
              const float v = eval_f32(&pl, boxf + c);
              if (dt == OR_F32) ((float *)out)[oi] = v;
              else ((uint16_t *)out)[oi] = f32_to_bf16(v);
            }
          }
    }
  }
}

/* The reference's accessor loop for the 7-point star, restated literally:
 * stencils/3axis.cpp:51-66, every access through the adjacency list
 * (brick.h:235-246).  Used as the CPU baseline and to cross-check the
 * gathered-box evaluation above. */
void or_brick_sweep_7pt_accessor(const uint32_t *adj, const uint32_t *list,
                                 uint32_t nlist, const double *in,
                                 size_t in_step, double *out, size_t out_step,
                                 const double *coeff) {
#define ACC(k, j, i)                                                           \
  in[(size_t)ab[(((k) + 8) >> 3) * 9 + (((j) + 8) >> 3) * 3 + (((i) + 8) >> 3)] * in_step + \
     (size_t)((((k) + 8) & 7) * 64 + (((j) + 8) & 7) * 8 + (((i) + 8) & 7))]
#pragma omp parallel for schedule(static)
  for (long l = 0; l < (long)nlist; ++l) {
    const uint32_t b = list[l];
    const uint32_t *ab = adj + (size_t)b * 27;
    for (long k = 0; k < 8; ++k)
      for (long j = 0; j < 8; ++j)
        for (long i = 0; i < 8; ++i)
          out[(size_t)b * out_step + (size_t)(k * 64 + j * 8 + i)] =
              coeff[5] * ACC(k + 1, j, i) + coeff[6] * ACC(k - 1, j, i) +
              coeff[3] * ACC(k, j + 1, i) + coeff[4] * ACC(k, j - 1, i) +
              coeff[1] * ACC(k, j, i + 1) + coeff[2] * ACC(k, j, i - 1) +
              coeff[0] * ACC(k, j, i);
  }
#undef ACC
}

/* ---------------------------------------------------------------------- */
/* array loop: 3axis.cpp:32-37 on a dense array with a ghost zone          */
/* ---------------------------------------------------------------------- */
/* arr extents n[a] + 2*g per axis (x fastest); points [g, g+n) computed. */
void or_array_sweep(int st, int mode, const long n[3], long g,
                    const double *in, double *out, const double *coeff,
                    double theta, int cond_mode) {
  plan_t pl;
  make_plan(&pl, st, mode, coeff, theta, cond_mode);
  const long e0 = n[0] + 2 * g, e1 = n[1] + 2 * g;
  const int P = pl.P;
  int aoff[125];
  for (int q = 0; q < pl.ntap; ++q) {
    /* decode the box offset back into (dk, dj, di) */
    int bo = pl.boff[q] + (pl.R * P + pl.R) * P + pl.R;
    int di = bo % P - pl.R, dj = (bo / P) % P - pl.R, dk = bo / (P * P) - pl.R;
    aoff[q] = (int)((dk * e1 + dj) * e0 + di);
  }
#pragma omp parallel
  {
    double x[OR_BOXMAX * OR_BOXMAX * OR_BOXMAX];
#pragma omp for collapse(2) schedule(static)
    for (long z = g; z < g + n[2]; ++z)
      for (long y = g; y < g + n[1]; ++y)
        for (long xi = g; xi < g + n[0]; ++xi) {
          const size_t c = (size_t)((z * e1 + y) * e0 + xi);
          /* place the taps where eval_f64 expects them in a box */
          const int R = pl.R;
          double *xc = x + ((R * P + R) * P + R);
          for (int q = 0; q < pl.ntap; ++q) xc[pl.boff[q]] = in[c + aoff[q]];
          out[c] = eval_f64(&pl, xc);
        }
  }
}

/* periodic ghost fill of a dense array: ghost points take x mod n */
void or_array_periodic_fill(double *arr, const long n[3], long g) {
  const long e0 = n[0] + 2 * g, e1 = n[1] + 2 * g, e2 = n[2] + 2 * g;
#pragma omp parallel for collapse(2)
  for (long z = 0; z < e2; ++z)
    for (long y = 0; y < e1; ++y)
      for (long x = 0; x < e0; ++x) {
        const int gz = z < g || z >= g + n[2], gy = y < g || y >= g + n[1],
                  gx = x < g || x >= g + n[0];
        if (!(gx || gy || gz)) continue;
        const long sx = ((x - g) % n[0] + n[0]) % n[0] + g;
        const long sy = ((y - g) % n[1] + n[1]) % n[1] + g;
        const long sz = ((z - g) % n[2] + n[2]) % n[2] + g;
        arr[(z * e1 + y) * e0 + x] = arr[(sz * e1 + sy) * e0 + sx];
      }
}

/* ---------------------------------------------------------------------- */
/* comparison                                                              */
/* ---------------------------------------------------------------------- */
/* compareBrick's rule (brickcompare.h:35-39):
 *   pass iff diff < tol || diff < (|a| + |b|) * tol   for every point.
 * Returns 1 when all points of the brick region pass.  Region as
 * copyFromBrick. */
int or_compare_brick(const long dims[3], const long padding[3],
                     const long ghost[3], const double *arr,
                     const uint32_t *grid, const double *dat, size_t step,
                     double tol) {
  long sa[3], sb[3], za = 1, zb = 1;
  for (int a = 0; a < 3; ++a) {
    sa[a] = za;
    sb[a] = zb;
    za *= dims[a] + 2 * (padding[a] + ghost[a]);
    zb *= (dims[a] + 2 * ghost[a]) / 8;
  }
  int ok = 1;
  const long k0 = ghost[2] / 8, k1 = (dims[2] + ghost[2]) / 8;
  const long j0 = ghost[1] / 8, j1 = (dims[1] + ghost[1]) / 8;
  const long i0 = ghost[0] / 8, i1 = (dims[0] + ghost[0]) / 8;
#pragma omp parallel for collapse(2) reduction(&& : ok)
  for (long tk = k0; tk < k1; ++tk)
    for (long tj = j0; tj < j1; ++tj)
      for (long ti = i0; ti < i1; ++ti) {
        const size_t b = grid[ti * sb[0] + tj * sb[1] + tk * sb[2]];
        for (long k = 0; k < 8; ++k)
          for (long j = 0; j < 8; ++j)
            for (long i = 0; i < 8; ++i) {
              const double a = arr[(padding[0] + ti * 8 + i) * sa[0] +
                                   (padding[1] + tj * 8 + j) * sa[1] +
                                   (padding[2] + tk * 8 + k) * sa[2]];
              const double v = dat[b * step + (size_t)(k * 64 + j * 8 + i)];
              const double diff = fabs(v - a);
              const int r = (diff < tol) || (diff < (fabs(v) + fabs(a)) * tol);
              ok = ok && r;
            }
      }
  return ok;
}

/* error statistics of field a against reference r over a brick list:
 * out = {max|a-r|, max |a-r|/max(|r|,floor), sum (a-r)^2, sum r^2,
 *        count(|a-r| > atol + rtol*|r|)} */
void or_field_compare(int dta, const void *a, size_t step_a, int dtr,
                      const void *r, size_t step_r, const uint32_t *list,
                      uint32_t nlist, size_t brick_elems, double floor_,
                      double atol, double rtol, double out[5]) {
  double mabs = 0, mrel = 0, s2 = 0, r2 = 0, bad = 0;
#pragma omp parallel for reduction(max : mabs, mrel) reduction(+ : s2, r2, bad)
  for (long l = 0; l < (long)nlist; ++l) {
    const size_t b = list[l];
    for (size_t e = 0; e < brick_elems; ++e) {
      const double va = ld_as_f64(dta, a, b * step_a + e);
      const double vr = ld_as_f64(dtr, r, b * step_r + e);
      const double d = fabs(va - vr);
      const double den = fabs(vr) > floor_ ? fabs(vr) : floor_;
      if (d > mabs) mabs = d;
      if (d / den > mrel) mrel = d / den;
      s2 += d * d;
      r2 += vr * vr;
      if (d > atol + rtol * fabs(vr)) bad += 1;
    }
  }
  out[0] = mabs; out[1] = mrel; out[2] = s2; out[3] = r2; out[4] = bad;
}

int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_num_threads(int n) { if (n > 0) omp_set_num_threads(n); }
double or_wtime(void) { return omp_get_wtime(); }
