// bk_cube.hip -- 125-point (5x5x5, radius 2) cube stencil sweep over 8x8x8
// bricks in the 8x8 fold, for gfx950.
//
// Replaces the codegen brick kernel of weak/main.hip.cpp:32-40 built from
// stencils/mpi125pt.py:12-162 (here one coefficient per offset).
//
// One wave64 per brick, four bricks per workgroup.  The wave gathers the
// brick's 12x12x12 neighbourhood (its own 512 points plus the 2-deep shell
// from all 26 neighbours, found through the adjacency table) into LDS, then
// lane L = j*8 + i accumulates the eight outputs of column (j, i): for each
// input plane kk it reads the 5x5 window once and feeds it to every output
// plane within reach, so each LDS value serves up to five FMAs.  Per point
// the taps accumulate in (dk, dj, di) lexicographic order, starting with a
// multiply -- the coefficient order of the API.
#include <cstddef>

#include "bk_device.h"
#include "bk_internal.h"

namespace bk {
namespace {

template <class T>
struct CubeArgs {
  using A = typename acc_of<T>::type;
  const uint32_t *__restrict__ adj;
  const uint32_t *__restrict__ list;
  const T *__restrict__ in;
  T *__restrict__ out;
  uint64_t in_step;
  uint64_t out_step;
  uint32_t nlist;
  // plan (quad kernel only)
  const uint32_t *__restrict__ anchor;
  const uint32_t *__restrict__ len;
  const uint32_t *__restrict__ mask;
  uint32_t nunits;
  uint32_t kcol;
  A c[125];
};

template <class T>
__global__ __launch_bounds__(256) void cube_sweep_lds(const CubeArgs<T> a) {
  using A = typename acc_of<T>::type;
  constexpr int R = 2, P = 8 + 2 * R;  // 12
  __shared__ A box[4][P][P][P];
  __shared__ uint32_t sadj[4][32];

  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int j = lane >> 3, i = lane & 7;
  const uint32_t slot = xcd_remap(blockIdx.x, gridDim.x) * 4u + (uint32_t)w;
  const bool active = slot < a.nlist;

  uint32_t b = 0;
  if (active) {
    b = uniform(a.list[slot]);
    if (lane < 27) sadj[w][lane] = a.adj[(size_t)b * 27 + lane];
  }
  __syncthreads();
  if (active) {
    const size_t st = a.in_step;
#pragma unroll
    for (int t = 0; t < 27; ++t) {  // 27 * 64 == 12^3
      const int q = t * 64 + lane;
      const int kk = q / (P * P), rem = q - kk * P * P;
      const int jj = rem / P, ii = rem - jj * P;
      const int k = kk - R + 8, jn = jj - R + 8, in_ = ii - R + 8;  // shifted by +8
      const int s = (k >> 3) * 9 + (jn >> 3) * 3 + (in_ >> 3);
      const int e = (k & 7) * 64 + (jn & 7) * 8 + (in_ & 7);
      box[w][kk][jj][ii] = widen(a.in[(size_t)sadj[w][s] * st + e]);
    }
  }
  __syncthreads();
  if (!active) return;

  A acc[8];
#pragma unroll
  for (int kk = 0; kk < P; ++kk) {
#pragma unroll
    for (int dj = 0; dj < 5; ++dj) {
#pragma unroll
      for (int di = 0; di < 5; ++di) {
        const A x = box[w][kk][j + dj][i + di];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int dk = kk - k;  // 0..4 when plane kk reaches output k
          if (dk < 0 || dk > 4) continue;
          const A c = a.c[dk * 25 + dj * 5 + di];
          if (dk == 0 && dj == 0 && di == 0) acc[k] = c * x;
          else acc[k] = fmadd(c, x, acc[k]);
        }
      }
    }
  }
  T *po = a.out + (size_t)b * a.out_step;
#pragma unroll
  for (int k = 0; k < 8; ++k) narrow(po[k * 64 + lane], acc[k]);
}


// Plane-streaming variant over 2x2 column quads (plan shape BK_PLAN_QUAD).
//
// One wave per plan unit: a 2x2 block of brick columns (members m0, m1 =
// i+1, m2 = j+1, m3 = i+1,j+1) over up to kcol levels.  The wave streams the
// input planes P = -2 .. 8L+1 of the 16x16 quad plane, each padded by the
// 2-deep halo to 20x20 and gathered through the adjacency into one of two
// LDS buffers (the next plane loads while the current one computes).  Lane
// (lj, li) owns the 2x2 points (2lj+pj, 2li+pi) of the quad plane and keeps
// five rolling accumulator slots per point, for output planes P-2 .. P+2:
// every 6x6 window it reads from LDS feeds up to 25 taps x 5 planes x 4
// points.  Per output the taps accumulate in (dk, dj, di) lexicographic
// order, the first as a multiply -- the same arithmetic as the per-brick
// kernel, so results are identical.
// The padded plane is 20 x 20 cells, stored with a row stride of 24 so the
// two rows a half-wave's LDS reads touch (2 apart) sit 32 banks apart in
// both FP64 (16-byte reads) and FP32 (8-byte reads): with stride 20 they
// overlapped and every read took a bank-conflict pass.
constexpr int kQuadPW = 20, kQuadStride = 24, kQuadCells = kQuadPW * kQuadStride;

// Two adjacent elements read as one load (16 bytes in FP64, 8 in FP32, 4 in
// BF16); only ever at an even element index of a brick row.
template <class T>
struct alignas(2 * sizeof(T)) Pair {
  T a, b;
};

// One input plane P of the quad unit: accumulate its taps into the five
// rolling slots (slot s holds output plane P - 2 + s, tap dk = 2 - s).
// Taps go row by row of the LDS plane (r), then slot by slot; per output
// point that is (dk, dj, di) lexicographic, the first tap a multiply.
//
// Edge planes (some slots outside the unit's output range): coefficients
// are read group by group (row r, slot s: up to ten values) through a
// constant-address pointer, so they arrive by scalar loads into SGPRs; the
// empty asm at each group hides the pointer's loop invariance so the
// compiler does not try to keep all 125 (250 SGPRs) live across planes.
template <class A>
__device__ __forceinline__ void quad_plane_edge(const A *pl, int lj, int li, kconst_ptr<A> cp,
                                                int P, int nout, A (&acc)[5][2][2]) {
  constexpr int PW = kQuadStride;
  int ro = (2 * lj) * PW + 2 * li;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    A x[6];
    if (r > 0) {
      // row r's reads wait for row r-1's FMAs (every accumulator is an
      // input): left alone, the scheduler issues all six rows' reads up
      // front, and their 72 live registers cost the FP64 kernel a wave
      ro += PW;
      asm volatile(""
                   : "+v"(ro)
                   : "v"(acc[0][0][0]), "v"(acc[0][0][1]), "v"(acc[0][1][0]), "v"(acc[0][1][1]),
                     "v"(acc[1][0][0]), "v"(acc[1][0][1]), "v"(acc[1][1][0]), "v"(acc[1][1][1]),
                     "v"(acc[2][0][0]), "v"(acc[2][0][1]), "v"(acc[2][1][0]), "v"(acc[2][1][1]),
                     "v"(acc[3][0][0]), "v"(acc[3][0][1]), "v"(acc[3][1][0]), "v"(acc[3][1][1]),
                     "v"(acc[4][0][0]), "v"(acc[4][0][1]), "v"(acc[4][1][0]), "v"(acc[4][1][1]));
    }
    // 16-byte LDS reads of the lane's two adjacent columns: 8 lanes of a
    // row cover 128 contiguous bytes, so a read is free of bank conflicts
#pragma unroll
    for (int c = 0; c < 6; c += 2) {
      typedef A A2 __attribute__((ext_vector_type(2)));
      const A2 v = *reinterpret_cast<const A2 *>(&pl[ro + c]);
      x[c] = v.x;
      x[c + 1] = v.y;
    }
#pragma unroll
    for (int s = 0; s < 5; ++s) {
      asm volatile("" : "+s"(cp));
      const int O = P - 2 + s;
      if (O < 0 || O >= nout) continue;
      const int dk = 2 - s;
#pragma unroll
      for (int pj = 0; pj < 2; ++pj) {
        const int dj = r - pj - 2;
        if (dj < -2 || dj > 2) continue;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
#pragma unroll
          for (int pi = 0; pi < 2; ++pi) {
            const int di = c - pi - 2;
            if (di < -2 || di > 2) continue;
            const A cf = cp[(dk + 2) * 25 + (dj + 2) * 5 + (di + 2)];
            if (s == 4 && dj == -2 && di == -2) acc[s][pj][pi] = cf * x[c];
            else acc[s][pj][pi] = fmadd(cf, x[c], acc[s][pj][pi]);
          }
        }
      }
    }
  }
}

// Coefficient group (r, s) of a plane: the rows dj = r-3 .. r-2 that exist
// (one row for r = 0 and r = 5) of tap plane dk = 2 - s -- up to ten
// consecutive values of the (dk, dj, di)-ordered table.
__device__ constexpr int quad_group_first_row(int r) { return r > 0 ? r - 1 : 0; }
__device__ constexpr int quad_group_count(int r) { return (r == 0 || r == 5) ? 5 : 10; }

// Steady-state planes (all five slots inside the unit): the same taps in
// the same per-point order, with the scalar loads of group g+1's
// coefficients and the LDS reads of row r+1 issued before group g's FMAs.
// Scalar loads and LDS reads share one wait counter and scalar loads
// complete out of order, so each group starts with one full wait.  The
// empty asm statements pin the schedule: a "use" of the group's registers
// puts the wait before the next group's loads; the launder of the pointer
// (and of the LDS row offset) takes an accumulator of the previous group
// as input, so the next loads cannot be hoisted above that group's FMAs
// (which would keep every group's coefficients live at once).
template <int G, class A>
__device__ __forceinline__ void quad_steady_group(const A *pl, int xo, kconst_ptr<A> &cp,
                                                  A (&cf)[2][10], A (&x)[2][6],
                                                  A (&acc)[5][2][2]) {
  constexpr int PW = kQuadStride;
  typedef A A2 __attribute__((ext_vector_type(2)));
  constexpr int r = G / 5, s = G % 5, cur = G & 1;
  // an accumulator the previous group (rp, sp) wrote
  constexpr int rp = (G + 29) / 5 % 6, sp = (G + 29) % 5, pjp = rp <= 4 ? 0 : 1;
  // wait here for this group's coefficients (and, at a row start, the row)
#pragma unroll
  for (int q = 0; q < quad_group_count(r); ++q) asm volatile("" ::"s"(cf[cur][q]));
  if constexpr (s == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) asm volatile("" ::"v"(x[r & 1][c]));
  }
  // issue the next group's scalar loads and, at a row start, the next row
  if constexpr (G + 1 < 30) {
    constexpr int rn = (G + 1) / 5, sn = (G + 1) % 5;
    if constexpr (G > 0) asm volatile("" : "+s"(cp) : "v"(acc[sp][pjp][0]));
    else asm volatile("" : "+s"(cp));
    constexpr int off = (4 - sn) * 25 + quad_group_first_row(rn) * 5;
#pragma unroll
    for (int q = 0; q < quad_group_count(rn); ++q) cf[cur ^ 1][q] = cp[off + q];
  }
  if constexpr (s == 0 && r + 1 < 6) {
    int ro = xo + (r + 1) * PW;
    if constexpr (G > 0) asm volatile("" : "+v"(ro) : "v"(acc[sp][pjp][0]));
    else asm volatile("" : "+v"(ro));
#pragma unroll
    for (int c = 0; c < 6; c += 2) {
      const A2 v = *reinterpret_cast<const A2 *>(&pl[ro + c]);
      x[(r + 1) & 1][c] = v.x;
      x[(r + 1) & 1][c + 1] = v.y;
    }
  }
  // this group's FMAs
  constexpr int row0 = quad_group_first_row(r);
#pragma unroll
  for (int pj = 0; pj < 2; ++pj) {
    const int dj = r - pj - 2;
    if (dj < -2 || dj > 2) continue;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int pi = 0; pi < 2; ++pi) {
        const int di = c - pi - 2;
        if (di < -2 || di > 2) continue;
        const A k = cf[cur][(dj + 2 - row0) * 5 + (di + 2)];
        if (s == 4 && dj == -2 && di == -2) acc[s][pj][pi] = k * x[r & 1][c];
        else acc[s][pj][pi] = fmadd(k, x[r & 1][c], acc[s][pj][pi]);
      }
    }
  }
  if constexpr (G + 1 < 30) quad_steady_group<G + 1, A>(pl, xo, cp, cf, x, acc);
}

template <class A>
__device__ __forceinline__ void quad_plane_steady(const A *pl, int lj, int li,
                                                  kconst_ptr<A> cp, A (&acc)[5][2][2]) {
  typedef A A2 __attribute__((ext_vector_type(2)));
  A cf[2][10];
  A x[2][6];
  const int xo = (2 * lj) * kQuadStride + 2 * li;
  asm volatile("" : "+s"(cp));
#pragma unroll
  for (int q = 0; q < quad_group_count(0); ++q) cf[0][q] = cp[4 * 25 + q];
#pragma unroll
  for (int c = 0; c < 6; c += 2) {
    const A2 v = *reinterpret_cast<const A2 *>(&pl[xo + c]);
    x[0][c] = v.x;
    x[0][c + 1] = v.y;
  }
  quad_steady_group<0, A>(pl, xo, cp, cf, x, acc);
}

// The other members of the quad anchored at m0 (m1 = i+1, m2 = j+1,
// m3 = i+1,j+1) and the brick tables of the 4x4 block of columns around it:
// level 0 in tbl[16..32), its k-1 neighbours (the input level of the first
// two planes) in tbl[0..16).  Entry (r, c) comes from one member's in-plane
// neighbours: m0 for r, c < 3, m1 for column 3, m2 for row 3, m3 for the
// corner -- one writer per entry.  Level 0 is taken from the in-plane
// neighbours directly, not by walking up from level -1 (where level -1 lies
// outside the grid its adjacency entries stand for no brick).  The plan
// builder has checked that every stored member's 27 neighbours are what
// these tables give.
__device__ __forceinline__ void quad_tables(const uint32_t *__restrict__ adj, uint32_t m0,
                                            uint32_t &m1, uint32_t &m2, uint32_t &m3,
                                            uint32_t *tbl, int lane) {
  m1 = adj[(size_t)m0 * 27 + 14];
  m2 = adj[(size_t)m0 * 27 + 16];
  m3 = adj[(size_t)m1 * 27 + 16];
  if (lane < 16) {
    const int r = lane >> 2, c = lane & 3;
    const uint32_t m = r < 3 ? (c < 3 ? m0 : m1) : (c < 3 ? m2 : m3);
    const int jd = r < 3 ? r : 2, id = c < 3 ? c : 2;
    const int jj = (r == 3) ? 2 : jd, ii = (c == 3) ? 2 : id;
    const uint32_t n0 = adj[(size_t)m * 27 + 9 + 3 * jj + ii];
    tbl[lane] = adj[(size_t)n0 * 27 + 4];
    tbl[16 + lane] = n0;
  }
  __syncthreads();
}

// One plan unit of the quad schedule: members m0.. over L levels, store
// masks mk[0..L), coefficients through the constant-address pointer cp.
// plane: two LDS plane buffers, tbl: 32-entry LDS brick table (the current
// input level's 4x4 block of columns, then the unit's level-0 block); one
// wave.
template <class T>
__device__ __forceinline__ void cube_quad_unit(
    const uint32_t *__restrict__ adj, const T *__restrict__ in, T *__restrict__ out, size_t st,
    size_t ost, uint32_t m0, int L, const uint32_t *__restrict__ mk,
    kconst_ptr<typename acc_of<T>::type> cp0, typename acc_of<T>::type (*plane)[kQuadCells],
    uint32_t *tbl, int lane) {
  using A = typename acc_of<T>::type;
  // the padded plane in 2-cell chunks: a 20-cell row is 10 chunks, each
  // inside one brick row and 2-cell aligned there (16 bytes in FP64)
  constexpr int CPR = kQuadPW / 2, NCH = kQuadPW * CPR, NLD = (NCH + 63) / 64;
  const int lj = lane >> 3, li = lane & 7;
  uint32_t m1, m2, m3;
  quad_tables(adj, m0, m1, m2, m3, tbl, lane);

  // Load t of a plane fetches chunk q = 64t + lane of the 20x20 padded
  // plane, two cells in one load (the last load only for the first
  // NCH - 64(NLD-1) lanes): four loads per plane instead of seven.  Its only
  // state across planes is one 32-bit element offset into the input field,
  // brick * step + in-brick offset (bk_plan_create and bk_batch_set reject
  // fields beyond 2^32 elements for the quad schedule), recomputed from the
  // brick table at each new input level.
  const bool tail = lane < NCH - 64 * (NLD - 1);
  const uint32_t st32 = (uint32_t)st;
  uint32_t off[NLD];
  auto set_offsets = [&]() {
    // the lane index through an empty asm: the compiler cannot hoist the
    // per-load cell arithmetic out of the plane loop and keep it live
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const bool tl = ln < NCH - 64 * (NLD - 1);
#pragma unroll
    for (int t = 0; t < NLD; ++t) {
      const int q = t * 64 + ln;
      const int qq = (t < NLD - 1 || tl) ? q : 0;
      const int J = qq / CPR + 6, I = (qq % CPR) * 2 + 6;  // cell + 8 in the 4x4 block
      off[t] = tbl[(J >> 3) * 4 + (I >> 3)] * st32 + (uint32_t)((J & 7) * 8 + (I & 7));
    }
  };
  auto lds_cell = [&](int t) {
    const int q = t * 64 + lane;
    return (q / CPR) * kQuadStride + (q % CPR) * 2;
  };
  typedef A A2 __attribute__((ext_vector_type(2)));
  auto load2 = [&](uint32_t e) -> A2 {
    const Pair<T> v = *reinterpret_cast<const Pair<T> *>(in + e);
    A2 r;
    r.x = widen(v.a);
    r.y = widen(v.b);
    return r;
  };
  set_offsets();
  uint32_t tn = 0;  // next level's table entry (lanes 0..15)
  if (lane < 16) tn = tbl[16 + lane];

  const int mm_l = (lj >> 2) * 2 + (li >> 2);
  const int jo = (2 * lj) & 7, io = (2 * li) & 7;
  uint32_t om0 = m0, om1 = m1, om2 = m2, om3 = m3;

  A2 nv[NLD];
#pragma unroll
  for (int t = 0; t < NLD; ++t)
    if (t < NLD - 1 || tail) nv[t] = load2(off[t] + 6 * 64);
#pragma unroll
  for (int t = 0; t < NLD; ++t)
    if (t < NLD - 1 || tail) *reinterpret_cast<A2 *>(&plane[0][lds_cell(t)]) = nv[t];
  __syncthreads();

  A acc[5][2][2];
#pragma unroll
  for (int s = 0; s < 5; ++s)
#pragma unroll
    for (int pj = 0; pj < 2; ++pj)
#pragma unroll
      for (int pi = 0; pi < 2; ++pi) acc[s][pj][pi] = A(0);

  const int nplanes = 8 * L + 4;
  const int nout = 8 * L;
  for (int it = 0; it < nplanes; ++it) {
    const int P = it - 2;
    const int buf = it & 1;
    const bool have_next = it + 1 < nplanes;
    if (have_next) {
      const int Pn = P + 1;
      if ((Pn & 7) == 0) {  // next plane starts a new input level
        if (lane < 16) tbl[lane] = tn;
        __syncthreads();
        set_offsets();
        if (lane < 16) tn = adj[(size_t)tbl[lane] * 27 + 22];
      }
      const int kp = Pn & 7;
#pragma unroll
      for (int t = 0; t < NLD; ++t)
        if (t < NLD - 1 || tail) nv[t] = load2(off[t] + kp * 64);
    }

    const A *pl = plane[buf];
    // FP64 keeps the edge form throughout: the pipelined form's extra live
    // registers (two rows, two coefficient groups) cost it a wave per SIMD
    // (155 VGPRs) for no gain; FP32 gains about 2% from it
    if (sizeof(A) == 4 && P >= 2 && P + 2 < nout) quad_plane_steady<A>(pl, lj, li, cp0, acc);
    else quad_plane_edge<A>(pl, lj, li, cp0, P, nout, acc);

    const int O = P - 2;  // slot 0 is complete
    if (O >= 0 && O < nout) {
      const int lo = O >> 3, kk = O & 7;
      if (kk == 0 && lo > 0) {
        om0 = adj[(size_t)om0 * 27 + 22];
        om1 = adj[(size_t)om1 * 27 + 22];
        om2 = adj[(size_t)om2 * 27 + 22];
        om3 = adj[(size_t)om3 * 27 + 22];
      }
      if ((mk[lo] >> mm_l) & 1u) {
        const uint32_t ob = mm_l == 0 ? om0 : (mm_l == 1 ? om1 : (mm_l == 2 ? om2 : om3));
        T *po = out + (size_t)ob * ost + kk * 64 + jo * 8 + io;
        narrow(po[0], acc[0][0][0]);
        narrow(po[1], acc[0][0][1]);
        narrow(po[8], acc[0][1][0]);
        narrow(po[9], acc[0][1][1]);
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int pj = 0; pj < 2; ++pj)
#pragma unroll
        for (int pi = 0; pi < 2; ++pi) acc[s][pj][pi] = acc[s + 1][pj][pi];

    if (have_next) {
#pragma unroll
      for (int t = 0; t < NLD; ++t)
        if (t < NLD - 1 || tail) *reinterpret_cast<A2 *>(&plane[buf ^ 1][lds_cell(t)]) = nv[t];
    }
    __syncthreads();
  }
}

// FP64 form of the quad unit.  The plane moves from the input field into
// LDS by direct loads (global_load_lds_dwordx4): load t, lane l fetches
// chunk q = 64t + l of the plane taken as rows of 12 two-cell chunks (10 of
// data, 2 of padding), and the hardware writes it at byte 16q of the LDS
// plane -- the 24-cell row stride of the padded plane, with no register
// round trip.  Padding chunks and chunks past row 19 fetch a cell of the
// plane again and land in cells no read touches.  Each plane's 6x6 window
// of the lane's 2x2 points is read into registers once; the single LDS
// plane then takes the next plane's loads while the window computes, slot
// by slot (the 25 coefficients of one output plane per scalar wait).
constexpr int kQuadDmaCPR = 12, kQuadDmaLoads = 4;
static_assert(kQuadDmaLoads * 64 * 2 <= 2 * kQuadCells, "the direct-load plane fits");
static_assert(kQuadDmaLoads * 64 >= kQuadPW * kQuadDmaCPR, "four loads cover the plane");
static_assert(kQuadDmaCPR * 2 == kQuadStride, "chunk rows are LDS rows");
typedef __attribute__((address_space(3))) void *lds_void_ptr;

// 16 bytes per lane from global memory straight into LDS: lane l's bytes land
// at lds + 16 l (lds is wave-uniform)
__device__ __forceinline__ void load_lds16(const void *g, void *lds) {
  __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)lds, 16, 0, 0);
}

// Slot s of input plane P (output plane P - 2 + s, tap plane dk = 2 - s):
// per output point the taps go (dj, di) lexicographic within the slot, and
// successive planes bring dk = -2 .. 2, so the order per point is the
// per-brick kernel's.
template <class A>
__device__ __forceinline__ void quad_plane_window(const A (&x)[6][6], kconst_ptr<A> cp, int P,
                                                  int nout, A (&acc)[5][2][2]) {
#pragma unroll
  for (int s = 0; s < 5; ++s) {
    const int O = P - 2 + s;
    if (O < 0 || O >= nout) continue;
    asm volatile("" : "+s"(cp));
    const int dk = 2 - s;
#pragma unroll
    for (int dj = -2; dj <= 2; ++dj)
#pragma unroll
      for (int di = -2; di <= 2; ++di) {
        const A cf = cp[(dk + 2) * 25 + (dj + 2) * 5 + (di + 2)];
#pragma unroll
        for (int pj = 0; pj < 2; ++pj)
#pragma unroll
          for (int pi = 0; pi < 2; ++pi) {
            const A xv = x[pj + dj + 2][pi + di + 2];
            if (s == 4 && dj == -2 && di == -2) acc[s][pj][pi] = cf * xv;
            else acc[s][pj][pi] = fmadd(cf, xv, acc[s][pj][pi]);
          }
      }
  }
}

template <class T>
__device__ __forceinline__ void cube_quad_unit_dma(
    const uint32_t *__restrict__ adj, const T *__restrict__ in, T *__restrict__ out, size_t st,
    size_t ost, uint32_t m0, int L, const uint32_t *__restrict__ mk,
    kconst_ptr<typename acc_of<T>::type> cp0, typename acc_of<T>::type *lds, uint32_t *tbl,
    int lane) {
  using A = typename acc_of<T>::type;
  static_assert(sizeof(T) == 8 && sizeof(A) == 8, "a 16-byte direct load carries two cells");
  constexpr int CPR = kQuadDmaCPR, NLD = kQuadDmaLoads;
  const int lj = lane >> 3, li = lane & 7;
  uint32_t m1, m2, m3;
  quad_tables(adj, m0, m1, m2, m3, tbl, lane);

  // one 32-bit element offset per load, recomputed at each new input level
  // (fields beyond 2^32 elements are rejected for the quad schedule)
  const uint32_t st32 = (uint32_t)st;
  uint32_t off[NLD];
  auto set_offsets = [&]() {
    int ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (int t = 0; t < NLD; ++t) {
      const int q = t * 64 + ln;
      const int row = min(q / CPR, kQuadPW - 1), cc = min(q % CPR, kQuadPW / 2 - 1);
      const int J = row + 6, I = cc * 2 + 6;  // cell + 8 in the 4x4 block
      off[t] = tbl[(J >> 3) * 4 + (I >> 3)] * st32 + (uint32_t)((J & 7) * 8 + (I & 7));
    }
  };
  auto load_plane = [&](uint32_t kofs) {
#pragma unroll
    for (int t = 0; t < NLD; ++t)
      load_lds16(in + off[t] + kofs, lds + t * 128);
  };
  set_offsets();
  uint32_t tn = 0;  // next level's table entry (lanes 0..15)
  if (lane < 16) tn = tbl[16 + lane];
  load_plane(6 * 64);

  const int mm_l = (lj >> 2) * 2 + (li >> 2);
  const int jo = (2 * lj) & 7, io = (2 * li) & 7;
  const int xo = (2 * lj) * kQuadStride + 2 * li;
  uint32_t om0 = m0, om1 = m1, om2 = m2, om3 = m3;

  A acc[5][2][2];
#pragma unroll
  for (int s = 0; s < 5; ++s)
#pragma unroll
    for (int pj = 0; pj < 2; ++pj)
#pragma unroll
      for (int pi = 0; pi < 2; ++pi) acc[s][pj][pi] = A(0);

  const int nplanes = 8 * L + 4;
  const int nout = 8 * L;
  for (int it = 0; it < nplanes; ++it) {
    const int P = it - 2;
    // the plane's direct loads have landed in LDS
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    A x[6][6];
    typedef A A2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; c += 2) {
        const A2 v = *reinterpret_cast<const A2 *>(&lds[xo + r * kQuadStride + c]);
        x[r][c] = v.x;
        x[r][c + 1] = v.y;
      }
    // the window is in registers before the next plane overwrites LDS
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    if (it + 1 < nplanes) {
      const int Pn = P + 1;
      if ((Pn & 7) == 0) {  // next plane starts a new input level
        if (lane < 16) tbl[lane] = tn;
        __syncthreads();
        set_offsets();
        if (lane < 16) tn = adj[(size_t)tbl[lane] * 27 + 22];
      }
      load_plane((uint32_t)((Pn & 7) * 64));
    }

    quad_plane_window<A>(x, cp0, P, nout, acc);

    const int O = P - 2;  // slot 0 is complete
    if (O >= 0 && O < nout) {
      const int lo = O >> 3, kk = O & 7;
      if (kk == 0 && lo > 0) {
        om0 = adj[(size_t)om0 * 27 + 22];
        om1 = adj[(size_t)om1 * 27 + 22];
        om2 = adj[(size_t)om2 * 27 + 22];
        om3 = adj[(size_t)om3 * 27 + 22];
      }
      if ((mk[lo] >> mm_l) & 1u) {
        const uint32_t ob = mm_l == 0 ? om0 : (mm_l == 1 ? om1 : (mm_l == 2 ? om2 : om3));
        T *po = out + (size_t)ob * ost + kk * 64 + jo * 8 + io;
        narrow(po[0], acc[0][0][0]);
        narrow(po[1], acc[0][0][1]);
        narrow(po[8], acc[0][1][0]);
        narrow(po[9], acc[0][1][1]);
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int pj = 0; pj < 2; ++pj)
#pragma unroll
        for (int pi = 0; pi < 2; ++pi) acc[s][pj][pi] = acc[s + 1][pj][pi];
  }
}

// FP64: 3 waves per SIMD (135 VGPRs); capping at 128 for a fourth wave
// spills and runs slower
template <class T>
__global__ __launch_bounds__(64) void cube_sweep_quads(const CubeArgs<T> a) {
  using A = typename acc_of<T>::type;
  __shared__ __attribute__((aligned(16))) A plane[2][kQuadCells];
  __shared__ uint32_t tbl[32];
  const uint32_t u = xcd_remap(blockIdx.x, gridDim.x);
  if (u >= a.nunits) return;
  // The argument struct is the kernel's only parameter and sits at offset 0
  // of the kernel-argument segment: its coefficients are read from there by
  // scalar loads.
  const kconst_ptr<A> cp =
      (kconst_ptr<A>)((kconst_ptr<char>)__builtin_amdgcn_kernarg_segment_ptr() +
                      offsetof(CubeArgs<T>, c));
  if constexpr (sizeof(T) == 8)
    cube_quad_unit_dma<T>(a.adj, a.in, a.out, a.in_step, a.out_step, a.anchor[u], (int)a.len[u],
                          a.mask + (size_t)u * a.kcol, cp, &plane[0][0], tbl, threadIdx.x);
  else
    cube_quad_unit<T>(a.adj, a.in, a.out, a.in_step, a.out_step, a.anchor[u], (int)a.len[u],
                      a.mask + (size_t)u * a.kcol, cp, plane, tbl, threadIdx.x);
}

// Batched variant: one persistent launch serves every cube request of a
// group table; work item = one quad-plan unit of one request.
template <class T>
__global__ __launch_bounds__(64) void cube_sweep_batched(const BatchTable *tab_,
                                                                      uint32_t tick) {
  using A = typename acc_of<T>::type;
  __shared__ __attribute__((aligned(16))) A plane[2][kQuadCells];
  __shared__ uint32_t tbl[32];
  const kconst_ptr<BatchTable> tab = (kconst_ptr<BatchTable>)tab_;
  const uint32_t nsegs = tab->nsegs;
  const uint32_t total = tab->prefix[nsegs];
  SegCursor cur;
  for (uint32_t u = xcd_remap(blockIdx.x, gridDim.x); u < total; u += gridDim.x) {
    cur.seek(tab->prefix, nsegs, u);
    const auto &sg = tab->seg[cur.s];
    if (tick >= sg.ticks) continue;
    const uint32_t ul = u - cur.lo;
    const bool odd = ((sg.parity + tick) & 1u) != 0;
    const T *in = static_cast<const T *>(odd ? sg.f1 : sg.f0);
    T *out = static_cast<T *>(odd ? sg.f0 : sg.f1);
    if constexpr (sizeof(T) == 8)
      cube_quad_unit_dma<T>(sg.adj, in, out, sg.step, sg.step, sg.items[ul], (int)sg.len[ul],
                            sg.mask + (size_t)ul * sg.kcol, (kconst_ptr<A>)sg.coeff,
                            &plane[0][0], tbl, threadIdx.x);
    else
      cube_quad_unit<T>(sg.adj, in, out, sg.step, sg.step, sg.items[ul], (int)sg.len[ul],
                        sg.mask + (size_t)ul * sg.kcol, (kconst_ptr<A>)sg.coeff, plane, tbl,
                        threadIdx.x);
  }
}

// the persistent grid: as many one-wave workgroups as fit on the device
template <class T>
int cube_batched_blocks() {
  static int nblk = 0;
  if (nblk == 0) {
    int dev = 0, cus = 256, per = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, cube_sweep_batched<T>, 64, 0);
    nblk = cus * (per > 0 ? per : 1);
  }
  return nblk;
}

template <class T>
hipError_t launch_cube_batched_t(const BatchTable *tab, uint32_t tick, hipStream_t s) {
  hipLaunchKernelGGL((cube_sweep_batched<T>), dim3(cube_batched_blocks<T>()), dim3(64), 0, s,
                     tab, tick);
  return hipGetLastError();
}

template <class T>
hipError_t launch_quads_t(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  using A = typename acc_of<T>::type;
  CubeArgs<T> a;
  a.adj = L.adj;
  a.list = nullptr;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = 0;
  a.anchor = P.anchor;
  a.len = P.len;
  a.mask = P.mask;
  a.nunits = P.nunits;
  a.kcol = P.kcol;
  for (int q = 0; q < 125; ++q) a.c[q] = (A)L.coeff[q];
  if (P.nunits == 0) return hipSuccess;
  hipLaunchKernelGGL((cube_sweep_quads<T>), dim3(P.nunits), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <class T>
hipError_t launch_t(const SweepLaunch &L, hipStream_t s) {
  using A = typename acc_of<T>::type;
  CubeArgs<T> a;
  a.adj = L.adj;
  a.list = L.list;
  a.in = static_cast<const T *>(L.in);
  a.out = static_cast<T *>(L.out);
  a.in_step = L.in_step;
  a.out_step = L.out_step;
  a.nlist = L.nlist;
  a.anchor = a.len = a.mask = nullptr;
  a.nunits = a.kcol = 0;
  for (int q = 0; q < 125; ++q) a.c[q] = (A)L.coeff[q];
  const uint32_t nwg = (L.nlist + 3u) / 4u;
  hipLaunchKernelGGL((cube_sweep_lds<T>), dim3(nwg), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_cube(const SweepLaunch &L, hipStream_t s) {
  switch (L.dtype) {
    case BK_F64: return launch_t<double>(L, s);
    case BK_F32: return launch_t<float>(L, s);
    case BK_BF16: return launch_t<bf16>(L, s);
    default: return hipErrorInvalidValue;
  }
}

uint32_t cube_batched_capacity(int dtype) {
  switch (dtype) {
    case BK_F64: return (uint32_t)cube_batched_blocks<double>();
    case BK_F32: return (uint32_t)cube_batched_blocks<float>();
    case BK_BF16: return (uint32_t)cube_batched_blocks<bf16>();
    default: return 0;
  }
}

hipError_t launch_cube_batched(int dtype, const BatchTable *tab, uint32_t tick, hipStream_t s) {
  switch (dtype) {
    case BK_F64: return launch_cube_batched_t<double>(tab, tick, s);
    case BK_F32: return launch_cube_batched_t<float>(tab, tick, s);
    case BK_BF16: return launch_cube_batched_t<bf16>(tab, tick, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_cube_quads(const SweepLaunch &L, const PlanView &P, hipStream_t s) {
  switch (L.dtype) {
    case BK_F64: return launch_quads_t<double>(L, P, s);
    case BK_F32: return launch_quads_t<float>(L, P, s);
    case BK_BF16: return launch_quads_t<bf16>(L, P, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bk
