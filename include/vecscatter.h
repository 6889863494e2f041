/**
 * @file vecscatter.h
 * @brief Element type and the brick(...) / tile(...) operator hooks.
 *
 * In bricklib (include/vecscatter.h:12-90) `bElem` is the element type,
 * `double` unless defined before inclusion, and `brick(file, vec, dims,
 * fold, b)` / `tile(file, vec, dims, iters)` are pragmas that a Python code
 * generator replaces with vector code for the stencil described in `file`.
 * Here there is no code generator.  The same call sites compile unchanged
 * and compute the stencil the file names, through a table of the library's
 * stencils keyed by file name (7pt.py and cond.py built in; others through
 * bk_vs::register_stencil):
 *
 *  - brick(...) inside a loop over bricks reads the fields and coefficients
 *    the stencil file names at the call site -- `bIn`, `bOut` and `coeff`
 *    for the reference's 7pt.py and cond.py (override with BK_VS_IN,
 *    BK_VS_OUT, BK_VS_COEFF).  Host-resident fields: brick b is computed on
 *    the host at once.  Device-resident fields (brick-hip.h): the brick is
 *    queued and the queued bricks run as one GPU sweep at the next flush
 *    point -- a call with other fields, a brick already queued (the next
 *    sweep of the same loop), brick_flush(), or any brick-hip.h call that
 *    moves, sweeps or refreshes device bricks.
 *  - tile(...) inside a loop over array tiles whose iterators are named
 *    `tk`, `tj`, `ti` (the names the reference's call sites give in `iters`)
 *    evaluates the tile through the call site's function-like macros
 *    `bIn(i, j, k)` / `bOut(i, j, k)`.
 *
 * Both evaluate a point as acc = c_0 * x_0, then acc = fma(c_t, x_t, acc)
 * over the remaining taps in the library's kernel order (centre, then
 * i+r, i-r, j+r, j-r, k+r, k-r for r = 1..R; the cube in (dk, dj, di)
 * order), the order of the GPU kernels, so host and device results agree
 * bit for bit.  The `vec` (vector ISA) and stride arguments are accepted and
 * ignored; the vector fold of the bricks comes from their Brick type.
 *
 * C++14 is enough for this header and brick.h, bricksetup.h, brickcompare.h
 * and multiarray.h (the reference's external example builds as C++14).
 */
#ifndef BK_VECSCATTER_H
#define BK_VECSCATTER_H

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "bk_api.h"

#ifndef bElem
#define bElem double
#endif

/// Stencil kind of a reference stencil file name ("7pt.py", "cond.py",
/// "mpi13pt.py", "mpi25pt.py", "mpi125pt.py", also without ".py" or "mpi"
/// and with a leading path).  Returns -1 for an unknown name.
inline int bk_stencil_from_file(const char *file) {
  if (!file) return -1;
  const char *base = std::strrchr(file, '/');
  base = base ? base + 1 : file;
  if (std::strncmp(base, "mpi", 3) == 0) base += 3;
  struct Entry {
    const char *name;
    int kind;
  };
  static const Entry table[] = {{"7pt", BK_STAR7},   {"13pt", BK_STAR13},   {"25pt", BK_STAR25},
                                {"125pt", BK_CUBE125}, {"cond", BK_COND7}};
  for (const Entry &e : table) {
    const size_t n = std::strlen(e.name);
    if (std::strncmp(base, e.name, n) == 0 && (base[n] == 0 || std::strcmp(base + n, ".py") == 0))
      return e.kind;
  }
  return -1;
}

/// Element type tag of bElem for the C-ABI.
template <typename T> struct bk_dtype_of;
template <> struct bk_dtype_of<double> { static constexpr int value = BK_F64; };
template <> struct bk_dtype_of<float> { static constexpr int value = BK_F32; };

/// Names the brick(...) hook reads at its call site (the Grid and ConstRef
/// names of the reference's 7pt.py and cond.py).
#ifndef BK_VS_IN
#define BK_VS_IN bIn
#endif
#ifndef BK_VS_OUT
#define BK_VS_OUT bOut
#endif
#ifndef BK_VS_COEFF
#define BK_VS_COEFF coeff
#endif

namespace bk_vs {

constexpr int kMaxTaps = 125;

/// Coefficients and radius of a stencil kind (bk_stencil_ncoeff /
/// bk_stencil_radius without the library: host-only users need not link it).
inline int ncoeff_of(int kind) {
  return kind == BK_STAR7 || kind == BK_COND7 ? 7
         : kind == BK_STAR13                 ? 13
         : kind == BK_STAR25                 ? 25
         : kind == BK_CUBE125                ? 125
                                             : 0;
}
inline int radius_of(int kind) {
  return kind == BK_STAR7 || kind == BK_COND7 ? 1
         : kind == BK_STAR13 || kind == BK_CUBE125 ? 2
         : kind == BK_STAR25                      ? 4
                                                  : 0;
}

/// What a stencil file stands for: a kernel and, for each of the kernel's
/// coefficients (bk_api.h order), the index of the call site's coefficient
/// array it comes from.
struct Spec {
  char name[96] = {0};  // file base name
  int kind = -1;        // cond is its reference form If(p > 0, p, -p) (cond.py)
  int ncoeff = 0;
  int coeff_index[kMaxTaps] = {0};
};

/// The taps of a kernel in its coefficient order: offsets (di, dj, dk).
struct Taps {
  int n = 0;
  int di[kMaxTaps], dj[kMaxTaps], dk[kMaxTaps];
  bool relu = false;    // cond: max(x, 0) on every tap
  bool select = false;  // cond (reference form): out = p > 0 ? p : -p
  int radius = 0;
};

inline Taps taps_of(int kind) {
  Taps t;
  if (kind == BK_CUBE125) {
    for (int dk = -2; dk <= 2; ++dk)
      for (int dj = -2; dj <= 2; ++dj)
        for (int di = -2; di <= 2; ++di) {
          t.di[t.n] = di;
          t.dj[t.n] = dj;
          t.dk[t.n] = dk;
          ++t.n;
        }
    t.radius = 2;
    return t;
  }
  const int R = radius_of(kind);
  t.radius = R;
  t.di[0] = t.dj[0] = t.dk[0] = 0;
  t.n = 1;
  for (int r = 1; r <= R; ++r) {
    const int d[6][3] = {{r, 0, 0}, {-r, 0, 0}, {0, r, 0}, {0, -r, 0}, {0, 0, r}, {0, 0, -r}};
    for (int q = 0; q < 6; ++q) {
      t.di[t.n] = d[q][0];
      t.dj[t.n] = d[q][1];
      t.dk[t.n] = d[q][2];
      ++t.n;
    }
  }
  t.relu = t.select = kind == BK_COND7;
  return t;
}

inline const char *base_name(const char *file) {
  const char *b = std::strrchr(file, '/');
  return b ? b + 1 : file;
}

struct Registry {
  std::mutex mu;
  std::vector<std::unique_ptr<Spec>> specs;
};

inline Spec make_spec(const char *file, int kind, const int *coeff_index) {
  Spec s;
  std::strncpy(s.name, base_name(file), sizeof(s.name) - 1);
  s.kind = kind;
  s.ncoeff = ncoeff_of(kind);
  for (int q = 0; q < s.ncoeff; ++q) s.coeff_index[q] = coeff_index ? coeff_index[q] : q;
  return s;
}

inline Registry &registry() {
  static Registry r;
  static std::once_flag once;
  std::call_once(once, [] {
    // the reference's stencils/7pt.py and cond.py read coeff[0..6] in the
    // kernel's own tap order (centre, i+1, i-1, j+1, j-1, k+1, k-1); cond in
    // its reference form If(p > 0, p, -p) over max(x, 0) taps
    r.specs.emplace_back(new Spec(make_spec("7pt.py", BK_STAR7, nullptr)));
    r.specs.emplace_back(new Spec(make_spec("cond.py", BK_COND7, nullptr)));
  });
  return r;
}

/// Map a stencil file (its base name) to a kernel: `kind` (enum bk_stencil;
/// BK_COND7 is cond's reference form) and, for kernel coefficient q, the
/// call site's coeff[coeff_index[q]] (NULL: coeff[q]).
/// `python -m bricklib_amd.ingest FILE.py` prints a description file's
/// kernel and coefficient order.  Replaces an earlier mapping of the same
/// name.
inline void register_stencil(const char *file, int kind, const int *coeff_index = nullptr) {
  if (!file || ncoeff_of(kind) == 0)
    throw std::invalid_argument("bk_vs::register_stencil: unknown stencil kind");
  Registry &r = registry();
  std::lock_guard<std::mutex> lock(r.mu);
  const Spec s = make_spec(file, kind, coeff_index);
  for (auto &p : r.specs)
    if (std::strcmp(p->name, s.name) == 0) {
      *p = s;
      return;
    }
  r.specs.emplace_back(new Spec(s));
}

/// The mapping of a stencil file (a copy); std::invalid_argument when the
/// file is not registered.
inline Spec spec_of(const char *file) {
  Registry &r = registry();
  std::lock_guard<std::mutex> lock(r.mu);
  const char *b = base_name(file);
  for (auto &p : r.specs)
    if (std::strcmp(p->name, b) == 0) return *p;
  throw std::invalid_argument(std::string("brick()/tile(): no stencil registered for ") + file +
                              " (bk_vs::register_stencil maps a file to a kernel)");
}

/// A field of bricks as the device hook sees it.
struct Field {
  const void *dat;
  size_t step;
  const unsigned *adj;  // device adjacency (movBrickInfo)
  unsigned nbricks;
  bool on_device;
  unsigned ndims, bricksize, veclen;
};

/// Device path of brick(...): installed by brick-hip.h.
typedef void (*DeviceBrick)(const Spec &sp, const Field &in, const Field &out, const double *c,
                            unsigned b);
inline DeviceBrick &device_hook() {
  static DeviceBrick f = nullptr;
  return f;
}

template <class BrickT>
Field field_of(BrickT &x) {
  Field f;
  f.dat = x.dat;
  f.step = x.step;
  f.adj = reinterpret_cast<const unsigned *>(x.bInfo->adj);
  f.nbricks = x.bInfo->nbricks;
  f.on_device = x.bStorage.on_device;
  f.ndims = BrickT::NDIMS;
  f.bricksize = BrickT::BRICKSIZE;
  f.veclen = BrickT::VECLEN;
  return f;
}

/// One brick on the host through the Brick accessor (the adjacency gives
/// the neighbours), in the kernel order.
template <class BrickT>
void host_brick(const Spec &sp, BrickT &in, BrickT &out, const double *c, unsigned b) {
  static_assert(BrickT::NDIMS == 3, "brick(): 3-D bricks");
  const Taps t = taps_of(sp.kind);
  const int K = (int)BrickT::extent(0), J = (int)BrickT::extent(1), I = (int)BrickT::extent(2);
  if (t.radius > K || t.radius > J || t.radius > I)
    throw std::invalid_argument("brick(): stencil radius larger than the brick");
  bElem cc[kMaxTaps];
  for (int q = 0; q < t.n; ++q) cc[q] = (bElem)c[q];
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < J; ++j)
      for (int i = 0; i < I; ++i) {
        bElem acc = 0;
        for (int q = 0; q < t.n; ++q) {
          const int x[3] = {k + t.dk[q], j + t.dj[q], i + t.di[q]};
          bElem v = in.elem(b, x);
          if (t.relu) v = v < 0 ? bElem(0) : v;
          acc = q == 0 ? cc[0] * v : std::fma(cc[q], v, acc);
        }
        if (t.select) acc = acc > 0 ? acc : -acc;
        const int xo[3] = {k, j, i};
        out.elem(b, xo) = acc;
      }
}

/// The brick(...) hook: brick b of `out` from `in` for the stencil `file`.
template <class BrickT, class CoeffT>
void brick_at(const char *file, BrickT &in, BrickT &out, const CoeffT &coeff, unsigned b) {
  const Spec sp = spec_of(file);
  double c[kMaxTaps];
  for (int q = 0; q < sp.ncoeff; ++q) c[q] = (double)coeff[sp.coeff_index[q]];
  if (in.bStorage.on_device || out.bStorage.on_device) {
    if (!in.bStorage.on_device || !out.bStorage.on_device)
      throw std::invalid_argument("brick(): one field is on the device, the other on the host");
    if (!device_hook())
      throw std::logic_error("brick(): device-resident fields need brick-hip.h");
    device_hook()(sp, field_of(in), field_of(out), c, b);
    return;
  }
  host_brick(sp, in, out, c, b);
}

/// tile(...): the stencil's taps and coefficients for a tile of `ext`
/// points (outermost first).
struct TileSpec {
  Taps t;
  double c[kMaxTaps];
  long ext[3];
};

template <class CoeffT>
TileSpec tile_spec(const char *file, const CoeffT &coeff, long e0, long e1, long e2) {
  const Spec sp = spec_of(file);
  TileSpec ts;
  ts.t = taps_of(sp.kind);
  for (int q = 0; q < sp.ncoeff; ++q) ts.c[q] = (double)coeff[sp.coeff_index[q]];
  ts.ext[0] = e0;
  ts.ext[1] = e1;
  ts.ext[2] = e2;
  return ts;
}

constexpr bool str_eq(const char *a, const char *b) {
  while (*a && *a == *b) {
    ++a;
    ++b;
  }
  return *a == *b;
}

/// The tile iterators the call site names must be the loop variables the
/// hook reads: tk, tj, ti.
constexpr bool tile_iters_named(const char *a, const char *b, const char *c) {
  return str_eq(a, "tk") && str_eq(b, "tj") && str_eq(c, "ti");
}

}  // namespace bk_vs

#define BK_VS_LIST(...) __VA_ARGS__
#define BK_VS_PICK5(_1, _2, _3, _4, _5, NAME, ...) NAME
#define BK_VS_PICK6(_1, _2, _3, _4, _5, _6, NAME, ...) NAME

/// brick(file, vec, dims, fold, b [, stride]) -- the per-brick hook.
#define brick(...) BK_VS_PICK6(__VA_ARGS__, _bk_vs_brick6, _bk_vs_brick5, _bk_vs_none)(__VA_ARGS__)
#define _bk_vs_brick5(file, vec, vsdim, vsfold, bidx)                                   \
  do {                                                                                  \
    ::bk_vs::brick_at(file, BK_VS_IN, BK_VS_OUT, BK_VS_COEFF, (unsigned)(bidx));        \
  } while (false)
#define _bk_vs_brick6(file, vec, vsdim, vsfold, bidx, stri) \
  _bk_vs_brick5(file, vec, vsdim, vsfold, bidx)

/// tile(file, vec, dims, iters [, stride]) -- the array-tile hook.
#define tile(...) BK_VS_PICK5(__VA_ARGS__, _bk_vs_tile5, _bk_vs_tile4, _bk_vs_none)(__VA_ARGS__)
#define _bk_vs_tile5(file, vec, vsdim, titer, stri) _bk_vs_tile4(file, vec, vsdim, titer)
#define _bk_vs_tile4(file, vec, vsdim, titer)                                                  \
  do {                                                                                         \
    static_assert(::bk_vs::tile_iters_named titer,                                             \
                  "tile(): the tile loop variables must be named tk, tj, ti");                 \
    const ::bk_vs::TileSpec bk_vs_ts_ =                                                        \
        ::bk_vs::tile_spec(file, BK_VS_COEFF, BK_VS_LIST vsdim);                               \
    for (long bk_vs_k = 0; bk_vs_k < bk_vs_ts_.ext[0]; ++bk_vs_k)                              \
      for (long bk_vs_j = 0; bk_vs_j < bk_vs_ts_.ext[1]; ++bk_vs_j)                            \
        for (long bk_vs_i = 0; bk_vs_i < bk_vs_ts_.ext[2]; ++bk_vs_i) {                        \
          const long bk_vs_x = ti + bk_vs_i, bk_vs_y = tj + bk_vs_j, bk_vs_z = tk + bk_vs_k;  \
          bElem bk_vs_acc = 0;                                                                 \
          for (int bk_vs_q = 0; bk_vs_q < bk_vs_ts_.t.n; ++bk_vs_q) {                          \
            bElem bk_vs_v = bIn(bk_vs_x + bk_vs_ts_.t.di[bk_vs_q],                             \
                                bk_vs_y + bk_vs_ts_.t.dj[bk_vs_q],                             \
                                bk_vs_z + bk_vs_ts_.t.dk[bk_vs_q]);                            \
            if (bk_vs_ts_.t.relu) bk_vs_v = bk_vs_v < 0 ? bElem(0) : bk_vs_v;                  \
            bk_vs_acc = bk_vs_q == 0 ? (bElem)bk_vs_ts_.c[0] * bk_vs_v                         \
                                     : std::fma((bElem)bk_vs_ts_.c[bk_vs_q], bk_vs_v,         \
                                                bk_vs_acc);                                    \
          }                                                                                    \
          if (bk_vs_ts_.t.select) bk_vs_acc = bk_vs_acc > 0 ? bk_vs_acc : -bk_vs_acc;          \
          bOut(bk_vs_x, bk_vs_y, bk_vs_z) = bk_vs_acc;                                         \
        }                                                                                      \
  } while (false)

#endif  // BK_VECSCATTER_H
