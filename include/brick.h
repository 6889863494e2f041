/**
 * @file brick.h
 * @brief Brick storage, adjacency metadata and the element accessor.
 *
 * Keeps the template API of bricklib's include/brick.h (BrickStorage
 * brick.h:53-82, BrickInfo brick.h:96-127, Dim / cal_size / cal_offs
 * brick.h:130-194, Brick<Dim<BDims...>,Dim<Folds...>> brick.h:353-395) so
 * user code written against it compiles unchanged, with the same memory
 * layout: element (x_0..x_{n-1}) (outermost first) of brick b lives at
 *   dat[adj[b][slot] * step + nvec * VECLEN + lane]
 * where each x_d in [-D_d, 2*D_d) picks the neighbour digit
 * (x_d + D_d) / D_d of the 3^n adjacency slot (outermost digit most
 * significant) and the in-brick coordinate (x_d + D_d) % D_d; the innermost
 * dims are split by the vector fold into the vector index nvec and the lane.
 * For the 8x8x8 brick in the 8x8 fold this is k*64 + j*8 + i.
 *
 * Host-only header (no GPU code).  Device residency comes from
 * brick-hip.h, which forwards to libbricks.so through bk_api.h.
 */
#ifndef BK_BRICK_H
#define BK_BRICK_H

#include <cstdlib>
#include <memory>
#include <type_traits>

#include "vecscatter.h"

/// BrickStorage allocation alignment in bytes (brick.h:15).
#ifndef ALIGN
#define ALIGN 2048
#endif

#if defined(__HIPCC__)
#define FORCUDA __host__ __device__
#else
#define FORCUDA
#endif

/// base^exp at compile time.
template <unsigned base, unsigned exp>
struct static_power {
  static constexpr unsigned value = base * static_power<base, exp - 1>::value;
};
template <unsigned base>
struct static_power<base, 0> {
  static constexpr unsigned value = 1;
};

/**
 * @brief Memory holding bricks: `chunks` chunks of `step` elements each.
 *
 * A chunk may interleave several fields (a field is an offset inside it).
 * `allocate` is the host allocator of the reference; device-resident
 * storage comes from BrickStorage::allocate_device / movBrickStorage in
 * brick-hip.h, which set `on_device`.
 */
struct BrickStorage {
  std::shared_ptr<bElem> dat;
  long chunks = 0;
  size_t step = 0;
  void *mmap_info = nullptr;
  bool on_device = false;

  static BrickStorage allocate(long chunks, size_t step) {
    BrickStorage b;
    b.chunks = chunks;
    b.step = step;
    size_t bytes = (size_t)chunks * step * sizeof(bElem);
    bytes = (bytes + ALIGN - 1) / ALIGN * ALIGN;
    b.dat = std::shared_ptr<bElem>(static_cast<bElem *>(::aligned_alloc(ALIGN, bytes)),
                                   [](bElem *p) { std::free(p); });
    return b;
  }
};

/**
 * @brief Adjacency list of `nbricks` bricks: adj[b][3^dims], slot 3^dims/2
 * is the brick itself.  `adj` is malloc'ed (the caller frees it, as in the
 * reference) or, after movBrickInfo, a device pointer.
 */
template <unsigned dims>
struct BrickInfo {
  typedef unsigned (*adjlist)[static_power<3, dims>::value];
  adjlist adj;
  unsigned nbricks;

  explicit BrickInfo(unsigned nbricks) : nbricks(nbricks) {
    adj = (adjlist)std::malloc((size_t)nbricks * static_power<3, dims>::value * sizeof(unsigned));
  }

  BrickStorage allocate(long step) { return BrickStorage::allocate(nbricks, step); }
};

/// Compile-time list of extents (outermost first).
template <unsigned... Ds>
struct Dim {};

/// Product of a list of numbers.
template <unsigned... xs>
struct cal_size;
template <unsigned x>
struct cal_size<x> {
  static constexpr unsigned value = x;
};
template <unsigned x, unsigned... xs>
struct cal_size<x, xs...> {
  static constexpr unsigned value = x * cal_size<xs...>::value;
};

/// Adjacency slot of neighbour digits offs (each 0..2, outermost first) in
/// `dim` dimensions: sum offs_d * 3^(dim-1-d).
template <unsigned dim, unsigned... offs>
struct cal_offs;
template <unsigned off>
struct cal_offs<1, off> {
  static constexpr unsigned value = off;
};
template <unsigned dim, unsigned off, unsigned... offs>
struct cal_offs<dim, off, offs...> {
  static constexpr unsigned value =
      off * static_power<3, dim - 1>::value + cal_offs<dim - 1, offs...>::value;
};

template <typename...>
struct Brick;

namespace bk_detail {

template <unsigned... Vs>
struct uarray {
  static constexpr unsigned n = sizeof...(Vs);
  static constexpr unsigned v[sizeof...(Vs) > 0 ? sizeof...(Vs) : 1] = {Vs...};
};

// Offset inside a brick (nvec * VECLEN + lane) of in-brick coordinates
// l[0..N) (outermost first) with the innermost NF dims folded by F.
template <unsigned N, unsigned NF>
FORCUDA inline unsigned fold_offset(const unsigned *D, const unsigned *F, const unsigned *l,
                                    unsigned veclen) {
  unsigned nvec = 0, lane = 0;
  for (unsigned d = 0; d < N; ++d) {
    if (d + NF < N) {
      nvec = nvec * D[d] + l[d];
    } else {
      const unsigned f = F[d + NF - N];
      nvec = nvec * (D[d] / f) + l[d] / f;
      lane = lane * f + l[d] % f;
    }
  }
  return nvec * veclen + lane;
}

}  // namespace bk_detail

/**
 * @brief Partially indexed brick: `Left` more subscripts to go.  The last
 * subscript returns the element reference.
 */
template <typename BrickT, unsigned Left>
struct _BrickAccessor {
  BrickT *par;
  unsigned b;
  int idx[BrickT::NDIMS];

  FORCUDA _BrickAccessor<BrickT, Left - 1> operator[](int x) const {
    _BrickAccessor<BrickT, Left - 1> r{par, b, {}};
    for (unsigned d = 0; d < BrickT::NDIMS; ++d) r.idx[d] = idx[d];
    r.idx[BrickT::NDIMS - Left] = x;
    return r;
  }
};

template <typename BrickT>
struct _BrickAccessor<BrickT, 1> {
  BrickT *par;
  unsigned b;
  int idx[BrickT::NDIMS];

  FORCUDA bElem &operator[](int x) const {
    int full[BrickT::NDIMS];
    for (unsigned d = 0; d + 1 < BrickT::NDIMS; ++d) full[d] = idx[d];
    full[BrickT::NDIMS - 1] = x;
    return par->elem(b, full);
  }
};

/**
 * @brief A field of bricks: metadata + storage + offset, indexed as
 * brick[b][x_0]...[x_{n-1}] with x_d in [-D_d, 2*D_d) reaching the 3^n
 * neighbours.
 */
template <unsigned... BDims, unsigned... Folds>
struct Brick<Dim<BDims...>, Dim<Folds...>> {
  typedef Brick<Dim<BDims...>, Dim<Folds...>> mytype;
  typedef BrickInfo<sizeof...(BDims)> myBrickInfo;

  static constexpr unsigned NDIMS = sizeof...(BDims);
  static constexpr unsigned NFOLDS = sizeof...(Folds);
  static constexpr unsigned VECLEN = cal_size<Folds...>::value;
  static constexpr unsigned BRICKSIZE = cal_size<BDims...>::value;
  static_assert(NFOLDS <= NDIMS, "more folds than brick dimensions");

  myBrickInfo *bInfo;
  size_t step;
  bElem *dat;
  BrickStorage bStorage;

  /// Extent of dimension d (outermost first).
  static constexpr unsigned extent(unsigned d) {
    constexpr unsigned D[NDIMS] = {BDims...};
    return D[d];
  }

  Brick(myBrickInfo *bInfo, const BrickStorage &brickStorage, unsigned offset) : bInfo(bInfo) {
    bStorage = brickStorage;
    dat = bStorage.dat.get() + offset;
    step = bStorage.step;
  }

  FORCUDA _BrickAccessor<mytype, NDIMS> operator[](unsigned b) {
    return _BrickAccessor<mytype, NDIMS>{this, b, {}};
  }

  /// Start of neighbour brick (Offsets: digits 0..2, outermost first).
  template <unsigned... Offsets>
  FORCUDA bElem *neighbor(unsigned b) {
    return &dat[(size_t)bInfo->adj[b][cal_offs<NDIMS, Offsets...>::value] * step];
  }

  /// Element at x[0..NDIMS) (outermost first) of brick b, through the
  /// adjacency.
  FORCUDA bElem &elem(unsigned b, const int *x) {
    static constexpr unsigned D[NDIMS] = {BDims...};
    static constexpr unsigned F[NFOLDS > 0 ? NFOLDS : 1] = {Folds...};
    unsigned l[NDIMS];
    unsigned slot = 0;
    for (unsigned d = 0; d < NDIMS; ++d) {
      const unsigned shifted = (unsigned)(x[d] + (int)D[d]);
      slot = slot * 3 + shifted / D[d];
      l[d] = shifted % D[d];
    }
    const unsigned off = bk_detail::fold_offset<NDIMS, NFOLDS>(D, F, l, VECLEN);
    return dat[(size_t)bInfo->adj[b][slot] * step + off];
  }
};

#endif  // BK_BRICK_H
