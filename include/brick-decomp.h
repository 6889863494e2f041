/**
 * @file brick-decomp.h
 * @brief BrickDecomp for a process that holds its whole periodic domain.
 *
 * Keeps the interface of bricklib's BrickDecomp (include/brick-mpi.h:178-713:
 * the constructor :304-320, initialize :322-460, exchange :466-500,
 * operator[] :502-505, getBrickInfo :511-513, and the public members ghost,
 * skin, sep_pos, skinlist, skin_size :185-198), of BitSet
 * (include/bitset.h:19-140) and of the skin lists skin3d_good,
 * skin3d_normal and skin3d_bad (src/brick-mpi.cpp:25-140), so code written
 * against them compiles unchanged:
 *
 *   BrickDecomp<3, 8, 8, 8> bDecomp({512, 512, 512}, 8);
 *   bDecomp.initialize(skin3d_good);
 *   BrickInfo<3> bInfo = bDecomp.getBrickInfo();
 *   ... grid[k][j][i] = bDecomp[k][j][i]; ...
 *   bDecomp.exchange(bStorage);
 *
 * The numbering (inner bricks, skin region by region in skin-list order,
 * then each neighbour's ghost bricks as contiguous runs) comes from
 * bk_decomp_layout_skin in libbricks.so.  Every neighbour of the one
 * subdomain is the subdomain itself, so exchange() copies each ghost run
 * from its skin run -- the periodic wrap -- on the device (one
 * bk_ghost_refresh launch over whole chunks, every interleaved field) when
 * the storage is device-resident, else on the host.  The reference's
 * multi-rank MPI exchange and its mmap views (exchangeView,
 * multiStageExchangeView, brick-mpi.h:521-686) have no counterpart: one
 * process drives one GPU here.
 *
 * Host-only header (no GPU code); link libbricks.so.  Failures throw
 * std::runtime_error or std::invalid_argument (the reference asserts or
 * throws std::runtime_error).
 */
#ifndef BK_BRICK_DECOMP_H
#define BK_BRICK_DECOMP_H

#include <unistd.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <vector>

#include "bk_api.h"
#include "brick.h"

/**
 * @brief A set of signed axis numbers: {1, -3} is "+i and -k" (axis 1 = i,
 * the contiguous one).  The bit layout is the reference's (element p > 0 at
 * bit p, -p at bit 31 + p), since user code compares and hashes `set`.
 */
struct BitSet {
  static const long zero = 31;  ///< -p lives at bit zero + p

  static inline uint64_t to_set(long pos) {
    return uint64_t(1) << (uint64_t)(pos < 0 ? zero - pos : pos);
  }

  uint64_t set;

  BitSet() : set(0) {}
  BitSet(uint64_t s) : set(s) {}
  BitSet(std::initializer_list<int> l) : set(0) {
    for (int p : l) set ^= to_set(p);
  }

  inline BitSet &flip(long pos) {
    set ^= to_set(pos);
    return *this;
  }
  inline long size() const { return __builtin_popcountll(set); }
  inline bool get(long pos) const { return (set & to_set(pos)) != 0; }
  inline BitSet operator&(BitSet a) const { return set & a.set; }
  inline BitSet operator|(BitSet a) const { return set | a.set; }
  inline BitSet operator^(BitSet a) const { return set ^ a.set; }
  inline operator bool() const { return set != 0; }
  /// subset: every element of this set is in a
  inline bool operator<=(BitSet a) const { return (set & a.set) == set; }
  /// superset
  inline bool operator>=(BitSet a) const { return (set & a.set) == a.set; }
  /// every element negated ({1, -2, 3} -> {-1, 2, -3}); elements +-1 .. +-31
  inline BitSet operator!() const {
    const uint64_t pos = set & 0xFFFFFFFEull;          // +1 .. +31 at bits 1 .. 31
    const uint64_t neg = (set >> 32) & 0x7FFFFFFFull;  // -1 .. -31 at bits 32 .. 62
    return BitSet((pos << zero) | (neg << 1));
  }
};

namespace bk_detail {
/// The 3-D regions as BitSets in (k, j, i) lexicographic order, -1 first,
/// i fastest, the empty centre included (27 entries).
inline std::vector<BitSet> lexicographic_regions() {
  std::vector<BitSet> out;
  for (int k = -1; k <= 1; ++k)
    for (int j = -1; j <= 1; ++j)
      for (int i = -1; i <= 1; ++i) {
        BitSet b;
        if (i) b.flip(i);
        if (j) b.flip(2 * j);
        if (k) b.flip(3 * k);
        out.push_back(b);
      }
  return out;
}
inline std::vector<BitSet> skin_normal() { return lexicographic_regions(); }
/// the 26 regions of the lexicographic order, an empty entry between each
/// two (no region shares a run with its successor)
inline std::vector<BitSet> skin_bad() {
  std::vector<BitSet> out;
  for (BitSet b : lexicographic_regions()) {
    if (b.set == 0) continue;  // (operator! negates elements; it is no emptiness test)
    if (!out.empty()) out.push_back(BitSet());
    out.push_back(b);
  }
  return out;
}
}  // namespace bk_detail

/// The recommended surface order: consecutive regions share axes, so each
/// neighbour's skin is one or two contiguous runs (src/brick-mpi.cpp:25-52).
inline const std::vector<BitSet> skin3d_good = {
    {1},       {1, -3},     {1, 2, -3},  {1, 2},     {1, 2, 3},   {2, 3},   {2},
    {2, -3},   {-1, 2, -3}, {-1, 2},     {-1, 2, 3}, {-1, 3},     {-1},     {-3},
    {-1, -3},  {-1, -2, -3}, {-1, -2},   {-1, -2, 3}, {-2, 3},    {-2},     {-2, -3},
    {1, -2, -3}, {1, -2},   {1, -2, 3},  {1, 3},     {3}};
/// Lexicographic order with the empty centre (src/brick-mpi.cpp:54-82).
inline const std::vector<BitSet> skin3d_normal = bk_detail::skin_normal();
/// Lexicographic order split by empty entries (src/brick-mpi.cpp:84-136).
inline const std::vector<BitSet> skin3d_bad = bk_detail::skin_bad();

template <unsigned dim, unsigned... BDims>
class BrickDecomp;

/**
 * @brief The 3-D decomposition of one subdomain into bricks of
 * BK x BJ x BI elements (outermost first, as Dim<BDims...>).
 */
template <unsigned BK, unsigned BJ, unsigned BI>
class BrickDecomp<3, BK, BJ, BI> {
 public:
  /// One exchange run (brick-mpi.h:185-193): bricks [pos, pos + len), the
  /// first first_pad and last last_pad of them padding.
  typedef struct {
    BitSet neighbor;  ///< ghost: the owner; skin: the receiver
    unsigned skin_st; ///< skin-list entries [skin_st, skin_ed) of the run
    unsigned skin_ed;
    unsigned pos;
    unsigned len;
    unsigned first_pad;
    unsigned last_pad;
  } g_region;
  std::vector<g_region> ghost;   ///< ghost runs, received from ghost[i].neighbor
  std::vector<g_region> skin;    ///< skin runs, skin[i] fills ghost[i] of skin[i].neighbor
  unsigned sep_pos[3];           ///< first skin id, first ghost id, end of ids
  std::vector<BitSet> skinlist;  ///< the surface order given to initialize
  std::vector<long> skin_size;   ///< ids per skin-list entry, padding included

  /**
   * @param dims interior size per axis (i, j, k) in elements
   * @param depth ghost depth in elements (a multiple of every brick extent)
   * @param numfield interleaved fields per chunk (sizes the page grouping)
   */
  BrickDecomp(const std::vector<unsigned> &dims, const unsigned depth, unsigned numfield = 1)
      : numfield(numfield) {
    if (dims.size() != 3) throw std::invalid_argument("BrickDecomp: dims needs 3 extents");
    const unsigned bdims[3] = {BI, BJ, BK};  // i, j, k
    for (int a = 0; a < 3; ++a) {
      if (depth % bdims[a] != 0 || dims[a] % bdims[a] != 0)
        throw std::invalid_argument("BrickDecomp: dims and depth must be multiples of the brick");
      this->dims.push_back(dims[a] / bdims[a]);
      g_depth.push_back(depth / bdims[a]);
    }
    if (g_depth[0] != g_depth[1] || g_depth[1] != g_depth[2])
      throw std::invalid_argument("BrickDecomp: the ghost depth must be the same number of "
                                  "bricks on every axis");
  }
  BrickDecomp(const BrickDecomp &) = delete;
  BrickDecomp &operator=(const BrickDecomp &) = delete;

  ~BrickDecomp() { release(); }

  /// Number the bricks in the surface order `skinlist` (each of the 26
  /// regions once; empty entries allowed).
  void initialize(const std::vector<BitSet> &skinlist) {
    release();
    this->skinlist = skinlist;
    std::vector<int32_t> enc;
    for (const BitSet &b : skinlist) {
      for (int a = 1; a <= 3; ++a) {
        const bool p = b.get(a), m = b.get(-a);
        if (p && m) throw std::invalid_argument("BrickDecomp: a skin region holds +a and -a");
        enc.push_back(p ? 1 : (m ? -1 : 0));
      }
      BitSet rest = b;
      for (int a = 1; a <= 3; ++a) {
        if (rest.get(a)) rest.flip(a);
        if (rest.get(-a)) rest.flip(-a);
      }
      if (rest) throw std::invalid_argument("BrickDecomp: skin elements must be +-1, +-2, +-3");
    }
    const long nb[3] = {(long)dims[0], (long)dims[1], (long)dims[2]};
    const size_t bytes = (size_t)BK * BJ * BI * sizeof(bElem) * numfield;
    const long page = sysconf(_SC_PAGESIZE);
    const int g = (int)g_depth[0];
    const int ns = (int)skinlist.size();
    uint32_t nbricks = 0, nreg = 0, sep[3];
    check(bk_decomp_layout_skin(nb, g, bytes, (size_t)page, enc.data(), ns, nullptr, nullptr,
                                &nbricks, sep, nullptr, &nreg, nullptr));
    t_dims.clear();
    stride.clear();
    unsigned cells = 1;
    for (int a = 0; a < 3; ++a) {
      t_dims.push_back(dims[a] + 2 * g_depth[a]);
      stride.push_back(cells);
      cells *= t_dims[a];
    }
    grid = new unsigned[cells];
    bInfo = new BrickInfo<3>(nbricks);
    std::vector<bk_decomp_region> regs(nreg);
    std::vector<uint32_t> sizes(ns);
    check(bk_decomp_layout_skin(nb, g, bytes, (size_t)page, enc.data(), ns, grid,
                                reinterpret_cast<uint32_t *>(bInfo->adj), &nbricks, sep,
                                regs.data(), &nreg, sizes.data()));
    for (int q = 0; q < 3; ++q) sep_pos[q] = sep[q];
    skin_size.assign(sizes.begin(), sizes.end());
    ghost.clear();
    skin.clear();
    for (const bk_decomp_region &r : regs) {
      BitSet n;
      for (int a = 0; a < 3; ++a)
        if (r.neighbor[a]) n.flip((a + 1) * r.neighbor[a]);
      g_region gr{n, r.skin_st, r.skin_ed, r.ghost_pos, r.len, r.first_pad, r.last_pad};
      g_region sr{!n, r.skin_st, r.skin_ed, r.skin_pos, r.len, r.first_pad, r.last_pad};
      ghost.push_back(gr);
      skin.push_back(sr);
    }
  }

  /**
   * @brief Fill every ghost brick from the skin brick it images (all
   * `step` elements of the chunk, so every interleaved field).  Blocking,
   * like the reference's exchange: a device-resident storage is refreshed
   * on stream s, which is then synchronised.
   */
  void exchange(BrickStorage &bStorage, bk_stream s = nullptr) {
    exchange_async(bStorage, s);
    if (bStorage.on_device) check(bk_stream_sync(s));
  }

  /// exchange() without the final synchronisation (device storage: the
  /// refresh is ordered on stream s before later work on s).
  void exchange_async(BrickStorage &bStorage, bk_stream s = nullptr) {
    require_initialized();
    if (bStorage.chunks < (long)sep_pos[2])
      throw std::invalid_argument("BrickDecomp::exchange: storage has fewer chunks than bricks");
    if (bStorage.on_device) {
      unsigned npairs = 0;
      const unsigned *d = device_pairs(npairs);
      check(bk_ghost_refresh(bk_dtype_of<bElem>::value, bStorage.dat.get(), bStorage.step,
                             bStorage.step, d, npairs, s));
      return;
    }
    bElem *dat = bStorage.dat.get();
    const size_t step = bStorage.step;
    for (size_t q = 0; q < ghost.size(); ++q) {
      const g_region &g = ghost[q], &k = skin[q];
      const size_t n = (size_t)(g.len - g.first_pad - g.last_pad) * step;
      std::memcpy(dat + (size_t)(g.pos + g.first_pad) * step,
                  dat + (size_t)(k.pos + k.first_pad) * step, n * sizeof(bElem));
    }
  }

  /**
   * @brief The exchange as (ghost, skin) brick pairs in HBM, for
   * brick_sweeps / bk_sweeps (their refresh before each sweep is this
   * exchange).  Built on first use; owned by the decomposition.
   */
  const unsigned *device_pairs(unsigned &npairs) {
    require_initialized();
    if (!d_pairs) {
      std::vector<unsigned> p;
      for (size_t q = 0; q < ghost.size(); ++q) {
        const g_region &g = ghost[q], &k = skin[q];
        for (unsigned b = g.first_pad; b + g.last_pad < g.len; ++b) {
          p.push_back(g.pos + b);
          p.push_back(k.pos + b);
        }
      }
      n_pairs = (unsigned)(p.size() / 2);
      void *d = nullptr;
      if (!p.empty()) {
        check(bk_alloc(p.size() * sizeof(unsigned), &d));
        check(bk_memcpy_h2d(d, p.data(), p.size() * sizeof(unsigned), nullptr));
      }
      d_pairs = static_cast<unsigned *>(d);
    }
    npairs = n_pairs;
    return d_pairs;
  }

  /// The interior bricks (ghost excluded) in grid order, k slowest.
  std::vector<unsigned> interior() const {
    require_initialized();
    std::vector<unsigned> out;
    const unsigned g = g_depth[0];
    for (unsigned k = g; k < g + dims[2]; ++k)
      for (unsigned j = g; j < g + dims[1]; ++j)
        for (unsigned i = g; i < g + dims[0]; ++i)
          out.push_back(grid[k * stride[2] + j * stride[1] + i]);
    return out;
  }

  /// grid[k][j][i] over the padded grid (ghost included), i contiguous.
  struct Row {
    const unsigned *p;
    unsigned operator[](int i) const { return p[i]; }
  };
  struct Plane {
    const unsigned *p;
    unsigned s;
    Row operator[](int j) const { return Row{p + (long)j * s}; }
  };
  Plane operator[](int k) {
    require_initialized();
    return Plane{grid + (long)k * stride[2], stride[1]};
  }

  /// The adjacency of the numbering (nbricks = sep_pos[2]).  Its adj array
  /// belongs to the decomposition: valid while it lives, not to be freed.
  BrickInfo<3> getBrickInfo() {
    require_initialized();
    return *bInfo;
  }

 private:
  std::vector<unsigned> dims;     ///< interior bricks per axis (i, j, k)
  std::vector<unsigned> t_dims;   ///< with the ghost zone
  std::vector<unsigned> g_depth;  ///< ghost bricks per axis
  std::vector<unsigned> stride;   ///< grid strides (i, j, k)
  unsigned *grid = nullptr;
  unsigned numfield;
  BrickInfo<3> *bInfo = nullptr;
  unsigned *d_pairs = nullptr;
  unsigned n_pairs = 0;

  static void check(int rc) {
    if (rc != BK_OK) throw std::runtime_error(std::string("BrickDecomp: ") + bk_last_error());
  }
  void require_initialized() const {
    if (!grid) throw std::runtime_error("BrickDecomp: call initialize(skinlist) first");
  }
  void release() {
    delete[] grid;
    grid = nullptr;
    if (bInfo) std::free(bInfo->adj);
    delete bInfo;
    bInfo = nullptr;
    if (d_pairs) bk_free(d_pairs);
    d_pairs = nullptr;
    n_pairs = 0;
  }
};

#endif  // BK_BRICK_DECOMP_H
