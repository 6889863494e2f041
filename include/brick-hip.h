/**
 * @file brick-hip.h
 * @brief Device residency and GPU sweeps for the brick headers (MI355X).
 *
 * Replaces bricklib's include/brick-gpu.h / brick-hip.h:
 *  - movBrickInfo / movBrickInfoDeep (brick-gpu.h:42-76): adjacency list
 *    to/from HBM;
 *  - movBrickStorage (brick-gpu.h:84-103): storage to/from HBM (sizes in
 *    size_t; the reference computes them in `unsigned`);
 *  - brick_sweep / brick_sweeps / brick_ghost_refresh: what the kernels of
 *    stencils/3axis.hip.cpp and weak/main.hip.cpp did through the
 *    brick(...) code generator, now one call over a list of bricks.
 *  - the device path of the brick(...) hook of vecscatter.h: bricks
 *    queued by brick(...) call sites on device-resident fields run as one
 *    sweep at the next flush point (brick_flush(), or any function here).
 * Everything goes through the C-ABI of bk_api.h (link libbricks.so); this
 * header contains no GPU code and compiles with any C++17 host compiler.
 * Failures throw std::runtime_error with the library's message (the
 * reference's gpuCheck printed and exited, brick-gpu.h:27-33).
 */
#ifndef BK_BRICK_HIP_H
#define BK_BRICK_HIP_H

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "bk_api.h"
#include "brick.h"

/// Copy directions (the values of hipMemcpyHostToDevice / DeviceToHost).
enum { gpuMemcpyHostToDevice = 1, gpuMemcpyDeviceToHost = 2 };

inline void bkCheck(int rc, const char *what) {
  if (rc != BK_OK) throw std::runtime_error(std::string(what) + ": " + bk_last_error());
}

/// Throws unless libbricks.so speaks the C-ABI version these headers were
/// written for (a descriptor struct laid out for another version would be
/// misread).
inline void bkCheckVersion() {
  static const int v = bk_version();
  if (v != BK_API_VERSION)
    throw std::runtime_error("libbricks.so has C-ABI version " + std::to_string(v) +
                             ", the headers expect " + std::to_string(BK_API_VERSION) +
                             ": rebuild the program against the library's include/");
}

namespace bk_detail {

/// Throws unless p is device memory (a kernel given a host pointer would
/// fault the GPU).
inline void require_device(const void *p, const char *what) {
  int dev = 0;
  bkCheck(bk_pointer_is_device(p, &dev), what);
  if (!dev)
    throw std::invalid_argument(std::string(what) +
                                ": not device memory (movBrickInfo / movBrickStorage first)");
}

// Bricks queued by brick(...) call sites on device-resident fields, run as
// one sweep at the next flush point.
struct PendingBricks {
  std::mutex mu;
  bool active = false;
  int kind = -1;
  int ncoeff = 0;
  double coeff[125];
  const void *in = nullptr, *out = nullptr;
  size_t in_step = 0, out_step = 0;
  const unsigned *adj = nullptr;
  std::vector<unsigned> ids;
  std::vector<unsigned char> seen;
  void *d_ids = nullptr;
  size_t d_cap = 0;
};

inline PendingBricks &pending_bricks() {
  static PendingBricks p;
  return p;
}

inline void flush_locked(PendingBricks &p) {
  if (!p.active) return;
  p.active = false;
  if (p.ids.empty()) return;
  std::vector<unsigned> ids;
  ids.swap(p.ids);
  for (unsigned b : ids) p.seen[b] = 0;
  const size_t bytes = ids.size() * sizeof(unsigned);
  if (bytes > p.d_cap) {
    bkCheck(bk_free(p.d_ids), "brick(): queued bricks");
    p.d_ids = nullptr;
    p.d_cap = 0;
    bkCheck(bk_alloc(bytes, &p.d_ids), "brick(): queued bricks");
    p.d_cap = bytes;
  }
  // stream-ordered after the previous flush's sweep, which read d_ids
  bkCheck(bk_memcpy_h2d(p.d_ids, ids.data(), bytes, nullptr), "brick(): queued bricks");
  require_device(p.adj, "brick(): adjacency");
  require_device(p.in, "brick(): input field");
  require_device(p.out, "brick(): output field");
  bk_sweep_desc d;
  std::memset(&d, 0, sizeof(d));
  d.stencil = p.kind;
  d.dtype = bk_dtype_of<bElem>::value;
  d.d_adj = reinterpret_cast<const uint32_t *>(p.adj);
  d.d_list = static_cast<const uint32_t *>(p.d_ids);
  d.nlist = (uint32_t)ids.size();
  d.d_in = p.in;
  d.in_step = p.in_step;
  d.d_out = const_cast<void *>(p.out);
  d.out_step = p.out_step;
  d.h_coeff = p.coeff;
  d.cond_mode = BK_COND_REFERENCE;
  bkCheck(bk_sweep(&d, nullptr), "brick()");
}

// The device path of the brick(...) hook (installed below).
inline void device_brick(const bk_vs::Spec &sp, const bk_vs::Field &in, const bk_vs::Field &out,
                         const double *c, unsigned b) {
  if (in.ndims != 3 || in.bricksize != 512 || (in.veclen != 64 && in.veclen != 8))
    throw std::invalid_argument(
        "brick(): GPU sweeps need Brick<Dim<8,8,8>, Dim<8,8>> (or Dim<8>) fields");
  if (b >= in.nbricks) throw std::out_of_range("brick(): brick index beyond the BrickInfo");
  PendingBricks &p = pending_bricks();
  std::lock_guard<std::mutex> lock(p.mu);
  const bool same = p.active && p.kind == sp.kind && p.in == in.dat && p.out == out.dat &&
                    p.in_step == in.step && p.out_step == out.step && p.adj == in.adj &&
                    std::equal(c, c + sp.ncoeff, p.coeff);
  if (!same || (b < p.seen.size() && p.seen[b])) flush_locked(p);
  if (!p.active) {
    p.kind = sp.kind;
    p.ncoeff = sp.ncoeff;
    std::copy(c, c + sp.ncoeff, p.coeff);
    p.in = in.dat;
    p.out = out.dat;
    p.in_step = in.step;
    p.out_step = out.step;
    p.adj = in.adj;
    if (p.seen.size() < in.nbricks) p.seen.assign(in.nbricks, 0);
    p.active = true;
  }
  p.ids.push_back(b);
  p.seen[b] = 1;
}

struct DeviceHookInstaller {
  DeviceHookInstaller() { bk_vs::device_hook() = &device_brick; }
};
static DeviceHookInstaller device_hook_installer;

}  // namespace bk_detail

/// Run the bricks that brick(...) call sites queued on device-resident
/// fields (one sweep, stream-ordered on the null stream).  Every function
/// of this header that moves, sweeps or refreshes bricks flushes first.
inline void brick_flush() {
  bk_detail::PendingBricks &p = bk_detail::pending_bricks();
  std::lock_guard<std::mutex> lock(p.mu);
  bk_detail::flush_locked(p);
}

/// Adjacency list to (kind == gpuMemcpyHostToDevice) or from HBM.  The
/// returned BrickInfo owns a new copy: bk_free(adj) for a device one,
/// free(adj) for a host one.
template <unsigned dims>
BrickInfo<dims> movBrickInfo(BrickInfo<dims> &bInfo, int kind) {
  brick_flush();
  if (kind != gpuMemcpyHostToDevice && kind != gpuMemcpyDeviceToHost)
    throw std::invalid_argument("movBrickInfo: kind must be host-to-device or device-to-host");
  BrickInfo<dims> ret(0);
  std::free(ret.adj);
  ret.nbricks = bInfo.nbricks;
  const size_t bytes = (size_t)bInfo.nbricks * static_power<3, dims>::value * sizeof(unsigned);
  if (kind == gpuMemcpyHostToDevice) {
    void *d = nullptr;
    bkCheck(bk_alloc(bytes, &d), "movBrickInfo");
    bkCheck(bk_memcpy_h2d(d, bInfo.adj, bytes, nullptr), "movBrickInfo");
    ret.adj = static_cast<typename BrickInfo<dims>::adjlist>(d);
  } else {
    ret.adj = static_cast<typename BrickInfo<dims>::adjlist>(std::malloc(bytes));
    bkCheck(bk_memcpy_d2h(ret.adj, bInfo.adj, bytes, nullptr), "movBrickInfo");
  }
  return ret;
}

/// The BrickInfo struct itself moved with its adjacency list
/// (brick-gpu.h:64-76).  Host to device: a device copy of the struct whose
/// adj points to HBM (release both with bk_free).  Device to host: bInfo is
/// a host struct whose adj is in HBM (movBrickInfo's result); returns a
/// malloc'ed struct whose adj is a malloc'ed host copy (release both with
/// free).
template <unsigned dims>
BrickInfo<dims> *movBrickInfoDeep(BrickInfo<dims> &bInfo, int kind) {
  BrickInfo<dims> tmp = movBrickInfo(bInfo, kind);
  if (kind == gpuMemcpyDeviceToHost) {
    void *h = std::malloc(sizeof(BrickInfo<dims>));
    if (!h) throw std::bad_alloc();
    std::memcpy(h, &tmp, sizeof(BrickInfo<dims>));
    return static_cast<BrickInfo<dims> *>(h);
  }
  void *d = nullptr;
  bkCheck(bk_alloc(sizeof(BrickInfo<dims>), &d), "movBrickInfoDeep");
  bkCheck(bk_memcpy_h2d(d, &tmp, sizeof(BrickInfo<dims>), nullptr), "movBrickInfoDeep");
  return static_cast<BrickInfo<dims> *>(d);
}

/// A device-resident BrickStorage (HBM, 2048-byte aligned).
inline BrickStorage allocateDeviceStorage(long chunks, size_t step) {
  void *d = nullptr;
  bkCheck(bk_storage_alloc(bk_dtype_of<bElem>::value, chunks, step, &d), "allocateDeviceStorage");
  BrickStorage b;
  b.chunks = chunks;
  b.step = step;
  b.on_device = true;
  b.dat = std::shared_ptr<bElem>(static_cast<bElem *>(d), [](bElem *p) { bk_free(p); });
  return b;
}

/// Storage to or from HBM: a new storage of the same shape holding a copy.
inline BrickStorage movBrickStorage(BrickStorage &bStorage, int kind) {
  brick_flush();
  const size_t bytes = bStorage.step * (size_t)bStorage.chunks * sizeof(bElem);
  if (kind == gpuMemcpyHostToDevice) {
    BrickStorage r = allocateDeviceStorage(bStorage.chunks, bStorage.step);
    bkCheck(bk_memcpy_h2d(r.dat.get(), bStorage.dat.get(), bytes, nullptr), "movBrickStorage");
    return r;
  }
  if (kind == gpuMemcpyDeviceToHost) {
    BrickStorage r = BrickStorage::allocate(bStorage.chunks, bStorage.step);
    bkCheck(bk_memcpy_d2h(r.dat.get(), bStorage.dat.get(), bytes, nullptr), "movBrickStorage");
    return r;
  }
  throw std::invalid_argument("movBrickStorage: kind must be host-to-device or device-to-host");
}

/// A stencil selected by its reference file name with its coefficients
/// (layout of bk_api.h: star c[0] centre then (i+r,i-r,j+r,j-r,k+r,k-r);
/// cube c[(dk+2)*25+(dj+2)*5+(di+2)]).
struct BrickStencil {
  int kind = -1;
  std::vector<double> coeff;
  double theta = 0.0;
  int cond_mode = BK_COND_REFERENCE;
};

inline BrickStencil brick_stencil(const char *file, const double *coeff, double theta = 0.0,
                                  int cond_mode = BK_COND_REFERENCE) {
  BrickStencil s;
  s.kind = bk_stencil_from_file(file);
  if (s.kind < 0) throw std::invalid_argument(std::string("unknown stencil file ") + file);
  s.coeff.assign(coeff, coeff + bk_stencil_ncoeff(s.kind));
  s.theta = theta;
  s.cond_mode = cond_mode;
  return s;
}

namespace bk_detail {
template <typename BrickT>
void check_layout() {
  // the kernels address element (k,j,i) at k*64 + j*8 + i: 8x8x8 bricks in
  // the 8x8 (or 8) fold
  static_assert(BrickT::NDIMS == 3 && BrickT::BRICKSIZE == 512 &&
                    (BrickT::VECLEN == 64 || BrickT::VECLEN == 8),
                "GPU sweeps need Brick<Dim<8,8,8>, Dim<8,8>> (or Dim<8>)");
}

template <typename BrickT>
bk_sweep_desc make_desc(const BrickStencil &st, BrickT &bOut, BrickT &bIn,
                        const unsigned *d_list, unsigned nlist, bk_plan plan,
                        int lane_shift = 0) {
  bkCheckVersion();
  check_layout<BrickT>();
  if (!bIn.bStorage.on_device || !bOut.bStorage.on_device)
    throw std::invalid_argument("brick_sweep: fields must be device-resident (movBrickStorage)");
  require_device(bIn.bInfo->adj, "brick_sweep: adjacency");
  bk_sweep_desc d;
  std::memset(&d, 0, sizeof(d));
  d.stencil = st.kind;
  d.dtype = bk_dtype_of<bElem>::value;
  d.d_adj = reinterpret_cast<const uint32_t *>(bIn.bInfo->adj);
  d.d_list = d_list;
  d.nlist = nlist;
  d.d_in = bIn.dat;
  d.in_step = bIn.step;
  d.d_out = bOut.dat;
  d.out_step = bOut.step;
  d.h_coeff = st.coeff.data();
  d.theta = st.theta;
  d.cond_mode = st.cond_mode;
  d.plan = plan;
  d.lane_shift = lane_shift;
  return d;
}
}  // namespace bk_detail

/// One sweep bOut = stencil(bIn) over the nlist bricks of d_list (a device
/// array) or over a plan's bricks.  bIn.bInfo must come from movBrickInfo.
/// lane_shift = 1 with a BK_PLAN_QUAD plan of a star or cond stencil: the
/// i-direction taps from DPP lane shifts (the same bits).
template <typename BrickT>
void brick_sweep(const BrickStencil &st, BrickT &bOut, BrickT &bIn, const unsigned *d_list,
                 unsigned nlist, bk_stream s = nullptr, bk_plan plan = nullptr,
                 int lane_shift = 0) {
  brick_flush();
  bk_sweep_desc d = bk_detail::make_desc(st, bOut, bIn, d_list, nlist, plan, lane_shift);
  bkCheck(bk_sweep(&d, s), "brick_sweep");
}

/// nsweeps ping-pong sweeps a -> b -> a ..., each after the periodic ghost
/// refresh of its input (d_pairs: bk_ghost_pairs on the device).  The
/// result is in b for odd nsweeps, in a for even.
template <typename BrickT>
void brick_sweeps(const BrickStencil &st, BrickT &a, BrickT &b, const unsigned *d_list,
                  unsigned nlist, const unsigned *d_pairs, unsigned npairs, int nsweeps,
                  bool use_graph = false, bk_stream s = nullptr, bk_plan plan = nullptr,
                  int lane_shift = 0) {
  brick_flush();
  bk_sweep_desc d = bk_detail::make_desc(st, b, a, d_list, nlist, plan, lane_shift);
  bkCheck(bk_sweeps(&d, d_pairs, npairs, nsweeps, use_graph ? 1 : 0, s), "brick_sweeps");
}

/// Periodic ghost refresh of one device field.
template <typename BrickT>
void brick_ghost_refresh(BrickT &field, const unsigned *d_pairs, unsigned npairs,
                         bk_stream s = nullptr) {
  bk_detail::check_layout<BrickT>();
  brick_flush();
  bkCheck(bk_ghost_refresh(bk_dtype_of<bElem>::value, field.dat, field.step, BrickT::BRICKSIZE,
                           d_pairs, npairs, s),
          "brick_ghost_refresh");
}

/// Upload a host array of uint32 to HBM (for brick lists and ghost pairs).
inline unsigned *bk_upload_u32(const std::vector<unsigned> &v) {
  void *d = nullptr;
  if (v.empty()) return nullptr;
  bkCheck(bk_alloc(v.size() * sizeof(unsigned), &d), "bk_upload_u32");
  bkCheck(bk_memcpy_h2d(d, v.data(), v.size() * sizeof(unsigned), nullptr), "bk_upload_u32");
  return static_cast<unsigned *>(d);
}

#endif  // BK_BRICK_HIP_H
