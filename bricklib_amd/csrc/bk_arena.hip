// Field arena: device memory for grids that come and go (a serving loop
// admitting and retiring requests) without a driver call per grid.
//
// The arena owns slabs obtained with hipMalloc (each at least slab_bytes,
// or one request if larger) and carves blocks out of them best-fit; a freed
// block merges with free neighbours of its slab at once.  Slabs stay until
// bk_arena_destroy.  All bookkeeping is on the host, under a mutex.
//
// Reuse is stream-ordered by the caller's discipline: a block freed while
// kernels that use it are queued on a stream may be handed out again at
// once, so the next user must issue its work on the same stream (or wait
// for that stream first) -- the same rule as hipFreeAsync/hipMallocAsync
// on one stream, without their per-call cost.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

#include "bk_api.h"

namespace bk {
int plan_fail(int code, const char *msg);
}

namespace {

constexpr size_t kAlign = 2048;  // block alignment (a brick of FP32 pairs)

size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct Slab {
  char *base = nullptr;
  size_t bytes = 0;
  std::map<size_t, size_t> free_by_off;  // offset -> size
};

struct Block {
  uint32_t slab;
  size_t off, size;
};

}  // namespace

struct bk_arena_s {
  int device = 0;
  size_t slab_bytes = 0;
  size_t capacity = 0;  // limit on the slabs' total (0: no limit)
  size_t reserved = 0;  // slabs' total
  size_t used = 0;      // bytes in live blocks
  std::vector<Slab> slabs;
  std::multimap<size_t, std::pair<uint32_t, size_t>> free_by_size;  // size -> (slab, off)
  std::map<char *, Block> live;
  std::mutex mu;

  void insert_free(uint32_t s, size_t off, size_t size) {
    slabs[s].free_by_off.emplace(off, size);
    free_by_size.emplace(size, std::make_pair(s, off));
  }
  void erase_free(uint32_t s, size_t off, size_t size) {
    slabs[s].free_by_off.erase(off);
    auto range = free_by_size.equal_range(size);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second.first == s && it->second.second == off) {
        free_by_size.erase(it);
        return;
      }
  }
};

namespace {

int fail(int code, const char *msg) { return bk::plan_fail(code, msg); }

}  // namespace

extern "C" {

int bk_arena_create(size_t slab_bytes, size_t capacity, bk_arena *out) {
  if (!out) return fail(BK_ERR_ARG, "bk_arena_create: null out pointer");
  *out = nullptr;
  if (slab_bytes == 0) return fail(BK_ERR_ARG, "bk_arena_create: slab_bytes must be positive");
  auto *a = new bk_arena_s();
  (void)hipGetDevice(&a->device);
  a->slab_bytes = round_up(slab_bytes);
  a->capacity = capacity;
  *out = a;
  return BK_OK;
}

int bk_arena_destroy(bk_arena a) {
  if (!a) return BK_OK;
  for (auto &s : a->slabs) (void)hipFree(s.base);
  delete a;
  return BK_OK;
}

int bk_arena_alloc(bk_arena a, size_t bytes, void **d_ptr) {
  if (!a || !d_ptr) return fail(BK_ERR_ARG, "bk_arena_alloc: null argument");
  *d_ptr = nullptr;
  if (bytes == 0) return BK_OK;
  const size_t need = round_up(bytes);
  std::lock_guard<std::mutex> lock(a->mu);
  auto it = a->free_by_size.lower_bound(need);
  if (it == a->free_by_size.end()) {
    // grow: one new slab, at least slab_bytes
    const size_t sb = need > a->slab_bytes ? need : a->slab_bytes;
    if (a->capacity && a->reserved + sb > a->capacity)
      return fail(BK_ERR_NOMEM, "bk_arena_alloc: arena capacity reached");
    Slab s;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&s.base), sb);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(e == hipErrorOutOfMemory ? BK_ERR_NOMEM : BK_ERR_HIP,
                  "bk_arena_alloc: slab allocation failed");
    }
    s.bytes = sb;
    a->slabs.push_back(std::move(s));
    a->reserved += sb;
    const uint32_t si = (uint32_t)(a->slabs.size() - 1);
    a->insert_free(si, 0, sb);
    it = a->free_by_size.lower_bound(need);
  }
  const size_t fsize = it->first;
  const uint32_t si = it->second.first;
  const size_t off = it->second.second;
  a->erase_free(si, off, fsize);
  if (fsize > need) a->insert_free(si, off + need, fsize - need);
  char *p = a->slabs[si].base + off;
  a->live.emplace(p, Block{si, off, need});
  a->used += need;
  *d_ptr = p;
  return BK_OK;
}

int bk_arena_free(bk_arena a, void *d_ptr) {
  if (!a) return fail(BK_ERR_ARG, "bk_arena_free: null arena");
  if (!d_ptr) return BK_OK;
  std::lock_guard<std::mutex> lock(a->mu);
  auto lv = a->live.find(static_cast<char *>(d_ptr));
  if (lv == a->live.end()) return fail(BK_ERR_ARG, "bk_arena_free: not a live block of this arena");
  Block b = lv->second;
  a->live.erase(lv);
  a->used -= b.size;
  Slab &s = a->slabs[b.slab];
  size_t off = b.off, size = b.size;
  // merge with the free block after, then the one before
  auto next = s.free_by_off.lower_bound(off);
  if (next != s.free_by_off.end() && next->first == off + size) {
    const size_t ns = next->second;
    a->erase_free(b.slab, next->first, ns);
    size += ns;
  }
  auto after = s.free_by_off.lower_bound(off);
  if (after != s.free_by_off.begin()) {
    auto prev = std::prev(after);
    if (prev->first + prev->second == off) {
      const size_t po = prev->first, ps = prev->second;
      a->erase_free(b.slab, po, ps);
      off = po;
      size += ps;
    }
  }
  a->insert_free(b.slab, off, size);
  return BK_OK;
}

int bk_arena_stats(bk_arena a, size_t *used, size_t *reserved, size_t *largest_free,
                   uint32_t *live_blocks) {
  if (!a) return fail(BK_ERR_ARG, "bk_arena_stats: null arena");
  std::lock_guard<std::mutex> lock(a->mu);
  if (used) *used = a->used;
  if (reserved) *reserved = a->reserved;
  if (largest_free) *largest_free = a->free_by_size.empty() ? 0 : a->free_by_size.rbegin()->first;
  if (live_blocks) *live_blocks = (uint32_t)a->live.size();
  return BK_OK;
}

}  // extern "C"
