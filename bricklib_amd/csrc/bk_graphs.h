// bk_graphs.h -- a bounded cache of instantiated hipGraphs.
//
// Every launch of a cached executable graph is followed by an event recorded
// on the same stream.  An entry is destroyed only after that event has
// completed, so a graph that is still queued or running (a pipelined caller
// may launch the next run before the previous one ends) is never freed under
// the device.  Eviction is least-recently-used, one entry at a time.  Each
// entry carries up to two tags (plan ids) so the owner of a tagged resource
// can drop the entries that captured it before freeing it.
//
// Not thread-safe by itself: callers hold their own lock.  There is no
// destructor work: caches that live until process exit are left to the
// runtime's teardown (destroying graphs after it has shut down would fault).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <utility>

namespace bk {

template <class Key>
class ExecCache {
 public:
  explicit ExecCache(size_t capacity) : cap_(capacity ? capacity : 1) {}

  // The cached graph for k, or nullptr.
  hipGraphExec_t find(const Key &k) {
    auto it = map_.find(k);
    if (it == map_.end()) return nullptr;
    it->second.last_use = ++clock_;
    return it->second.exec;
  }

  // Take ownership of ge under k, evicting the least recently used entry
  // first when the cache is full.
  hipError_t insert(const Key &k, hipGraphExec_t ge, uint64_t tag0 = 0, uint64_t tag1 = 0) {
    while (map_.size() >= cap_) {
      auto lru = map_.begin();
      for (auto it = map_.begin(); it != map_.end(); ++it)
        if (it->second.last_use < lru->second.last_use) lru = it;
      hipError_t e = destroy(lru->second);
      map_.erase(lru);
      if (e != hipSuccess) {
        (void)hipGraphExecDestroy(ge);
        return e;
      }
    }
    Entry en;
    en.exec = ge;
    en.tag0 = tag0;
    en.tag1 = tag1;
    en.last_use = ++clock_;
    hipError_t e = hipEventCreateWithFlags(&en.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipGraphExecDestroy(ge);
      return e;
    }
    map_.emplace(k, en);
    return hipSuccess;
  }

  // Launch the graph cached under k on s and mark the launch's end.
  hipError_t launch(const Key &k, hipStream_t s) {
    auto it = map_.find(k);
    if (it == map_.end()) return hipErrorInvalidValue;
    hipError_t e = hipGraphLaunch(it->second.exec, s);
    if (e != hipSuccess) return e;
    it->second.launched = true;
    return hipEventRecord(it->second.done, s);
  }

  // Drop every entry tagged with `tag` (after its last launch completed).
  hipError_t forget(uint64_t tag) {
    hipError_t first = hipSuccess;
    for (auto it = map_.begin(); it != map_.end();) {
      if (tag != 0 && (it->second.tag0 == tag || it->second.tag1 == tag)) {
        hipError_t e = destroy(it->second);
        if (first == hipSuccess) first = e;
        it = map_.erase(it);
      } else {
        ++it;
      }
    }
    return first;
  }

  hipError_t clear() {
    hipError_t first = hipSuccess;
    for (auto &kv : map_) {
      hipError_t e = destroy(kv.second);
      if (first == hipSuccess) first = e;
    }
    map_.clear();
    return first;
  }

  size_t size() const { return map_.size(); }

 private:
  struct Entry {
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    uint64_t tag0 = 0, tag1 = 0;
    uint64_t last_use = 0;
    bool launched = false;
  };

  static hipError_t destroy(Entry &en) {
    hipError_t e = hipSuccess;
    if (en.launched) e = hipEventSynchronize(en.done);
    (void)hipEventDestroy(en.done);
    hipError_t e2 = hipGraphExecDestroy(en.exec);
    return e != hipSuccess ? e : e2;
  }

  size_t cap_;
  uint64_t clock_ = 0;
  std::map<Key, Entry> map_;
};

}  // namespace bk
