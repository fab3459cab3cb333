// residency.hpp — which scenes stay resident in device memory under a byte budget.
//
// Host-only policy (no HIP), used by the serving loop: scenes are keyed by the caller's
// scene key and carry a device-memory size, an LRU stamp and a pin count (in-flight
// batches that read or build the scene). Admitting a scene that is not resident evicts
// least-recently-used unpinned scenes until it fits. A victim whose grid shape (n, levels)
// matches the newcomer's is handed over for reuse (its allocations fit as they are, so
// nothing is freed and no device synchronization happens); other victims are returned
// to be released. Nothing pinned, and nothing the caller protects, is ever evicted.
#pragma once
#include <stdint.h>

#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace vgi {

struct SceneShape {
  uint32_t n = 0, levels = 0;
  bool operator==(const SceneShape& o) const { return n == o.n && levels == o.levels; }
};

template <class Handle>
class Residency {
 public:
  struct Entry {
    Handle handle{};
    uint64_t bytes = 0;
    uint64_t last_use = 0;
    int pins = 0;
    bool building = false;  // a build of this scene is in flight
    SceneShape shape;
  };

  enum class Admit { kOk, kWait, kOverBudget };

  struct AdmitResult {
    Admit status = Admit::kWait;
    bool reused = false;           // handle is a recycled victim's (or the key's own)
    Handle handle{};
    std::vector<Handle> release;   // victims the caller must free (different shape)
    std::vector<uint64_t> evicted; // keys evicted by this admission
  };

  explicit Residency(uint64_t budget) : budget_(budget) {}

  Entry* find(uint64_t key) {
    auto it = map_.find(key);
    return it == map_.end() ? nullptr : &it->second;
  }

  void touch(uint64_t key) {
    if (Entry* e = find(key)) e->last_use = ++clock_;
  }

  // Makes `key` resident with `shape` and `bytes`. A key already resident with the same
  // shape keeps its handle (rebuild in place); with another shape it is replaced (its old
  // handle is returned for release). `protect` names keys that must not be evicted.
  AdmitResult admit(uint64_t key, SceneShape shape, uint64_t bytes, const std::unordered_set<uint64_t>& protect) {
    AdmitResult r;
    if (bytes > budget_) {
      r.status = Admit::kOverBudget;
      return r;
    }
    uint64_t own = 0;  // bytes of the key's current entry when it is replaced
    bool replace = false;
    if (Entry* e = find(key)) {
      if (e->pins > 0 || e->building) return r;  // in use: wait
      if (e->shape == shape) {
        r.status = Admit::kOk;
        r.reused = true;
        r.handle = e->handle;
        e->last_use = ++clock_;
        return r;
      }
      own = e->bytes;
      replace = true;
    }
    bool have = false;
    Handle reuse{};
    std::vector<uint64_t> victims;
    uint64_t freed = 0;
    // choose victims (without committing) until the newcomer fits
    while (used_ - own - freed + bytes > budget_) {
      const uint64_t v = lru_victim(key, protect, victims);
      if (v == kNone) return r;  // nothing evictable: wait (nothing was changed)
      victims.push_back(v);
      const Entry& e = map_.at(v);
      freed += e.bytes;
      if (!have && e.shape == shape) {  // its allocations carry over to the newcomer
        have = true;
        reuse = e.handle;
      }
    }
    if (replace) {
      r.release.push_back(map_.at(key).handle);
      used_ -= own;
      map_.erase(key);
    }
    for (uint64_t v : victims) {
      Entry& e = map_.at(v);
      if (!(have && e.handle == reuse)) r.release.push_back(e.handle);
      used_ -= e.bytes;
      map_.erase(v);
      r.evicted.push_back(v);
      evictions_++;
    }
    Entry n;
    n.bytes = bytes;
    n.shape = shape;
    n.last_use = ++clock_;
    if (have) {
      n.handle = reuse;
      r.reused = true;
      r.handle = reuse;
    }
    used_ += bytes;
    peak_ = used_ > peak_ ? used_ : peak_;
    map_.emplace(key, n);
    r.status = Admit::kOk;
    return r;
  }

  // Sets the handle of a newly admitted key (when admit() did not reuse one).
  void set_handle(uint64_t key, Handle h) {
    if (Entry* e = find(key)) e->handle = h;
  }

  // Drops a key (e.g. its build failed); returns its handle for release.
  bool remove(uint64_t key, Handle* h) {
    auto it = map_.find(key);
    if (it == map_.end()) return false;
    if (h) *h = it->second.handle;
    used_ -= it->second.bytes;
    map_.erase(it);
    return true;
  }

  template <class F>
  void for_each(F f) {
    for (auto& kv : map_) f(kv.first, kv.second);
  }

  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  uint64_t budget() const { return budget_; }
  uint64_t evictions() const { return evictions_; }
  size_t size() const { return map_.size(); }

 private:
  static constexpr uint64_t kNone = ~0ull;

  uint64_t lru_victim(uint64_t self, const std::unordered_set<uint64_t>& protect,
                      const std::vector<uint64_t>& taken) const {
    uint64_t best = kNone, stamp = ~0ull;
    for (const auto& kv : map_) {
      const Entry& e = kv.second;
      if (kv.first == self || e.pins > 0 || e.building || protect.count(kv.first)) continue;
      bool t = false;
      for (uint64_t k : taken) t = t || k == kv.first;
      if (t) continue;
      if (e.last_use < stamp) {
        stamp = e.last_use;
        best = kv.first;
      }
    }
    return best;
  }

  std::unordered_map<uint64_t, Entry> map_;
  uint64_t budget_;
  uint64_t used_ = 0, peak_ = 0, clock_ = 0, evictions_ = 0;
};

}  // namespace vgi
