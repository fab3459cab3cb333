// Unit test of the serving loop's residency policy (final184_amd/csrc/residency.hpp):
// LRU order, budget, pins and protection, same-shape recycling, over-budget refusal,
// waiting without side effects, rebuild in place, and randomized invariants.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <unordered_set>

#include "../../final184_amd/csrc/residency.hpp"

using R = vgi::Residency<int>;
using vgi::SceneShape;

#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::printf("CHECK failed line %d: %s\n", __LINE__, #c);    \
      std::exit(1);                                               \
    }                                                             \
  } while (0)

static const SceneShape S1{128, 7}, S2{256, 8};
static const std::unordered_set<uint64_t> kNone;

int main() {
  {  // LRU eviction, release of another shape, recycling of the same shape
    R r(100);
    auto a = r.admit(1, S1, 40, kNone);
    CHECK(a.status == R::Admit::kOk && !a.reused);
    r.set_handle(1, 11);
    auto b = r.admit(2, S1, 40, kNone);
    r.set_handle(2, 12);
    CHECK(b.status == R::Admit::kOk && r.used() == 80);
    auto c = r.admit(3, S2, 40, kNone);  // evicts 1 (LRU), another shape: released
    CHECK(c.status == R::Admit::kOk && !c.reused);
    CHECK(c.release.size() == 1 && c.release[0] == 11 && c.evicted.size() == 1 && c.evicted[0] == 1);
    r.set_handle(3, 13);
    CHECK(r.used() == 80 && r.size() == 2 && !r.find(1));
    r.touch(2);  // 3 is now LRU
    auto d = r.admit(4, S2, 40, kNone);  // evicts 3, same shape: recycled
    CHECK(d.status == R::Admit::kOk && d.reused && d.handle == 13 && d.release.empty());
    CHECK(r.find(4)->handle == 13 && r.used() == 80 && r.evictions() == 2);
  }
  {  // pins and protection are never evicted; waiting changes nothing
    R r(100);
    r.admit(1, S1, 50, kNone);
    r.set_handle(1, 1);
    r.admit(2, S1, 50, kNone);
    r.set_handle(2, 2);
    r.find(1)->pins = 1;
    std::unordered_set<uint64_t> prot{2};
    auto w = r.admit(3, S1, 50, prot);
    CHECK(w.status == R::Admit::kWait && w.release.empty() && w.evicted.empty());
    CHECK(r.size() == 2 && r.used() == 100 && !r.find(3));
    auto ok = r.admit(3, S1, 50, kNone);  // 2 is the only candidate
    CHECK(ok.status == R::Admit::kOk && ok.handle == 2 && ok.evicted[0] == 2);
    r.find(1)->pins = 0;
    r.find(1)->building = true;  // a build in flight is not evictable either
    r.find(3)->pins = 1;
    CHECK(r.admit(4, S1, 50, kNone).status == R::Admit::kWait);
  }
  {  // over budget; rebuild in place; rebuild with another shape
    R r(100);
    CHECK(r.admit(9, S1, 101, kNone).status == R::Admit::kOverBudget && r.size() == 0);
    r.admit(1, S1, 60, kNone);
    r.set_handle(1, 5);
    auto same = r.admit(1, S1, 60, kNone);
    CHECK(same.status == R::Admit::kOk && same.reused && same.handle == 5 && r.used() == 60);
    auto other = r.admit(1, S2, 90, kNone);
    CHECK(other.status == R::Admit::kOk && !other.reused && other.release.size() == 1 && other.release[0] == 5);
    CHECK(r.used() == 90 && r.size() == 1);
    r.find(1)->pins = 1;
    CHECK(r.admit(1, S1, 60, kNone).status == R::Admit::kWait && r.find(1)->bytes == 90);
  }
  {  // randomized: used == sum of entries <= budget, pinned entries survive, handles unique
    std::mt19937_64 rng(184);
    R r(1000);
    int next_handle = 1;
    for (int step = 0; step < 20000; step++) {
      const uint64_t key = rng() % 40;
      const SceneShape sh = (rng() & 1) ? S1 : S2;
      const uint64_t bytes = sh.n == 128 ? 100 : 300;
      std::map<uint64_t, int> pinned;
      r.for_each([&](uint64_t k, R::Entry& e) {
        if (e.pins) pinned[k] = e.handle;
      });
      auto a = r.admit(key, sh, bytes, kNone);
      if (a.status == R::Admit::kOk && !a.reused) r.set_handle(key, next_handle++);
      for (auto& kv : pinned) CHECK(r.find(kv.first) && r.find(kv.first)->handle == kv.second);
      uint64_t sum = 0;
      std::unordered_set<int> handles;
      r.for_each([&](uint64_t, R::Entry& e) {
        sum += e.bytes;
        CHECK(handles.insert(e.handle).second);
      });
      CHECK(sum == r.used() && r.used() <= r.budget());
      // pin / unpin at random
      r.for_each([&](uint64_t, R::Entry& e) {
        if (rng() % 7 == 0) e.pins = e.pins ? 0 : 1;
      });
    }
    CHECK(r.evictions() > 100 && r.peak() <= 1000);
  }
  std::printf("RESIDENCY OK\n");
  return 0;
}
