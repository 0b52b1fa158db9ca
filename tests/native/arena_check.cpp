// Randomised check of the serving arena (pyloworder_amd/csrc/arena.hpp)
// against a plain model: every block inside the arena, granule aligned, no
// two live blocks overlapping, an allocation failing only when no free range
// can hold it, free ranges disjoint and coalesced, both indexes of the free
// ranges in step, and the arena whole again once every block is released.
// Built with g++ -fsanitize=address,undefined by tests/test_arena.py.
#include <cstdio>
#include <random>
#include <vector>

#include "arena.hpp"

using pcserve::Arena;

static int fail(const char* what, int trial, int step) {
  std::printf("FAIL %s (trial %d, step %d)\n", what, trial, step);
  return 1;
}

int main() {
  std::mt19937_64 rng(20261203);
  char* const base = reinterpret_cast<char*>(uintptr_t(1) << 40);  // never dereferenced
  for (int trial = 0; trial < 300; trial++) {
    Arena a;
    const int64_t size = int64_t(1 + rng() % 1000) * Arena::kGran + int64_t(rng() % Arena::kGran);
    a.init(base, size);
    if (a.size != size / Arena::kGran * Arena::kGran) return fail("size rounding", trial, -1);
    std::vector<std::pair<int64_t, int64_t>> live;  // offset, length
    for (int step = 0; step < 2000; step++) {
      if (live.empty() || rng() % 2) {
        const int64_t want = int64_t(rng() % uint64_t(a.size / 4 + 2)) - 1;  // includes 0 and -1
        const int64_t len = Arena::round(want);
        bool fits = false;
        for (auto& f : a.free_at) fits |= f.second >= len;
        void* p = a.alloc(want);
        if (!p) {
          if (fits) return fail("allocation refused although a free range fits", trial, step);
          continue;
        }
        const int64_t off = static_cast<char*>(p) - base;
        if (off < 0 || off % Arena::kGran || off + len > a.size) return fail("block outside the arena", trial, step);
        for (auto& b : live)
          if (off < b.first + b.second && b.first < off + len) return fail("overlapping blocks", trial, step);
        live.push_back({off, len});
      } else {
        const size_t i = rng() % live.size();
        a.release(base + live[i].first);
        live.erase(live.begin() + long(i));
      }
      int64_t total = 0, prev_end = -1;
      for (auto& f : a.free_at) {
        if (f.first <= prev_end) return fail(f.first == prev_end ? "adjacent free ranges" : "free ranges overlap", trial, step);
        prev_end = f.first + f.second;
        total += f.second;
      }
      for (auto& u : a.used) total += u.second;
      if (total != a.size) return fail("free + used != size", trial, step);
      if (a.free_len.size() != a.free_at.size() || a.used.size() != live.size()) return fail("index sizes", trial, step);
      for (auto& f : a.free_len) {
        auto it = a.free_at.find(f.second);
        if (it == a.free_at.end() || it->second != f.first) return fail("length index out of step", trial, step);
      }
    }
    a.release(nullptr);            // ignored
    a.release(base + a.size + 8);  // not a block: ignored
    for (auto& b : live) a.release(base + b.first);
    if (a.size > 0 && (a.free_at.size() != 1 || a.free_at.begin()->second != a.size))
      return fail("arena not whole after releasing everything", trial, -1);
  }
  std::printf("arena check passed\n");
  return 0;
}
