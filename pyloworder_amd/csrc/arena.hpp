// Sub-allocator of the native request loop (serve.cpp): carves blocks out of
// one device allocation made at server creation.  Host-only bookkeeping (no
// HIP calls), so it is also compiled and checked on its own on the CPU.
#pragma once
#include <cstdint>
#include <map>

namespace pcserve {

// Sub-allocator over one device allocation.  Free ranges are kept twice: by
// offset, to coalesce a freed block with its neighbours, and by length, for
// the best fit (the shortest free range that holds the block).  Every block
// is a whole number of granules, so all of them are 4 KiB aligned.
struct Arena {
  static constexpr int64_t kGran = 4096;
  char* base = nullptr;
  int64_t size = 0;
  std::map<int64_t, int64_t> free_at;         // offset -> length
  std::multimap<int64_t, int64_t> free_len;   // length -> offset
  std::map<int64_t, int64_t> used;            // offset -> length

  static int64_t round(int64_t bytes) { return bytes <= 0 ? kGran : (bytes + kGran - 1) / kGran * kGran; }

  void init(void* p, int64_t bytes) {
    base = static_cast<char*>(p);
    size = bytes / kGran * kGran;
    free_at.clear();
    free_len.clear();
    used.clear();
    if (size > 0) add_free(0, size);
  }
  void add_free(int64_t off, int64_t len) {
    free_at[off] = len;
    free_len.emplace(len, off);
  }
  void drop_free(int64_t off, int64_t len) {
    free_at.erase(off);
    auto r = free_len.equal_range(len);
    for (auto it = r.first; it != r.second; ++it)
      if (it->second == off) {
        free_len.erase(it);
        return;
      }
  }
  void* alloc(int64_t bytes) {
    const int64_t len = round(bytes);
    auto it = free_len.lower_bound(len);
    if (it == free_len.end()) return nullptr;
    const int64_t off = it->second, have = it->first;
    drop_free(off, have);
    if (have > len) add_free(off + len, have - len);
    used[off] = len;
    return base + off;
  }
  void release(void* p) {
    if (!p) return;
    auto u = used.find(static_cast<char*>(p) - base);
    if (u == used.end()) return;
    int64_t off = u->first, len = u->second;
    used.erase(u);
    auto nx = free_at.find(off + len);  // the free range right after
    if (nx != free_at.end()) {
      const int64_t nlen = nx->second;
      drop_free(nx->first, nlen);
      len += nlen;
    }
    auto pv = free_at.lower_bound(off);  // the free range right before
    if (pv != free_at.begin()) {
      --pv;
      if (pv->first + pv->second == off) {
        const int64_t poff = pv->first, plen = pv->second;
        drop_free(poff, plen);
        off = poff;
        len += plen;
      }
    }
    add_free(off, len);
  }
};

}  // namespace pcserve
