// Sub-allocator of the serving loop's device budget (host-side bookkeeping
// only: offsets into one allocation made at server start).
#pragma once
#include <cstddef>
#include <iterator>
#include <map>

namespace tsw {

// First-fit sub-allocator over one device allocation made at server start:
// per-request hipMalloc / hipFree of multi-GB snapshot arenas would stall the
// serving loop for milliseconds each.  Free blocks are coalesced on release.
// A request's working block is carved first-fit from the bottom and its result
// block (kept until the caller fetches it) last-fit from the top: results of
// finished requests left between working blocks would fragment the space, and
// a queued request that no gap can hold would wait with nothing running.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  static size_t up(size_t b) { return (b + 255) & ~(size_t)255; }
  bool fits(size_t bytes) const {
    bytes = up(bytes);
    for (const auto& f : free_)
      if (f.second >= bytes) return true;
    return false;
  }
  char* alloc(size_t bytes) {
    bytes = up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      return base + off;
    }
    return nullptr;
  }
  // Last fit, from the top of the highest free block that holds `bytes`.
  // Long-lived blocks (results kept until fetched) go to the top so that they
  // do not split the space the short-lived working blocks are carved from.
  char* alloc_high(size_t bytes) {
    bytes = up(bytes);
    for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rest = it->second - bytes;
      free_.erase(std::next(it).base());
      if (rest) free_[off] = rest;
      return base + off + rest;
    }
    return nullptr;
  }
  void release(char* p, size_t bytes) {
    if (!p) return;
    bytes = up(bytes);
    size_t off = (size_t)(p - base);
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + bytes == nx->first) {  // merge with the next block
      bytes += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {  // merge with the previous block
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
};

}  // namespace tsw
