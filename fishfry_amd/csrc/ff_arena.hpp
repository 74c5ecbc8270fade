// Host-side pieces of the serving loop that need no GPU: the first-fit
// arena that sub-allocates the server's resident device block, and the
// size classes requests are grouped by.  Kept free of HIP so the CPU tests
// exercise them directly.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <iterator>
#include <map>

namespace ff {

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// First-fit allocator over [0, size) with coalescing free list.
class Arena {
 public:
  void reset(size_t size) {
    size_ = size;
    free_.clear();
    sizes_.clear();
    free_[0] = size;
    used_ = 0;
    peak_ = 0;
  }
  bool alloc(size_t bytes, size_t& off) {
    bytes = align_up(std::max<size_t>(bytes, 256), 256);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        const size_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        sizes_[off] = bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off) {
    auto s = sizes_.find(off);
    if (s == sizes_.end()) return;
    size_t bytes = s->second;
    sizes_.erase(s);
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first == off + bytes) {
      bytes += next->second;
      free_.erase(next);
    }
    auto it = free_.lower_bound(off);
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }

 private:
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;   // offset -> length
  std::map<size_t, size_t> sizes_;  // allocated offset -> length
};

// Size class of a request: precision, refinement steps and the octave of its
// mean edge (log2 of the point count / 3), so a group holds solves of similar
// cost and memory.
inline int size_class(int precision, int refine_steps, const int n[3]) {
  const double pts = double(n[0]) * n[1] * n[2];
  const int lg = (int)std::floor(std::log2(std::max(1.0, pts)) / 3.0);
  return (precision * 16 + std::min(refine_steps, 15)) * 64 + lg;
}

}  // namespace ff
