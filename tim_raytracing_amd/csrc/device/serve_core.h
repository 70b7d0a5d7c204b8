// Host-only pieces of the served form: the device-arena range allocator and the admission
// policy. Kept free of HIP so they can be unit-tested on a machine without a GPU.
#pragma once
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>

namespace rtq {

// First-fit allocator over [0, capacity) in 256-byte granules; free ranges are coalesced.
class RangeAllocator {
 public:
  static constexpr uint64_t kGranule = 256;
  static uint64_t round(uint64_t b) { return (b + kGranule - 1) & ~(kGranule - 1); }

  void reset(uint64_t capacity) {
    free_.clear();
    capacity_ = capacity & ~(kGranule - 1);
    used_ = 0;
    if (capacity_) free_[0] = capacity_;
  }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = round(bytes);
    if (bytes == 0) bytes = kGranule;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes, start = it->first + bytes;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = round(bytes);
    if (bytes == 0) bytes = kGranule;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t capacity() const { return capacity_; }
  uint64_t used() const { return used_; }
  size_t fragments() const { return free_.size(); }

 private:
  std::map<uint64_t, uint64_t> free_;
  uint64_t capacity_ = 0, used_ = 0;
};

// Device bytes a request needs: rays (32 B), results (16 B closest / 4 B any) and, for closest
// hits, leaf node ids (4 B), each sub-buffer 256-byte aligned.
inline uint64_t requestBytes(uint32_t numRays, bool closest) {
  const uint64_t n = numRays;
  return RangeAllocator::round(32 * n) + RangeAllocator::round((closest ? 16 : 4) * n) +
         (closest ? RangeAllocator::round(4 * n) : 0);
}

// FIFO admission: admit from the head while the head fits; a request that does not fit blocks the
// ones behind it (no overtaking, so a large request is never starved by small ones).
template <class Req, class Fits>
void admitFifo(std::deque<Req>& queue, Fits&& tryAdmit) {
  while (!queue.empty()) {
    if (!tryAdmit(queue.front())) break;
    queue.pop_front();
  }
}

}  // namespace rtq
