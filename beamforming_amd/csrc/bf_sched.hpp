// bf_sched.hpp -- device-independent parts of the serving loop: the
// sub-allocator that hands out pieces of the server's HBM arena, and the
// admission queue that decides which waiting requests may take memory and a
// slot. Plain C++ (no HIP), so the policy is testable on a host without a GPU.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <vector>

namespace bf {

// First-fit allocator over [0, capacity) with coalescing of free neighbours.
// Offsets and sizes are multiples of `align`.
class ArenaAllocator {
public:
  explicit ArenaAllocator(size_t capacity = 0, size_t align = 256) { reset(capacity, align); }

  void reset(size_t capacity, size_t align = 256) {
    align_ = align ? align : 1;
    capacity_ = capacity / align_ * align_;
    free_.clear();
    used_.clear();
    if (capacity_) free_[0] = capacity_;
    in_use_ = peak_ = 0;
  }

  size_t round(size_t bytes) const { return (bytes + align_ - 1) / align_ * align_; }

  // Returns false when no free range is large enough (nothing changes).
  bool alloc(size_t bytes, size_t *offset) {
    const size_t need = round(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > need) free_[off + need] = len - need;
      used_[off] = need;
      in_use_ += need;
      if (in_use_ > peak_) peak_ = in_use_;
      *offset = off;
      return true;
    }
    return false;
  }

  // Returns false for an offset that is not allocated.
  bool release(size_t offset) {
    auto u = used_.find(offset);
    if (u == used_.end()) return false;
    size_t off = u->first, len = u->second;
    used_.erase(u);
    in_use_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return true;
      }
    }
    free_[off] = len;
    return true;
  }

  size_t capacity() const { return capacity_; }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }
  size_t largest_free() const {
    size_t m = 0;
    for (const auto &f : free_) m = f.second > m ? f.second : m;
    return m;
  }
  size_t free_ranges() const { return free_.size(); }

private:
  size_t align_ = 256, capacity_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_; // offset -> length
  std::map<size_t, size_t> used_; // offset -> length
};

// What one request needs from the server: bytes of arena (one or two
// pieces) and a slot.
struct AdmitNeed {
  size_t bytes[2] = {0, 0}; // piece sizes; 0 = no piece
};
struct AdmitGrant {
  size_t offset[2] = {0, 0};
};

// First-in first-out admission: the oldest waiting request is admitted when
// a slot is free and the arena has room for all of its pieces; younger
// requests never overtake it (no starvation of large chunks). admit() is
// called whenever memory or slots are returned.
template <class Id> class AdmissionQueue {
public:
  AdmissionQueue(ArenaAllocator *arena, int max_slots) : arena_(arena), max_slots_(max_slots) {}

  void push(Id id, const AdmitNeed &need) { waiting_.push_back({id, need}); }
  size_t waiting() const { return waiting_.size(); }
  int slots_used() const { return slots_used_; }

  // Admits as many requests as fit, in order; calls grant(id, AdmitGrant).
  template <class F> int admit(F &&grant) {
    int n = 0;
    while (!waiting_.empty() && slots_used_ < max_slots_) {
      const Entry &e = waiting_.front();
      AdmitGrant g;
      bool ok = true;
      int got = 0;
      for (int i = 0; i < 2 && ok; ++i) {
        if (!e.need.bytes[i]) continue;
        ok = arena_->alloc(e.need.bytes[i], &g.offset[i]);
        if (ok) ++got;
      }
      if (!ok) { // undo partial grants; wait for memory
        for (int i = 0, k = 0; i < 2 && k < got; ++i)
          if (e.need.bytes[i]) {
            arena_->release(g.offset[i]);
            ++k;
          }
        break;
      }
      ++slots_used_;
      const Id id = e.id;
      waiting_.pop_front();
      grant(id, g);
      ++n;
    }
    return n;
  }

  // A finished request gives back its slot (its memory is released to the
  // arena separately, when the caller is done with the output).
  void finish_slot() {
    if (slots_used_ > 0) --slots_used_;
  }

private:
  struct Entry {
    Id id;
    AdmitNeed need;
  };
  ArenaAllocator *arena_;
  int max_slots_;
  int slots_used_ = 0;
  std::deque<Entry> waiting_;
};

// Nearest-rank percentile (q in [0, 1]) of a sample, sorted in place.
inline double percentile(std::vector<double> &v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)(q * (double)v.size() + 0.999999999);
  if (k < 1) k = 1;
  if (k > v.size()) k = v.size();
  return v[k - 1];
}

} // namespace bf
