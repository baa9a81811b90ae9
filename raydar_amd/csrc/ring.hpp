// Ring allocator for the radiance arena of a streaming engine.
//
// Tickets (passes) take contiguous slot ranges in issue order and release
// them in any order; a range's space is reused once it and every older range
// are released.  Live ranges never overlap.
#pragma once

#include <cstdint>
#include <deque>

namespace rdh {

class RadianceRing {
 public:
  explicit RadianceRing(uint64_t capacity = 0) : cap_(capacity) {}
  void reset(uint64_t capacity) {
    cap_ = capacity;
    live_.clear();
  }
  uint64_t capacity() const { return cap_; }
  bool empty() const { return live_.empty(); }

  // Reserves n slots for `ticket`; false when they do not fit now.
  bool alloc(uint64_t n, uint32_t ticket, uint64_t *off) {
    if (n == 0 || n > cap_) return false;
    uint64_t at;
    if (live_.empty()) {
      at = 0;
    } else {
      const Range &oldest = live_.front(), &newest = live_.back();
      const uint64_t tail = oldest.off, head = newest.off + newest.size;
      if (newest.off >= oldest.off) {
        // not wrapped: in use [tail, head); free [head, cap) and [0, tail)
        if (cap_ - head >= n) at = head;
        else if (tail >= n) at = 0;
        else return false;
      } else {
        // wrapped: in use [tail, cap) and [0, head); free [head, tail)
        if (tail - head >= n) at = head;
        else return false;
      }
    }
    live_.push_back(Range{at, n, ticket, false});
    *off = at;
    return true;
  }

  // Marks the ticket's range free; its space returns once older ranges have.
  void release(uint32_t ticket) {
    for (Range &r : live_)
      if (r.ticket == ticket) r.freed = true;
    while (!live_.empty() && live_.front().freed) live_.pop_front();
  }

 private:
  struct Range {
    uint64_t off, size;
    uint32_t ticket;
    bool freed;
  };
  uint64_t cap_;
  std::deque<Range> live_;
};

}  // namespace rdh
