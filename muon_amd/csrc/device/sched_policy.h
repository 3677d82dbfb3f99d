// Host-side policies of the scheduler that do not touch the device: the
// first-fit range allocator behind the film pool (admission by HBM
// footprint) and the choice of the next slice of samples to hand to the
// engine. Plain C++, checked on the CPU.
#ifndef MUONX_DEVICE_SCHED_POLICY_H_
#define MUONX_DEVICE_SCHED_POLICY_H_

#include <cstdint>
#include <iterator>
#include <map>
#include <vector>

namespace muonx {

// First-fit allocator of aligned byte ranges in [0, capacity), with free
// ranges coalesced on release.
class RangeAllocator {
 public:
  static constexpr uint64_t kAlign = 256;
  static uint64_t Round(uint64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
  void Reset(uint64_t capacity) {
    cap_ = Round(capacity);
    used_ = 0;
    free_.clear();
    if (cap_) free_[0] = cap_;
  }
  uint64_t capacity() const { return cap_; }
  uint64_t in_use() const { return used_; }
  // Returns false when no free range is large enough.
  bool Alloc(uint64_t bytes, uint64_t *offset) {
    bytes = Round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *offset = it->first;
      const uint64_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void Free(uint64_t offset, uint64_t bytes) {
    bytes = Round(bytes);
    used_ -= bytes;
    auto it = free_.emplace(offset, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {  // merge with the following range
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {  // merge with the preceding range
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  const std::map<uint64_t, uint64_t> &free_ranges() const { return free_; }

 private:
  uint64_t cap_ = 0, used_ = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> bytes
};

// A running render as the slice choice sees it (in arrival order).
struct SliceCandidate {
  uint64_t work_left;  // pixels x samples not yet handed to the engine (0: nothing left)
  uint32_t passes;     // slices given to newer work since its last slice, while it had work left
  uint64_t age = 0;    // pixels x samples handed to the engine (for any render) since it was submitted
};

// Which running render gets the next slice (after forced and new renders).
enum class SliceOrder {
  kOldestFirst,     // arrival order
  kLeastWorkFirst,  // least work left
  kResponseRatio,   // highest (age + work left) / work left, oldest on ties
};

// A running render passed over this many times by newer work (a newly
// admitted render, or a later render chosen by least-work-first) goes first,
// so that no stream of new or smaller requests can starve it. Under
// oldest-first it is passed over only by new admissions, which a render
// sees a few of in its lifetime at ordinary loads.
constexpr uint32_t kSlicePassLimit = 16;

enum class SliceChoice { kAdmitPending, kRunning, kNone };

// Chooses what the next slice of the engine's backlog is for: the oldest
// running render passed over kSlicePassLimit times (`*index`); else a
// pending render whose film fits (its first slice); else a running render
// with work left, by `order`. The response ratio (highest response ratio
// next, Brinch Hansen 1971, with work in place of time: age is the work the
// engine was given since the render's submission, which grows at the
// engine's rate while it waits) sets a render's wait against its own
// remaining work: a small render goes ahead of a large one soon after it
// arrives, and a large one goes ahead once it has waited in proportion to
// its size, so neither the median (as with oldest-first) nor the tail (as
// with least-work-first) is given up for the other.
inline SliceChoice PickSlice(const std::vector<SliceCandidate> &running, bool pending_fits, SliceOrder order,
                             size_t *index) {
  for (size_t k = 0; k < running.size(); ++k) {
    if (running[k].work_left && running[k].passes >= kSlicePassLimit) {
      *index = k;
      return SliceChoice::kRunning;
    }
  }
  if (pending_fits) return SliceChoice::kAdmitPending;
  bool found = false;
  for (size_t k = 0; k < running.size(); ++k) {
    const SliceCandidate &c = running[k];
    if (!c.work_left) continue;
    if (order == SliceOrder::kOldestFirst) {
      *index = k;
      return SliceChoice::kRunning;
    }
    bool better = !found;
    if (found) {
      const SliceCandidate &b = running[*index];
      if (order == SliceOrder::kLeastWorkFirst) {
        better = c.work_left < b.work_left;
      } else {  // c.age / c.work_left > b.age / b.work_left, exactly
        better = static_cast<unsigned __int128>(c.age) * b.work_left >
                 static_cast<unsigned __int128>(b.age) * c.work_left;
      }
    }
    if (better) {
      *index = k;
      found = true;
    }
  }
  return found ? SliceChoice::kRunning : SliceChoice::kNone;
}

// Updates the pass counts after a choice: every render older than the
// chosen one (all of them when a new render was admitted) that had work left
// was passed over; the chosen render's count restarts.
inline void RecordPick(std::vector<SliceCandidate> &running, SliceChoice choice, size_t index) {
  const size_t older = choice == SliceChoice::kAdmitPending ? running.size()
                       : choice == SliceChoice::kRunning    ? index
                                                            : 0;
  for (size_t k = 0; k < older; ++k)
    if (running[k].work_left) ++running[k].passes;
  if (choice == SliceChoice::kRunning) running[index].passes = 0;
}

}  // namespace muonx

#endif
