// CPU checks of the engine's index arithmetic and the scheduler's policies
// (muon_amd/csrc/device/engine_plan.h, sched_policy.h), the code the
// engine's kernels and host side index their buffers with. Built with the
// host compiler and run by tests/test_engine_plan.py: prints "ok" and exits
// 0, or names the failed check and exits 1.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "engine_plan.h"
#include "sched_policy.h"

using namespace muonx;

#define CHECK(c)                                                     \
  do {                                                               \
    if (!(c)) {                                                      \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);       \
      std::exit(1);                                                  \
    }                                                                \
  } while (0)

// Every array of the path and shadow buffers starts on a 256-byte boundary,
// holds its rows and ends before the next one starts; the buffer's size is
// at least the path budget's per-path bytes times the rows; the largest
// shadow-slot index the engine admits fits 32 bits.
static void CheckLayout() {
  CHECK(PathSlotBytes(0) == 248 && PathSlotBytes(4) == 248 + 4 * 34);
  for (size_t P : {size_t(1), size_t(7), size_t(1024), size_t(22800000), size_t(1) << 26}) {
    for (size_t K : {size_t(1), size_t(2), size_t(4), size_t(9)}) {
      for (int which = 0; which < 2; ++which) {
        const ArraySpec *spec = which ? kShadowArrays : kPathArrays;
        const int n = which ? kNumShadowArrays : kNumPathArrays;
        const size_t rows = which ? P * K : P;
        size_t off[64];
        const size_t bytes = LayoutOffsets(spec, n, rows, off);
        uint64_t per_row = 0;
        for (int k = 0; k < n; ++k) {
          CHECK(off[k] % kCarveAlign == 0);
          const size_t end = off[k] + rows * spec[k].per_path * spec[k].elem_bytes;
          CHECK(end <= (k + 1 < n ? off[k + 1] : bytes));
          per_row += uint64_t(spec[k].per_path) * spec[k].elem_bytes;
        }
        CHECK(bytes >= rows * per_row && bytes <= rows * per_row + size_t(n) * kCarveAlign);
      }
      CHECK(PathSlotBytes(int(K)) * P <= PathSlotBytes(int(K)) * P);  // no overflow at these sizes
    }
  }
  // Engine::Create admits capacity * kmax <= 2^31: the largest slot index
  const uint32_t caps[] = {1u << 28, 1u << 30, 715827882u};
  const uint32_t kms[] = {8u, 2u, 3u};
  for (int t = 0; t < 3; ++t) {
    const uint64_t P = caps[t], K = kms[t];
    CHECK(P * K <= kMaxShadowSlots);
    CHECK(uint64_t(ShadowSlot(uint32_t(K - 1), uint32_t(P), uint32_t(P - 1))) == (K - 1) * P + (P - 1));
  }
}

// The admission window, plan and fill of the engine, simulated with random
// job sizes, queue capacities and job-ring sizes: every generated sample
// maps to the admitted job that owns it (the ring slot still holds that
// job), to a pixel inside its film and a sample inside its slice; the queue
// never exceeds its capacity; every item of every job is generated exactly
// once, in order.
static void CheckWindow(std::mt19937_64 &rng) {
  struct Job {
    unsigned long long first, nitems;
    uint32_t npix, sample_base, count;
    unsigned long long generated;
  };
  for (uint32_t max_jobs : {1u, 2u, 3u, 8u, 64u}) {
    for (uint32_t capacity : {1u, 5u, 64u, 1000u}) {
      std::vector<unsigned long long> ring(max_jobs, ~0ull);
      std::vector<uint64_t> ring_seq(max_jobs, ~0ull);
      std::map<uint64_t, Job> jobs;  // by sequence number
      unsigned long long total = 0, cursor = 0, lo = 0, hi = 0;
      uint32_t queued = 0;
      for (int iter = 0; iter < 400; ++iter) {
        // admissions while job slots are free
        while (hi - lo < max_jobs && rng() % 3 != 0) {
          Job j;
          j.npix = 1 + uint32_t(rng() % 300);
          j.sample_base = uint32_t(rng() % 5);
          j.count = 1 + uint32_t(rng() % 4);
          j.first = total;
          j.nitems = uint64_t(j.npix) * j.count;
          j.generated = 0;
          ring[hi % max_jobs] = j.first;
          ring_seq[hi % max_jobs] = hi;
          jobs[hi] = j;
          total += j.nitems;
          ++hi;
        }
        // continuing paths: any number up to the current queue
        const uint32_t cont = queued ? uint32_t(rng() % (queued + 1)) : 0u;
        const uint32_t n = FillCount(total, cursor, capacity, cont);
        CHECK(uint64_t(cont) + n <= capacity);
        CHECK(n == std::min<unsigned long long>(total - cursor, capacity - cont));
        for (uint32_t k = 0; k < n; ++k) {
          const unsigned long long item = cursor + k;
          CHECK(lo < hi);
          const unsigned long long s = FindJobSeq(ring.data(), max_jobs, lo, hi, item);
          CHECK(s >= lo && s < hi);
          CHECK(ring_seq[s % max_jobs] == s);  // the slot was not reused
          Job &j = jobs.at(s);
          CHECK(j.first <= item && item < j.first + j.nitems);
          CHECK(item - j.first == j.generated);  // in order, each once
          ++j.generated;
          const ItemPos ip = ItemPosition(item - j.first, j.npix, j.sample_base);
          CHECK(ip.pixel < j.npix);
          CHECK(ip.sample >= j.sample_base && ip.sample < j.sample_base + j.count);
        }
        cursor += n;
        queued = cont + n;
        // the window's low end moves past jobs whose samples were all
        // generated (and, here at random, finished)
        while (lo < hi && jobs.at(lo).generated == jobs.at(lo).nitems && rng() % 2 == 0) ++lo;
      }
      for (auto &kv : jobs) {
        if (kv.first < lo) CHECK(kv.second.generated == kv.second.nitems);
      }
    }
  }
}

// Random allocations and releases against a byte-level model: ranges are
// aligned, inside the capacity and disjoint; accounting matches; after
// releasing everything one free range spans the pool.
static void CheckRanges(std::mt19937_64 &rng) {
  for (uint64_t cap : {uint64_t(256), uint64_t(4096), uint64_t(1) << 20}) {
    RangeAllocator ra;
    ra.Reset(cap);
    std::map<uint64_t, uint64_t> live;  // offset -> rounded bytes
    uint64_t used = 0;
    for (int step = 0; step < 3000; ++step) {
      if (live.empty() || rng() % 2) {
        const uint64_t want = 1 + rng() % (cap / 4 + 1);
        uint64_t off = 0;
        if (ra.Alloc(want, &off)) {
          const uint64_t b = RangeAllocator::Round(want);
          CHECK(off % RangeAllocator::kAlign == 0 && off + b <= ra.capacity());
          auto nx = live.lower_bound(off);
          if (nx != live.end()) CHECK(off + b <= nx->first);
          if (nx != live.begin()) {
            auto pv = std::prev(nx);
            CHECK(pv->first + pv->second <= off);
          }
          live[off] = b;
          used += b;
        } else {
          // first fit failed: no free range is large enough
          for (auto &fr : ra.free_ranges()) CHECK(fr.second < RangeAllocator::Round(want));
        }
      } else {
        auto it = live.begin();
        std::advance(it, long(rng() % live.size()));
        ra.Free(it->first, it->second);
        used -= it->second;
        live.erase(it);
      }
      CHECK(ra.in_use() == used);
      uint64_t fsum = 0;
      for (auto &fr : ra.free_ranges()) fsum += fr.second;
      CHECK(fsum + used == ra.capacity());
    }
    for (auto &kv : live) ra.Free(kv.first, kv.second);
    CHECK(ra.in_use() == 0 && ra.free_ranges().size() == 1 && ra.free_ranges().begin()->second == ra.capacity());
  }
}

// Slice choice: new admissions first, then the oldest, least-work or
// highest-response-ratio render; and under a stream that always has a new
// render waiting, a running large render still gets a slice at least every
// kSlicePassLimit + 1 picks (without the pass count it would get none).
static void CheckSlicePolicy(std::mt19937_64 &rng) {
  constexpr SliceOrder kOldest = SliceOrder::kOldestFirst, kLeast = SliceOrder::kLeastWorkFirst,
                       kRatio = SliceOrder::kResponseRatio;
  size_t idx = 99;
  std::vector<SliceCandidate> c = {{50, 0, 100}, {0, 0, 500}, {20, 0, 10}};
  CHECK(PickSlice(c, true, kOldest, &idx) == SliceChoice::kAdmitPending);
  CHECK(PickSlice(c, true, kRatio, &idx) == SliceChoice::kAdmitPending);
  CHECK(PickSlice(c, false, kOldest, &idx) == SliceChoice::kRunning && idx == 0);
  CHECK(PickSlice(c, false, kLeast, &idx) == SliceChoice::kRunning && idx == 2);
  // ratios 100/50 = 2 and 10/20 = 0.5 (render 1 has no work left)
  CHECK(PickSlice(c, false, kRatio, &idx) == SliceChoice::kRunning && idx == 0);
  c[2].age = 41;  // 41/20 > 2
  CHECK(PickSlice(c, false, kRatio, &idx) == SliceChoice::kRunning && idx == 2);
  c[2].age = 40;  // a tie: the older render
  CHECK(PickSlice(c, false, kRatio, &idx) == SliceChoice::kRunning && idx == 0);
  c[2].age = 10;
  CHECK(PickSlice({{0, 0}}, false, kLeast, &idx) == SliceChoice::kNone);
  CHECK(PickSlice({{0, 0, 7}}, false, kRatio, &idx) == SliceChoice::kNone);
  // the ratio pick against exact rational comparison on random candidates
  // (ages and work up to 2^40, as a large scene's samples reach)
  for (int rep = 0; rep < 20000; ++rep) {
    std::vector<SliceCandidate> r(1 + rng() % 9);
    for (auto &x : r) {
      x.work_left = rng() % 4 == 0 ? 0 : 1 + rng() % (1ull << (1 + rng() % 40));
      x.age = rng() % (1ull << (rng() % 41));
      x.passes = 0;
    }
    size_t k = 99;
    const SliceChoice ch = PickSlice(r, false, kRatio, &k);
    size_t best = r.size();
    for (size_t i = 0; i < r.size(); ++i) {
      if (!r[i].work_left) continue;
      if (best == r.size() || static_cast<unsigned __int128>(r[i].age) * r[best].work_left >
                                  static_cast<unsigned __int128>(r[best].age) * r[i].work_left)
        best = i;
    }
    CHECK(best == r.size() ? ch == SliceChoice::kNone : (ch == SliceChoice::kRunning && k == best));
  }
  // passes: a new admission passes every running render with work left
  RecordPick(c, SliceChoice::kAdmitPending, 0);
  CHECK(c[0].passes == 1 && c[1].passes == 0 && c[2].passes == 1);
  // least-work-first chose render 2: render 0 was passed over, 2 restarts
  RecordPick(c, SliceChoice::kRunning, 2);
  CHECK(c[0].passes == 2 && c[2].passes == 0);
  c[0].passes = kSlicePassLimit;
  CHECK(PickSlice(c, true, kLeast, &idx) == SliceChoice::kRunning && idx == 0);
  CHECK(PickSlice(c, true, kRatio, &idx) == SliceChoice::kRunning && idx == 0);
  for (SliceOrder order : {kOldest, kLeast, kRatio}) {
    // one large render with 1000 slices of work; at every pick a new
    // two-slice render is waiting and fits; every pick hands out one slice,
    // which ages every render submitted before it
    std::vector<SliceCandidate> run = {{1000, 0, 0}};
    uint64_t last = 0, max_gap = 0, picks_large = 0;
    for (uint64_t pick = 1; pick <= 5000; ++pick) {
      const SliceChoice ch = PickSlice(run, true, order, &idx);
      RecordPick(run, ch, idx);
      for (auto &x : run) ++x.age;
      if (ch == SliceChoice::kAdmitPending) {
        run.push_back({2, 0, 0});
      } else {
        CHECK(ch == SliceChoice::kRunning);
        run[idx].work_left -= 1;
        if (idx == 0) {
          ++picks_large;
          max_gap = std::max(max_gap, pick - last);
          last = pick;
        }
      }
    }
    CHECK(picks_large > 0 && max_gap <= kSlicePassLimit + 1);
  }
}

// A path's job word: job and scene slot below 2^16 each come back out.
static void CheckPathJobWord(std::mt19937_64 &rng) {
  for (int rep = 0; rep < 100000; ++rep) {
    const uint32_t job = uint32_t(rng() % kMaxPathSlotIndex), scene = uint32_t(rng() % kMaxPathSlotIndex);
    const uint32_t w = PackPathJob(job, scene);
    CHECK(PathJobSlot(w) == job && PathSceneSlot(w) == scene);
  }
  CHECK(PathJobSlot(PackPathJob(kMaxPathSlotIndex - 1, kMaxPathSlotIndex - 1)) == kMaxPathSlotIndex - 1);
  CHECK(PathSceneSlot(PackPathJob(0, kMaxPathSlotIndex - 1)) == kMaxPathSlotIndex - 1);
}

// The admission flush's slot runs: every job of [lo, hi) exactly once, in
// job order, inside the ring, in at most two runs.
static void CheckSlotRuns(std::mt19937_64 &rng) {
  for (int rep = 0; rep < 20000; ++rep) {
    const uint32_t M = 1 + uint32_t(rng() % 300);
    const unsigned long long lo = rng() % 100000ull + (rep & 1 ? (1ull << 40) : 0ull);
    const unsigned long long hi = lo + rng() % (M + 1ull);
    SlotRun runs[2] = {};
    const int n = SlotRuns(lo, hi, M, runs);
    CHECK(n >= 0 && n <= 2);
    CHECK((hi == lo) == (n == 0));
    unsigned long long q = lo;
    for (int r = 0; r < n; ++r) {
      CHECK(runs[r].count > 0 && runs[r].first < M && runs[r].first + runs[r].count <= M);
      for (uint32_t k = 0; k < runs[r].count; ++k, ++q) CHECK(runs[r].first + k == uint32_t(q % M));
    }
    CHECK(q == hi);
  }
}

int main() {
  std::mt19937_64 rng(12345);
  CheckLayout();
  CheckWindow(rng);
  CheckRanges(rng);
  CheckSlicePolicy(rng);
  CheckPathJobWord(rng);
  CheckSlotRuns(rng);
  std::printf("ok\n");
  return 0;
}
