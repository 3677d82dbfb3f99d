// Host-side scheduler over one Engine: admission of render jobs by film
// memory and job slots, the iterate / snapshot loop, and completion
// (film resolve into the job's output). Single-threaded; the server wraps it
// in its own thread, muonx_render and muonx_render_batch drive it inline.
#ifndef MUONX_DEVICE_SCHEDULER_H_
#define MUONX_DEVICE_SCHEDULER_H_

#include <cstdint>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "sched_policy.h"

namespace muonx {

// Films of in-flight jobs: one device buffer carved by a first-fit range
// allocator (sched_policy.h).
class FilmPool {
 public:
  static constexpr uint64_t kAlign = RangeAllocator::kAlign;
  int Init(uint64_t bytes);
  uint64_t capacity() const { return ranges_.capacity(); }
  uint64_t in_use() const { return ranges_.in_use(); }
  bool empty() const { return ranges_.in_use() == 0; }
  // Returns false when no free range is large enough.
  bool Alloc(uint64_t bytes, uint64_t *offset) { return ranges_.Alloc(bytes, offset); }
  void Free(uint64_t offset, uint64_t bytes) { ranges_.Free(offset, bytes); }
  unsigned long long *at(uint64_t offset) const {
    return reinterpret_cast<unsigned long long *>(buf_.as<char>() + offset);
  }
  static uint64_t Round(uint64_t b) { return RangeAllocator::Round(b); }

 private:
  DevBuf buf_;
  RangeAllocator ranges_;
};

struct SchedJob {
  uint64_t ticket = 0;
  muonx_job spec{};
  uint32_t scene_slot = 0;
  uint32_t npix = 0;
  uint64_t film_off = 0, film_bytes = 0;
  uint32_t next_sample = 0;  // samples handed to the engine so far
  uint32_t open_slices = 0;  // of those, slices not finished yet
  uint32_t passes = 0;  // slices given to newer work since its last slice (PickSlice)
  uint64_t handed_at_submit = 0;  // the scheduler's handed_ when it was submitted (PickSlice's age)
  int state = 0;  // kPending, kActive, kResolving, kDone
  muonx_job_result result{};
  hipEvent_t ev = nullptr;  // resolve finished
};

// Consecutive samples of one job admitted as one engine job.
struct Slice {
  uint64_t ticket = 0;      // its job
  uint64_t nitems = 0;      // pixels x samples
  uint64_t admit_snap = 0;  // first done-snapshot taken after admission
  uint64_t seen_done = 0;   // finished samples at the last snapshot
};

class Scheduler {
 public:
  enum { kPending = 0, kActive = 1, kResolving = 2, kDone = 3 };

  static int Create(muonx_ctx *ctx, const muonx_tuning &tuning, const muonx_engine_config &cfg, Scheduler **out);
  ~Scheduler();

  // Checks that the job can ever run on this scheduler (params, film size,
  // shadow rays per vertex).
  int Validate(const muonx_job &job) const;
  // Queues a validated job; `submit_ns` is its arrival time.
  int Submit(const muonx_job &job, uint64_t ticket, uint64_t submit_ns);
  // Removes a job that is still waiting for admission; false if it is not
  // pending (admitted, finished or unknown). `submit_ns` gets its arrival.
  bool Cancel(uint64_t ticket, uint64_t *submit_ns);
  // One poll: admits work, launches `iterations_per_poll` engine
  // iterations, harvests the previous poll's done counters and finished
  // resolves. Jobs that completed are appended to `finished`.
  //
  // Admission: a job enters in arrival order once its film fits the film
  // budget; its samples go to the engine in slices of about a quarter of the
  // path capacity, and the engine is given new slices only while fewer than
  // 1.5 capacities of samples are unfinished. A newly admitted job's first
  // slice comes before further slices of running jobs (which continue
  // oldest first), so a small request waits behind about one queue fill,
  // not behind the whole of a large one, while large requests still finish
  // in arrival order. A running job passed over kSlicePassLimit times by
  // newer work goes first (PickSlice), so no stream of new or smaller
  // requests starves it. Images do not depend on the slicing.
  int Step(std::vector<SchedJob> *finished);
  bool Idle() const { return pending_.empty() && jobs_.empty(); }
  size_t pending() const { return pending_.size(); }
  size_t in_flight() const { return jobs_.size(); }
  Engine *engine() const { return eng_; }
  uint64_t iterations() const { return iterations_; }
  const unsigned long long *rays() const { return rays_; }
  uint64_t film_in_use() const { return films_.in_use(); }
  uint64_t film_capacity() const { return films_.capacity(); }
  const muonx_engine_config &config() const { return cfg_; }

 private:
  Scheduler() = default;
  int AdmitWork();
  int AdmitSlice(SchedJob &j);
  int Harvest(int b);
  int FinishResolves(std::vector<SchedJob> *finished, bool wait_all);
  void AdvanceWindow();

  muonx_ctx *ctx_ = nullptr;
  muonx_engine_config cfg_{};
  Engine *eng_ = nullptr;
  FilmPool films_;
  std::deque<SchedJob> pending_;
  std::map<uint64_t, SchedJob> jobs_;  // admitted, by ticket (arrival order)
  std::map<uint64_t, Slice> slices_;   // open slices by engine sequence
  uint64_t open_items_ = 0;            // unfinished samples of the open slices
  uint64_t handed_ = 0;                // samples (pixels x spp) handed to the engine, all renders
  std::vector<hipEvent_t> free_events_;
  int snap_ = 0;
  bool snap_pending_[2] = {false, false};
  uint64_t snap_id_[2] = {0, 0};
  uint64_t snap_count_ = 0;
  uint64_t iterations_ = 0;
  uint64_t progress_iter_ = 0;  // iteration count when a sample last finished
  unsigned long long rays_[kRayCounters] = {};
};

uint64_t NowNs();

}  // namespace muonx

#endif
