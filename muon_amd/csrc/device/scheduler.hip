// Scheduler over the persistent engine (see scheduler.h).
#include "scheduler.h"
#include "engine_plan.h"

#include <algorithm>
#include <chrono>
#include <string>

namespace muonx {

uint64_t NowNs() {
  return uint64_t(
      std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
          .count());
}

// ---------------------------------------------------------------- film pool
int FilmPool::Init(uint64_t bytes) {
  bytes = Round(std::max<uint64_t>(bytes, kAlign));
  MX_HIP(buf_.Alloc(bytes));
  ranges_.Reset(bytes);
  return MUONX_OK;
}

// ---------------------------------------------------------------- scheduler
namespace {

uint64_t FilmBytes(const muonx_render_params &p) {
  return FilmPool::Round(uint64_t(p.width) * uint64_t(p.height) * 3 * sizeof(unsigned long long));
}

}  // namespace

int Scheduler::Create(muonx_ctx *ctx, const muonx_tuning &tuning, const muonx_engine_config &in, Scheduler **out) {
  *out = nullptr;
  muonx_engine_config cfg = in;
  if (cfg.slice_order != MUONX_SLICES_OLDEST_FIRST && cfg.slice_order != MUONX_SLICES_LEAST_WORK_FIRST &&
      cfg.slice_order != MUONX_SLICES_RESPONSE_RATIO) {
    SetError("engine config: unknown slice_order");
    return MUONX_ERR_INVALID;
  }
  if (cfg.max_jobs == 0) cfg.max_jobs = 1024;
  if (cfg.shadow_per_vertex <= 0) cfg.shadow_per_vertex = 4;
  if (cfg.iterations_per_poll == 0) cfg.iterations_per_poll = 2;
  if (cfg.film_budget_bytes == 0) cfg.film_budget_bytes = uint64_t(1) << 30;
  if (cfg.path_slots == 0) {
    const uint64_t budget = uint64_t(std::max(16, tuning.path_budget_mb)) << 20;
    cfg.path_slots = uint32_t(std::min<uint64_t>(budget / PathSlotBytes(cfg.shadow_per_vertex), uint64_t(1) << 26));
    cfg.path_slots = std::max<uint32_t>(cfg.path_slots, 1024);
  }
  Scheduler *s = new Scheduler();
  s->ctx_ = ctx;
  s->cfg_ = cfg;
  int rc = Engine::Create(ctx, tuning, cfg.path_slots, cfg.max_jobs, cfg.shadow_per_vertex, cfg.use_graph != 0,
                          &s->eng_);
  if (rc == MUONX_OK) rc = s->films_.Init(cfg.film_budget_bytes);
  if (rc != MUONX_OK) {
    delete s;
    return rc;
  }
  *out = s;
  return MUONX_OK;
}

Scheduler::~Scheduler() {
  if (eng_) (void)hipStreamSynchronize(eng_->stream());
  for (auto &kv : jobs_)
    if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
  for (auto ev : free_events_) (void)hipEventDestroy(ev);
  delete eng_;
}

int Scheduler::Validate(const muonx_job &job) const {
  if (!job.scene || !job.accum_rgb) {
    SetError("job: scene and accum_rgb must be set");
    return MUONX_ERR_INVALID;
  }
  if (job.scene->ctx != ctx_) {
    SetError("job: the scene belongs to another context");
    return MUONX_ERR_INVALID;
  }
  const muonx_render_params &p = job.params;
  if (p.width <= 0 || p.height <= 0 || p.spp <= 0) {
    SetError("job: width, height and spp must be positive");
    return MUONX_ERR_INVALID;
  }
  if (uint64_t(p.width) * uint64_t(p.height) > (uint64_t(1) << 28)) {
    SetError("job: more than 2^28 pixels");
    return MUONX_ERR_INVALID;
  }
  if (!p.has_camera) {
    SetError("job: scene has no camera");
    return MUONX_ERR_INVALID;
  }
  if (p.integrator != MUONX_INTEGRATOR_PATHTRACER) {
    SetError("job: the engine runs the path tracer only (integrator " + std::to_string(p.integrator) + ")");
    return MUONX_ERR_INVALID;
  }
  if (FilmBytes(p) > films_.capacity()) {
    SetError("job: film of " + std::to_string(FilmBytes(p)) + " bytes exceeds the film budget of " +
             std::to_string(films_.capacity()));
    return MUONX_ERR_CAPACITY;
  }
  const int K = Engine::ShadowRaysPerVertex(job.scene, p);
  if (K > eng_->kmax()) {
    SetError("job: needs " + std::to_string(K) + " shadow rays per vertex, the engine holds " +
             std::to_string(eng_->kmax()));
    return MUONX_ERR_CAPACITY;
  }
  return MUONX_OK;
}

bool Scheduler::Cancel(uint64_t ticket, uint64_t *submit_ns) {
  for (auto it = pending_.begin(); it != pending_.end(); ++it) {
    if (it->ticket == ticket) {
      if (submit_ns) *submit_ns = it->result.submit_ns;
      pending_.erase(it);
      return true;
    }
  }
  return false;
}

int Scheduler::Submit(const muonx_job &job, uint64_t ticket, uint64_t submit_ns) {
  SchedJob j;
  j.ticket = ticket;
  j.spec = job;
  j.film_bytes = FilmBytes(job.params);
  j.npix = uint32_t(job.params.width) * uint32_t(job.params.height);
  j.state = kPending;
  j.result.status = MUONX_PENDING;
  j.result.submit_ns = submit_ns;
  j.handed_at_submit = handed_;
  pending_.push_back(j);
  return MUONX_OK;
}

int Scheduler::AdmitSlice(SchedJob &j) {
  const uint64_t slice_items = std::max<uint64_t>(cfg_.path_slots / 4, 1);
  const uint32_t remaining = uint32_t(j.spec.params.spp) - j.next_sample;
  const uint32_t n = uint32_t(std::min<uint64_t>(remaining, std::max<uint64_t>(1, slice_items / j.npix)));
  JobSpec js;
  js.scene_slot = j.scene_slot;
  js.params = j.spec.params;
  js.key = j.spec.key;
  js.film = films_.at(j.film_off);
  js.sample_begin = j.next_sample;
  js.sample_count = n;
  js.clear_film = j.next_sample == 0;
  uint64_t seq = 0;
  const int rc = eng_->Admit(js, &seq);
  if (rc != MUONX_OK) return rc;
  Slice sl;
  sl.ticket = j.ticket;
  sl.nitems = uint64_t(j.npix) * n;
  sl.admit_snap = snap_count_;
  if (slices_.empty()) progress_iter_ = iterations_;
  slices_.emplace(seq, sl);
  open_items_ += sl.nitems;
  handed_ += sl.nitems;
  j.next_sample += n;
  ++j.open_slices;
  return MUONX_OK;
}

int Scheduler::AdmitWork() {
  const uint64_t backlog = 3 * uint64_t(cfg_.path_slots) / 2;
  std::vector<SliceCandidate> cands;
  std::vector<SchedJob *> running;
  while (open_items_ < backlog && eng_->seq_hi() - eng_->seq_lo() < eng_->max_jobs()) {
    cands.clear();
    running.clear();
    for (auto &kv : jobs_) {  // ticket (arrival) order
      SchedJob &j = kv.second;
      if (j.state != kActive) continue;
      const uint32_t spp = uint32_t(j.spec.params.spp);
      const uint64_t left = j.next_sample < spp ? uint64_t(spp - j.next_sample) * j.npix : 0;
      cands.push_back(SliceCandidate{left, j.passes, handed_ - j.handed_at_submit});
      running.push_back(&j);
    }
    // the oldest waiting job's film, if it fits (admission by film memory,
    // in arrival order)
    uint64_t off = 0;
    const bool fits = !pending_.empty() && films_.Alloc(pending_.front().film_bytes, &off);
    size_t idx = 0;
    const SliceOrder order = cfg_.slice_order == MUONX_SLICES_LEAST_WORK_FIRST ? SliceOrder::kLeastWorkFirst
                             : cfg_.slice_order == MUONX_SLICES_RESPONSE_RATIO ? SliceOrder::kResponseRatio
                                                                               : SliceOrder::kOldestFirst;
    const SliceChoice ch = PickSlice(cands, fits, order, &idx);
    if (ch != SliceChoice::kAdmitPending && fits) films_.Free(off, pending_.front().film_bytes);
    RecordPick(cands, ch, idx);
    for (size_t k = 0; k < running.size(); ++k) running[k]->passes = cands[k].passes;
    if (ch == SliceChoice::kAdmitPending) {
      SchedJob &p = pending_.front();
      uint32_t slot = 0;
      int rc = eng_->AddScene(p.spec.scene, &slot);
      if (rc != MUONX_OK) {
        films_.Free(off, p.film_bytes);
        return rc;
      }
      p.film_off = off;
      p.scene_slot = slot;
      p.state = kActive;
      p.result.admit_ns = NowNs();
      SchedJob &j = jobs_.emplace(p.ticket, p).first->second;
      pending_.pop_front();
      if ((rc = AdmitSlice(j)) != MUONX_OK) return rc;
      continue;
    }
    if (ch == SliceChoice::kNone) break;
    const int rc = AdmitSlice(*running[idx]);
    if (rc != MUONX_OK) return rc;
  }
  return MUONX_OK;
}

void Scheduler::AdvanceWindow() { eng_->Retire(slices_.empty() ? eng_->seq_hi() : slices_.begin()->first); }

int Scheduler::Harvest(int b) {
  const unsigned long long *done = eng_->done_host(b);
  if (slices_.empty()) progress_iter_ = iterations_;
  std::vector<uint64_t> complete;  // renders whose every sample is in the film
  for (auto it = slices_.begin(); it != slices_.end();) {
    Slice &sl = it->second;
    if (sl.admit_snap > snap_id_[b]) {  // snapshot older than the slice's admission
      ++it;
      continue;
    }
    const uint64_t d = done[it->first % eng_->max_jobs()];
    if (d > sl.seen_done) {
      open_items_ -= std::min(d, sl.nitems) - std::min(sl.seen_done, sl.nitems);
      sl.seen_done = d;
      progress_iter_ = iterations_;
    }
    if (d < sl.nitems) {
      ++it;
      continue;
    }
    const uint64_t ticket = sl.ticket;
    SchedJob &j = jobs_.at(ticket);
    it = slices_.erase(it);  // sl is gone from here on
    --j.open_slices;
    if (j.open_slices != 0 || j.next_sample < uint32_t(j.spec.params.spp)) continue;
    complete.push_back(ticket);
  }
  // their films, resolved together (a few launches for all of them)
  std::vector<FilmResolve> resolves;
  resolves.reserve(complete.size());
  for (uint64_t t : complete) {
    const SchedJob &j = jobs_.at(t);
    resolves.push_back(FilmResolve{films_.at(j.film_off), j.spec.accum_rgb, j.npix});
  }
  if (!resolves.empty()) {
    const int rc = eng_->ResolveFilms(resolves.data(), resolves.size());
    if (rc != MUONX_OK) return rc;
  }
  for (uint64_t t : complete) {
    SchedJob &j = jobs_.at(t);
    if (free_events_.empty()) {
      hipEvent_t ev;
      MX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      free_events_.push_back(ev);
    }
    j.ev = free_events_.back();
    free_events_.pop_back();
    MX_HIP(hipEventRecord(j.ev, eng_->stream()));
    // later work on the engine stream (a new job's film clear) runs after the resolve
    films_.Free(j.film_off, j.film_bytes);
    j.state = kResolving;
  }
  // Every generated path retires within kMaxBounces + 2 iterations, so a
  // long stretch without a finished sample means the engine lost paths.
  if (iterations_ - progress_iter_ > 2 * uint64_t(kMaxBounces) + 64) {
    SetError("engine: no sample finished in " + std::to_string(iterations_ - progress_iter_) + " iterations");
    return MUONX_ERR_STATE;
  }
  const unsigned long long *r = eng_->rays_host(b);
  for (int k = 0; k < kRayCounters; ++k) rays_[k] = r[k];
  AdvanceWindow();
  return MUONX_OK;
}

int Scheduler::FinishResolves(std::vector<SchedJob> *finished, bool wait_all) {
  for (auto it = jobs_.begin(); it != jobs_.end();) {
    SchedJob &j = it->second;
    if (j.state == kResolving) {
      const hipError_t q = wait_all ? hipEventSynchronize(j.ev) : hipEventQuery(j.ev);
      if (q == hipSuccess) {
        j.result.done_ns = NowNs();
        j.result.status = MUONX_OK;
        j.state = kDone;
        free_events_.push_back(j.ev);
        j.ev = nullptr;
        finished->push_back(j);
        it = jobs_.erase(it);
        continue;
      }
      if (q == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: keep it out of later launch checks
      } else {
        SetError(std::string("engine: ") + hipGetErrorString(q));
        return MUONX_ERR_HIP;
      }
    }
    ++it;
  }
  return MUONX_OK;
}

int Scheduler::Step(std::vector<SchedJob> *finished) {
  const bool idle = slices_.empty();  // every admitted sample has finished: no path in flight
  int rc = AdmitWork();
  if (rc != MUONX_OK) return rc;
  if (idle && !slices_.empty() && (rc = eng_->Prime()) != MUONX_OK) return rc;
  if (!slices_.empty()) {
    for (uint32_t i = 0; i < cfg_.iterations_per_poll; ++i) {
      if ((rc = eng_->Iterate()) != MUONX_OK) return rc;
      ++iterations_;
    }
    const int b = snap_;
    if ((rc = eng_->SnapshotDone(b)) != MUONX_OK) return rc;
    snap_id_[b] = snap_count_++;
    snap_pending_[b] = true;
    snap_ ^= 1;
    // the previous poll's counters, read while this poll's iterations run
    const int pb = b ^ 1;
    if (snap_pending_[pb]) {
      if ((rc = eng_->WaitDone(pb)) != MUONX_OK) return rc;
      snap_pending_[pb] = false;
      if ((rc = Harvest(pb)) != MUONX_OK) return rc;
    }
  } else {
    for (int k = 0; k < 2; ++k) {
      const int b = snap_ ^ 1 ^ k;  // older snapshot first
      if (!snap_pending_[b]) continue;
      if ((rc = eng_->WaitDone(b)) != MUONX_OK) return rc;
      snap_pending_[b] = false;
      if ((rc = Harvest(b)) != MUONX_OK) return rc;
    }
  }
  return FinishResolves(finished, slices_.empty() && pending_.empty());
}

}  // namespace muonx
