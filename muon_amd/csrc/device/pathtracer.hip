// Path tracer entry points over the persistent engine (engine.h):
//   muonx_render (integrator pathtracer)  one job, driven inline
//   muonx_render_batch                    n jobs, driven inline
//   muonx_server_*                        request loop on its own thread
// The context keeps one scheduler (engine + film pool) for the first two and
// rebuilds it when a job needs more than it holds or the tuning changed.
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "render.h"
#include "scheduler.h"

namespace muonx {

void DestroyContextScheduler(RenderWorkspace *ws) {
  delete ws->sched;
  ws->sched = nullptr;
}

namespace {

bool SameTuning(const muonx_tuning &a, const muonx_tuning &b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }
bool SameConfig(const muonx_engine_config &a, const muonx_engine_config &b) {
  return a.path_slots == b.path_slots && a.max_jobs == b.max_jobs && a.shadow_per_vertex == b.shadow_per_vertex &&
         a.iterations_per_poll == b.iterations_per_poll && a.film_budget_bytes == b.film_budget_bytes &&
         a.use_graph == b.use_graph;
}

// The context's scheduler for `jobs`: with `want`, one built from exactly
// that configuration; otherwise one built for the plain path (capacity from
// the tuning, not a caller's batch configuration) that can run every job.
// Called with the context lock held.
int ContextScheduler(muonx_ctx *ctx, const muonx_engine_config *want, const muonx_job *jobs, uint32_t n,
                     Scheduler **out) {
  RenderWorkspace *ws = GetWorkspace(ctx);
  Scheduler *s = ws->sched;
  bool ok = s && SameTuning(ws->sched_tuning, ctx->tuning);
  if (ok) ok = want ? SameConfig(ws->sched_want, *want) : ws->sched_plain;
  for (uint32_t i = 0; ok && i < n; ++i) ok = s->Validate(jobs[i]) == MUONX_OK;
  if (ok) {
    *out = s;
    return MUONX_OK;
  }
  DestroyContextScheduler(ws);
  muonx_engine_config cfg{};
  if (want) {
    cfg = *want;
  } else {
    // one job at a time: room for its film and shadow rays, few job slots
    uint64_t film = 0;
    int kmax = 4;
    for (uint32_t i = 0; i < n; ++i) {
      const muonx_render_params &p = jobs[i].params;
      if (jobs[i].scene && p.width > 0 && p.height > 0) {
        film = std::max<uint64_t>(film, uint64_t(p.width) * uint64_t(p.height) * 24);
        kmax = std::max(kmax, Engine::ShadowRaysPerVertex(jobs[i].scene, p));
      }
    }
    cfg.max_jobs = 64;
    cfg.shadow_per_vertex = kmax;
    cfg.iterations_per_poll = 2;
    cfg.film_budget_bytes = std::max<uint64_t>(film, uint64_t(64) << 20);
    cfg.use_graph = 1;
  }
  int rc = Scheduler::Create(ctx, ctx->tuning, cfg, &ws->sched);
  if (rc != MUONX_OK) return rc;
  ws->sched_tuning = ctx->tuning;
  ws->sched_want = want ? *want : muonx_engine_config{};
  ws->sched_plain = want == nullptr;
  for (uint32_t i = 0; i < n; ++i)
    if ((rc = ws->sched->Validate(jobs[i])) != MUONX_OK) return rc;
  *out = ws->sched;
  return MUONX_OK;
}

// Runs `jobs` to completion on `s`; results[i] is job i's.
int RunInline(Scheduler *s, const muonx_job *jobs, uint32_t n, muonx_job_result *results,
              unsigned long long rays[kRayCounters]) {
  int rc = s->engine()->ResetRayCounts();
  if (rc != MUONX_OK) return rc;
  const uint64_t t0 = NowNs();
  for (uint32_t i = 0; i < n; ++i)
    if ((rc = s->Validate(jobs[i])) != MUONX_OK) return rc;
  // Jobs without NEE first (in their order), then the others: while only the
  // former are admitted the engine runs its single-shade-kernel iterations.
  // A job's image does not depend on the order.
  for (int with_nee = 0; with_nee < 2; ++with_nee)
    for (uint32_t i = 0; i < n; ++i)
      if ((jobs[i].params.nee != MUONX_NEE_OFF) == (with_nee != 0) && (rc = s->Submit(jobs[i], i, t0)) != MUONX_OK)
        return rc;
  std::vector<SchedJob> finished;
  finished.reserve(n);
  while (!s->Idle())
    if ((rc = s->Step(&finished)) != MUONX_OK) return rc;
  if (results)
    for (const SchedJob &j : finished) results[j.ticket] = j.result;
  if (rays) std::memcpy(rays, s->rays(), kRayCounters * sizeof(unsigned long long));
  return MUONX_OK;
}

}  // namespace

int RenderPathTracer(muonx_scene *scene, const muonx_render_params &p, float *accum, hipStream_t stream,
                     unsigned long long rays[kRayCounters]) {
  muonx_ctx *ctx = scene->ctx;
  muonx_job job{};
  job.scene = scene;
  job.params = p;
  job.key = 0;
  job.accum_rgb = accum;
  Scheduler *s = nullptr;
  int rc = ContextScheduler(ctx, nullptr, &job, 1, &s);
  if (rc != MUONX_OK) return rc;
  RenderWorkspace *ws = GetWorkspace(ctx);
  if (!ws->ev_in) MX_HIP(hipEventCreateWithFlags(&ws->ev_in, hipEventDisableTiming));
  if (!ws->ev_out) MX_HIP(hipEventCreateWithFlags(&ws->ev_out, hipEventDisableTiming));
  hipStream_t es = s->engine()->stream();
  // the engine's stream runs after the caller's earlier work ...
  MX_HIP(hipEventRecord(ws->ev_in, stream));
  MX_HIP(hipStreamWaitEvent(es, ws->ev_in, 0));
  rc = RunInline(s, &job, 1, nullptr, rays);
  if (rc != MUONX_OK) return rc;
  // ... and the caller's later work after the film is written
  MX_HIP(hipEventRecord(ws->ev_out, es));
  MX_HIP(hipStreamWaitEvent(stream, ws->ev_out, 0));
  return MUONX_OK;
}

}  // namespace muonx

using namespace muonx;

extern "C" int muonx_render_batch(muonx_ctx *ctx, const muonx_job *jobs, uint32_t n, const muonx_engine_config *cfg,
                                  muonx_job_result *results, muonx_trace_stats *stats) {
  if (!ctx || (n && !jobs)) {
    SetError("muonx_render_batch: null argument");
    return MUONX_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  MX_HIP(hipSetDevice(ctx->device));
  Scheduler *s = nullptr;
  int rc = ContextScheduler(ctx, cfg, jobs, n, &s);
  if (rc != MUONX_OK) return rc;
  const uint64_t t0 = NowNs();
  unsigned long long rays[kRayCounters] = {};
  rc = RunInline(s, jobs, n, results, rays);
  if (rc != MUONX_OK) return rc;
  if (stats) {
    stats->primary_rays = rays[0];
    stats->secondary_rays = rays[1];
    stats->shadow_rays = rays[2];
    stats->culled_rays = rays[3];
    stats->render_ns = NowNs() - t0;
    stats->object_tests = rays[kCtrObjTests];
    stats->object_hits = rays[kCtrObjHits];
    stats->bounds_tests = rays[kCtrBoxTests];
    stats->bounds_hits = rays[kCtrBoxHits];
  }
  return MUONX_OK;
}

// ---------------------------------------------------------------- server
struct muonx_server {
  muonx_ctx *ctx = nullptr;
  Scheduler *sched = nullptr;
  std::thread thread;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  struct Incoming {
    uint64_t ticket, submit_ns;
    muonx_job job;
  };
  std::deque<Incoming> incoming;
  std::unordered_map<uint64_t, muonx_job_result> results;  // finished, not yet collected
  std::unordered_map<uint64_t, muonx_scene *> outstanding;  // submitted, not finished
  std::vector<uint64_t> cancels;  // cancel requests for jobs handed to the scheduler
  uint64_t next_ticket = 1;
  bool stop = false;
  int fatal = MUONX_OK;
  std::string fatal_msg;
  muonx_server_stats stats{};
};

namespace {

// Records a job dropped before admission (caller holds srv->mu).
void PostCancelled(muonx_server *srv, uint64_t ticket, uint64_t submit_ns) {
  muonx_job_result r{};
  r.status = MUONX_ERR_CANCELLED;
  r.submit_ns = submit_ns;
  r.done_ns = NowNs();
  srv->results[ticket] = r;
  auto it = srv->outstanding.find(ticket);
  if (it != srv->outstanding.end()) {
    it->second->inflight.fetch_sub(1);
    srv->outstanding.erase(it);
  }
  ++srv->stats.cancelled;
}

void ServerLoop(muonx_server *srv) {
  if (hipSetDevice(srv->ctx->device) != hipSuccess) {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->fatal = MUONX_ERR_HIP;
    srv->fatal_msg = "server: hipSetDevice failed";
    srv->cv_done.notify_all();
    return;
  }
  Scheduler *s = srv->sched;
  std::vector<SchedJob> finished;
  std::vector<muonx_server::Incoming> take;
  std::vector<uint64_t> cancels;
  std::vector<std::pair<uint64_t, uint64_t>> dropped;  // (ticket, submit_ns)
  while (true) {
    {
      std::unique_lock<std::mutex> lock(srv->mu);
      srv->cv_work.wait(lock, [&] { return srv->stop || !srv->incoming.empty() || !s->Idle(); });
      if (srv->stop && srv->incoming.empty() && s->Idle()) break;
      take.assign(srv->incoming.begin(), srv->incoming.end());
      srv->incoming.clear();
      cancels.swap(srv->cancels);
      srv->cancels.clear();
    }
    int rc = MUONX_OK;
    for (const auto &in : take)
      if (rc == MUONX_OK) rc = s->Submit(in.job, in.ticket, in.submit_ns);
    // cancel requests: jobs still waiting for admission are dropped; the
    // others (admitted or finished) are left to complete
    dropped.clear();
    for (const uint64_t t : cancels) {
      uint64_t submit_ns = 0;
      if (s->Cancel(t, &submit_ns)) dropped.emplace_back(t, submit_ns);
    }
    cancels.clear();
    finished.clear();
    if (rc == MUONX_OK) rc = s->Step(&finished);
    std::lock_guard<std::mutex> lock(srv->mu);
    for (const auto &d : dropped) PostCancelled(srv, d.first, d.second);
    for (const SchedJob &j : finished) {
      srv->results[j.ticket] = j.result;
      srv->outstanding.erase(j.ticket);
      j.spec.scene->inflight.fetch_sub(1);
      ++srv->stats.completed;
    }
    srv->stats.pending = s->pending() + srv->incoming.size();
    srv->stats.in_flight = s->in_flight();
    srv->stats.iterations = s->iterations();
    srv->stats.primary_rays = s->rays()[0];
    srv->stats.secondary_rays = s->rays()[1];
    srv->stats.shadow_rays = s->rays()[2];
    srv->stats.film_bytes_in_use = s->film_in_use();
    if (rc != MUONX_OK) {
      // the engine is in an unknown state: fail every job that is not finished
      srv->fatal = rc;
      srv->fatal_msg = muonx_last_error();
      for (const auto &kv : srv->outstanding) {
        muonx_job_result r{};
        r.status = rc;
        srv->results[kv.first] = r;
        kv.second->inflight.fetch_sub(1);
      }
      srv->outstanding.clear();
      srv->incoming.clear();
      srv->cv_done.notify_all();
      return;
    }
    if (!finished.empty() || !dropped.empty()) srv->cv_done.notify_all();
  }
}

}  // namespace

extern "C" int muonx_server_create(muonx_ctx *ctx, const muonx_engine_config *cfg, muonx_server **out) {
  if (!ctx || !out) {
    SetError("muonx_server_create: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = nullptr;
  muonx_tuning tuning;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    tuning = ctx->tuning;
  }
  MX_HIP(hipSetDevice(ctx->device));
  muonx_engine_config c = cfg ? *cfg : muonx_engine_config{};
  if (!cfg) c.use_graph = 1;
  muonx_server *srv = new muonx_server();
  srv->ctx = ctx;
  int rc = Scheduler::Create(ctx, tuning, c, &srv->sched);
  if (rc != MUONX_OK) {
    delete srv;
    return rc;
  }
  srv->stats.film_budget_bytes = srv->sched->film_capacity();
  srv->stats.path_slots = srv->sched->config().path_slots;
  srv->stats.shadow_per_vertex = srv->sched->config().shadow_per_vertex;
  srv->thread = std::thread(ServerLoop, srv);
  *out = srv;
  return MUONX_OK;
}

extern "C" int muonx_server_submit(muonx_server *srv, const muonx_job *job, uint64_t *ticket) {
  if (!srv || !job || !ticket) {
    SetError("muonx_server_submit: null argument");
    return MUONX_ERR_INVALID;
  }
  const uint64_t now = NowNs();
  // Validate reads only immutable scheduler state (capacities, kmax)
  const int rc = srv->sched->Validate(*job);
  if (rc != MUONX_OK) return rc;
  std::lock_guard<std::mutex> lock(srv->mu);
  if (srv->fatal != MUONX_OK) {
    SetError("server stopped after an error: " + srv->fatal_msg);
    return MUONX_ERR_STATE;
  }
  if (srv->stop) {
    SetError("server is shutting down");
    return MUONX_ERR_STATE;
  }
  *ticket = srv->next_ticket++;
  job->scene->inflight.fetch_add(1);  // muonx_scene_destroy waits for it
  srv->incoming.push_back({*ticket, now, *job});
  srv->outstanding[*ticket] = job->scene;
  ++srv->stats.submitted;
  srv->cv_work.notify_one();
  return MUONX_OK;
}

extern "C" int muonx_server_wait(muonx_server *srv, uint64_t ticket, int64_t timeout_ms, muonx_job_result *result) {
  if (!srv) {
    SetError("muonx_server_wait: null server");
    return MUONX_ERR_INVALID;
  }
  std::unique_lock<std::mutex> lock(srv->mu);
  auto ready = [&] { return srv->results.count(ticket) != 0; };
  if (!ready() && !srv->outstanding.count(ticket)) {
    SetError("muonx_server_wait: unknown or already collected ticket " + std::to_string(ticket));
    return MUONX_ERR_INVALID;
  }
  if (timeout_ms < 0) {
    srv->cv_done.wait(lock, ready);
  } else if (!srv->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) {
    return MUONX_PENDING;
  }
  const muonx_job_result r = srv->results[ticket];
  srv->results.erase(ticket);
  if (result) *result = r;
  if (r.status == MUONX_ERR_CANCELLED) {
    SetError("muonx_server_wait: ticket " + std::to_string(ticket) + " was cancelled");
    return r.status;
  }
  if (r.status != MUONX_OK) {
    SetError("job failed: " + srv->fatal_msg);
    return r.status;
  }
  return MUONX_OK;
}

extern "C" int muonx_server_cancel(muonx_server *srv, uint64_t ticket) {
  if (!srv) {
    SetError("muonx_server_cancel: null server");
    return MUONX_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(srv->mu);
  if (srv->results.count(ticket)) return MUONX_OK;  // finished: nothing to cancel
  if (!srv->outstanding.count(ticket)) {
    SetError("muonx_server_cancel: unknown or already collected ticket " + std::to_string(ticket));
    return MUONX_ERR_INVALID;
  }
  for (auto it = srv->incoming.begin(); it != srv->incoming.end(); ++it) {
    if (it->ticket == ticket) {  // not yet handed to the scheduler
      const uint64_t submit_ns = it->submit_ns;
      srv->incoming.erase(it);
      PostCancelled(srv, ticket, submit_ns);
      srv->cv_done.notify_all();
      return MUONX_OK;
    }
  }
  srv->cancels.push_back(ticket);  // the server thread drops it if it is still pending
  srv->cv_work.notify_one();
  return MUONX_OK;
}

extern "C" int muonx_server_get_stats(muonx_server *srv, muonx_server_stats *out) {
  if (!srv || !out) {
    SetError("muonx_server_get_stats: null argument");
    return MUONX_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(srv->mu);
  *out = srv->stats;
  return MUONX_OK;
}

extern "C" void muonx_server_destroy(muonx_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lock(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  if (srv->thread.joinable()) srv->thread.join();
  delete srv->sched;
  delete srv;
}
