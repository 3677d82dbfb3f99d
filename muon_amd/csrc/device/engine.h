// Persistent wavefront engine: the path tracer (PathTracer,
// muon/integration.cc:205-630) over a fixed-capacity queue of paths shared
// by any number of render jobs of any scenes.
//
// Each iteration advances every queued path by one bounce:
//   extend   closest hit of the queue's rays
//   shade    emission and NEE shadow rays; then the MIS BRDF ray and the
//            continuation: a continuing path is appended to the next queue
//            together with its state
//   shadow   any hit of the compacted shadow rays
//   mis      closest hit of the MIS rays
//   retire   adds the NEE terms and, for MIS rays that hit a light, the
//            weighted emission; continuing paths carry their radiance
//            into the next queue, finished ones add it to their job's film
//   plan     decides how many new samples fit behind the continuing paths
//   fill     appends camera samples of admitted jobs, in work order
// so new jobs join between any two bounces (continuous batching). Path state
// moves with the queue (compacted every bounce), so every per-path access is
// coalesced. Films are accumulated in 64-bit fixed point (2^-24 units):
// integer addition is associative, so a job's film does not depend on what
// else is in flight. Random numbers are keyed by (job seed, job key, pixel,
// sample).
#ifndef MUONX_DEVICE_ENGINE_H_
#define MUONX_DEVICE_ENGINE_H_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../../include/muonx.h"
#include "render.h"
#include "runtime.h"
#include "traverse_launch.h"

namespace muonx {

constexpr int kEngBlock = 128;
constexpr int kShadeBlock = 512;  // shade kernels: one queue-append atomic per block and append
constexpr int kShadeWaves = kShadeBlock / 64;
constexpr int kMaxBounces = 1024;  // hard cap for max_depth = -1 without Russian roulette
constexpr double kFilmScale = 16777216.0;  // fixed-point units per radiance unit

// Counters of the engine: primary, secondary (continuation and MIS rays, as
// the reference counts them), shadow rays, the MIS rays among the secondary
// ones resolved without a traversal, and the traversals' work in the device
// tree: primitive tests and hits, child boxes tested and passed.
constexpr int kRayCounters = kEngineCounters;
constexpr int kCtrObjTests = 4, kCtrObjHits = 5, kCtrBoxTests = 6, kCtrBoxHits = 7;

struct JobDev {
  CamParams cam;
  uint32_t scene;  // scene table slot
  uint32_t npix;
  uint32_t seed, key;
  int min_depth, cutoff, nee, is_mode, rr, light_samples, strata, per_section, K;
  unsigned long long first_item;  // global index of the job's first sample
  unsigned long long nitems;      // W * H * (samples of this slice)
  uint32_t sample_base;           // first sample index of the slice
  unsigned long long *film;       // W * H * 3 fixed-point sums
  unsigned long long *done;       // finished samples
  uint32_t clear_film;            // zero the film at admission (the first slice of a render)
};
constexpr uint32_t kClearBlocksPerJob = 64;  // blocks per admitted job in k_admit_clear

// One finished render's film and its float output, for k_film_resolve.
struct FilmResolve {
  const unsigned long long *film;
  float *out;
  uint32_t npix;
};
constexpr int kResolveBatch = 16;
struct FilmResolveBatch {
  FilmResolve r[kResolveBatch];
};

// Device-side scalars of the engine (one struct so that the host can update
// the admission window with a single copy).
// A counter on a 256-byte line of its own: atomics on one L2 line are
// executed one after another, so counters that many wavefronts update are
// kept apart.
struct alignas(256) HotCounter {
  uint32_t v;
  uint32_t pad[63];
};

struct alignas(256) HotCounter64 {
  unsigned long long v;
  unsigned long long pad[31];
};
constexpr int kRaySlots = 16;  // partial ray counts per counter, blocks spread over them

struct EngCounters {
  HotCounter qcount[2];  // path queues
  HotCounter mis_count;
  HotCounter shadow_count;
  HotCounter claim[3];   // work claimed by the per-lane-fetch extend / shadow / mis kernels
  uint32_t plan_n, plan_base;
  unsigned long long plan_cursor;
  unsigned long long cursor;          // next work item to generate
  unsigned long long total_items;     // admitted work (host-written)
  unsigned long long seq_lo, seq_hi;  // admitted jobs (host-written)
  unsigned long long rays[kRayCounters];  // primary, secondary, shadow, culled MIS, traversal work
  // This iteration's counts by kernel block (block % kRaySlots), folded into
  // `rays` by the plan kernel: one counter updated by every wavefront of a
  // launch would serialise its atomics.
  HotCounter64 ray_slots[kRayCounters][kRaySlots];
};

// Per-path state, struct of arrays over queue positions.
struct PathState {
  uint32_t *job, *pixel, *ctr;
  int32_t *depth;
  float *rad, *thr;       // 3 per path
  uint32_t *key_a, *key_b;  // the sample's random-stream key (path_key), set at fill
};

struct EngArgs {
  const DevSceneView *scenes;
  const JobDev *jobs;
  const unsigned long long *job_first;  // first_item by job sequence % max_jobs
  uint32_t max_jobs;
  uint32_t capacity;
  int kmax;
  EngCounters *c;
  PathState st[2];    // state of queue 0 / 1
  float4 *q_ray[2];   // 2 per path: (o.xyz, 0), (d.xyz, 0)
  int32_t *hit_prim;  // closest hit of path i of the current queue
  uint32_t *nidx;     // its position in the next queue, or ~0 when it ends
  float4 *m_ray;      // MIS rays: (o.xyz, path index), (d.xyz, 0)
  float *m_fac;       // MIS weight times the BRDF-sampled throughput
  int32_t *m_hit;     // closest hit of the MIS ray, -1 for none
  float4 *shift;      // origin of path i's shadow and MIS rays (xyz), material id (w)
  float4 *hnorm;      // shading normal at path i's hit
  // shadow rays, index k * capacity + i, and the queue of those to trace
  float4 *s_dir;      // (direction, tmax)
  float *s_contrib;
  uint8_t *s_vis, *s_valid;
  uint32_t *s_queue;
};

// One engine job: samples [sample_begin, sample_begin + sample_count) of
// every pixel of a render (a slice; a render may be admitted as several).
struct JobSpec {
  uint32_t scene_slot;
  muonx_render_params params;
  uint32_t key;
  unsigned long long *film;  // npix * 3 fixed-point sums, zeroed before the next launch when clear_film
  uint32_t sample_begin = 0, sample_count = 0;
  bool clear_film = true;
};

class Engine {
 public:
  // capacity: paths in flight; max_jobs: job slots; kmax: shadow rays per vertex.
  // The traversal variant follows `tuning` (a copy is kept).
  static int Create(muonx_ctx *ctx, const muonx_tuning &tuning, uint32_t capacity, uint32_t max_jobs, int kmax,
                    bool use_graph, Engine **out);
  ~Engine();

  // Registers a resident scene (by id); returns its slot. Entries of
  // destroyed scenes are reused.
  int AddScene(muonx_scene *scene, uint32_t *slot);
  // Shadow rays one vertex of a job generates; admission requires <= kmax.
  static int ShadowRaysPerVertex(const muonx_scene *scene, const muonx_render_params &p);
  // Admits a job (its done counter, and with clear_film its film, are
  // zeroed on the engine stream before the next launch).
  int Admit(const JobSpec &job, uint64_t *seq);
  // Advances the admission window past finished jobs (all of their samples
  // generated and retired).
  void Retire(uint64_t seq_lo) { seq_lo_ = seq_lo; }
  // Launches one iteration (all kernels) on the engine stream.
  int Iterate();
  // With no path in flight (every admitted sample finished): generates the
  // first paths of newly admitted work without running the traversal and
  // shading kernels of an iteration over an empty queue.
  int Prime();
  // Enqueues a copy of the per-job-slot done counters (index = seq %
  // max_jobs) into host buffer `b` (0 or 1) and marks it with an event.
  int SnapshotDone(int b);
  // Waits for snapshot `b`; done_host(b) is then valid.
  int WaitDone(int b);
  const unsigned long long *done_host(int b) const { return h_done_[b]; }
  // Ray counters (primary, secondary, shadow) taken with snapshot `b`.
  const unsigned long long *rays_host(int b) const { return h_done_[b] + max_jobs_; }
  // Converts finished renders' fixed-point films to float sums on the engine
  // stream, up to kResolveBatch renders per launch.
  int ResolveFilms(const FilmResolve *films, size_t n);
  // Ray counters (primary, secondary, shadow) since the last reset; waits.
  int RayCounts(unsigned long long out[kRayCounters]);
  int ResetRayCounts();

  hipStream_t stream() const { return stream_; }
  uint32_t capacity() const { return capacity_; }
  uint32_t max_jobs() const { return max_jobs_; }
  int kmax() const { return kmax_; }
  uint64_t seq_lo() const { return seq_lo_; }
  uint64_t seq_hi() const { return seq_hi_; }
  uint64_t total_items() const { return total_items_; }
  muonx_ctx *ctx() const { return ctx_; }
  uint64_t device_bytes() const { return device_bytes_; }

 private:
  Engine() = default;
  // True when no job of the admission window has NEE on.
  bool PlainWindow() const;
  int LaunchIteration(int cur, bool plain);
  // Copies the records of jobs admitted since the last launch to the device
  // (one copy per contiguous run of slots) and clears their films and done
  // counters with one kernel; called before every launch.
  int FlushAdmitted();
  int SyncWindow();
  // Destroys the cached iteration graphs (they capture the plan and tables).
  void DropGraphs();

  muonx_ctx *ctx_ = nullptr;
  muonx_tuning tuning_{};
  int num_cus_ = 0;
  hipStream_t stream_ = nullptr;
  uint32_t capacity_ = 0, max_jobs_ = 0;
  int kmax_ = 1;
  bool use_graph_ = false;
  // iteration graphs by kernel set (0: NEE / MIS kernels, 1: plain) and queue parity
  hipGraphExec_t graph_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  uint32_t max_depth_ = 1, max_wstack_ = 1;
  uint64_t device_bytes_ = 0;
  int cur_ = 0;
  uint64_t seq_lo_ = 0, seq_hi_ = 0, total_items_ = 0;
  uint64_t flushed_hi_ = 0;  // jobs below it are on the device
  uint64_t window_written_[3] = {~0ull, ~0ull, ~0ull};
  TraverseCfg cfg_{};
  EngArgs a_{};
  DevBuf scenes_, jobs_, job_first_, counters_, paths_, shadow_, spill_, done_;
  unsigned long long *h_done_[2] = {nullptr, nullptr};
  hipEvent_t done_ev_[2] = {nullptr, nullptr};
  std::vector<DevSceneView> host_scenes_;
  std::vector<muonx_scene *> scene_ptrs_;
  std::vector<uint64_t> scene_ids_;
  // pinned host staging
  JobDev *h_jobs_ = nullptr;
  unsigned long long *h_first_ = nullptr;
  unsigned long long *h_window_ = nullptr;  // [seq_lo, seq_hi, total_items] x 4 ring
  int window_slot_ = 0;
  hipEvent_t window_ev_[4] = {nullptr, nullptr, nullptr, nullptr};
};

}  // namespace muonx

#endif
