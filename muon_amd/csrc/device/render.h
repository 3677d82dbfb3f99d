// Shared pieces of the GPU render path: camera, sample positions, the
// per-context workspace of queues and counters.
#ifndef MUONX_DEVICE_RENDER_H_
#define MUONX_DEVICE_RENDER_H_

#include "../common/vecmath.h"
#include "rng.h"
#include "runtime.h"

namespace muonx {

constexpr int kSimpleBlock = 128;

// Device counters (u64): rays by kind, plus scratch.
constexpr int kCtrPrimary = 0;
constexpr int kCtrSecondary = 1;
constexpr int kCtrShadow = 2;
constexpr int kCtrObjectTests = 3;  // the reference's TraceStats test counters
constexpr int kCtrObjectHits = 4;
constexpr int kCtrBoundsTests = 5;
constexpr int kCtrBoundsHits = 6;
constexpr int kNumCounters = 8;

struct CamParams {
  V3 eye, u, v, w;
  float tan_fov_x, tan_fov_y;
  int width, height;
};

// muon/camera.cc:37-62
__device__ __forceinline__ void cast_ray(const CamParams &c, float x, float y, V3 &o, V3 &d) {
  const float hw = c.width / 2.0f;
  const float hh = c.height / 2.0f;
  const float ox = (x - hw) / hw;
  const float oy = (hh - y) / hh;
  const float alpha = c.tan_fov_x * ox;
  const float beta = c.tan_fov_y * oy;
  d = normalize((alpha * c.u + beta * c.v) - c.w);
  o = c.eye;
}

// muon/sampling.cc:86-108: sample 0 at the pixel centre, later samples
// jittered uniformly inside the pixel (counter-based draws instead of the
// per-tile mt19937 stream), with the same guard against rounding onto the
// next pixel.
__device__ __forceinline__ void sample_position(int px, int py, int sidx, PathKey key, float &x, float &y) {
  if (sidx == 0) {
    x = float(px) + 0.5f;
    y = float(py) + 0.5f;
    return;
  }
  uint32_t ctr = 0x40000000u;  // draw indices reserved for the sampler
  const float u = rng_uniform(key, ctr);
  const float v = rng_uniform(key, ctr);
  x = float(px) + u;
  y = float(py) + v;
  if (x == float(px) + 1.0f) x = nextafterf(x, -__builtin_inff());
  if (y == float(py) + 1.0f) y = nextafterf(y, -__builtin_inff());
}

class Scheduler;

// Per-context device workspace: grown on demand, reused across renders.
struct RenderWorkspace {
  DevBuf counters;
  DevBuf film;  // staging film of muonx_render_host
  // path-tracer engine of muonx_render / muonx_render_batch
  Scheduler *sched = nullptr;
  muonx_tuning sched_tuning{};
  muonx_engine_config sched_want{};
  bool sched_plain = false;  // built by the plain path (muonx_render), not for a batch configuration
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

RenderWorkspace *GetWorkspace(muonx_ctx *ctx);
void DestroyContextScheduler(RenderWorkspace *ws);
CamParams MakeCam(const muonx_render_params &p);
// Counters of the path-tracer engine: primary, secondary, shadow and culled
// MIS rays, then the traversals' primitive tests and hits and child boxes
// tested and passed (in the device tree).
constexpr int kEngineCounters = 8;
// One path-tracer job on the context's engine, ordered after earlier work on
// `stream` (later work on `stream` waits for the film); its counters in rays.
int RenderPathTracer(muonx_scene *scene, const muonx_render_params &p, float *accum, hipStream_t stream,
                     unsigned long long rays[kEngineCounters]);

}  // namespace muonx

#endif
