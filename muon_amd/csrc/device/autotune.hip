// Autotuner: measures candidate launch configurations on the device and
// installs the fastest in the context's tuning. No counterpart in the
// reference (its only knob is --parallelism).
//
// Traversal (BVH node width and box precision, stack form, while-while loop,
// blocks per CU, per-lane ray fetch) is timed on incoherent closest-hit rays generated over
// the scene's bounds: origins uniform in the box, directions uniform on the
// sphere (counter-based draws). With a render job, the engine's settings
// (node precision, fetch threshold, blocks per CU, path-queue capacity) are
// then timed on
// that job (at most 2^26 samples) and the job's best is installed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "query.h"
#include "render.h"
#include "rng.h"
#include "runtime.h"

namespace muonx {
namespace {

__global__ void k_tune_rays(float *o, float *d, uint32_t n, float x0, float y0, float z0, float x1, float y1,
                            float z1, uint32_t seed) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const PathKey key = path_key(seed, 0x7475u, i, 0);
    uint32_t ctr = 0;
    const float u0 = rng_uniform(key, ctr), u1 = rng_uniform(key, ctr), u2 = rng_uniform(key, ctr);
    const float u3 = rng_uniform(key, ctr), u4 = rng_uniform(key, ctr);
    o[3 * i] = x0 + (x1 - x0) * u0;
    o[3 * i + 1] = y0 + (y1 - y0) * u1;
    o[3 * i + 2] = z0 + (z1 - z0) * u2;
    const float z = 1.0f - 2.0f * u3, phi = 6.28318530718f * u4;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    d[3 * i] = r * cp;
    d[3 * i + 1] = r * sp;
    d[3 * i + 2] = z;
  }
}

struct Candidate {
  muonx_tuning t;
  float ms;
};

}  // namespace
}  // namespace muonx

using namespace muonx;

extern "C" int muonx_autotune(muonx_scene *scene, const muonx_render_params *job, uint32_t n_rays,
                              muonx_autotune_report *report) {
  if (!scene) {
    SetError("muonx_autotune: null scene");
    return MUONX_ERR_INVALID;
  }
  if (n_rays == 0) n_rays = 1u << 22;
  if (n_rays > (1u << 26)) {
    SetError("muonx_autotune: at most 2^26 sample rays");
    return MUONX_ERR_INVALID;
  }
  muonx_ctx *ctx = scene->ctx;
  muonx_tuning base;
  DevBuf rays, prim, t;
  std::vector<Candidate> cands;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MUONX_OK;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    MX_HIP(hipSetDevice(ctx->device));
    base = ctx->tuning;
    MX_HIP(rays.Alloc(size_t(n_rays) * 6 * sizeof(float)));
    MX_HIP(prim.Alloc(size_t(n_rays) * sizeof(int32_t)));
    MX_HIP(t.Alloc(size_t(n_rays) * sizeof(float)));
    MX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    MX_HIP(hipEventCreate(&e0));
    MX_HIP(hipEventCreate(&e1));
    float *o = rays.as<float>(), *d = o + size_t(n_rays) * 3;
    const float *b = scene->bounds;
    hipLaunchKernelGGL(k_tune_rays, dim3(1024), dim3(256), 0, stream, o, d, n_rays, b[0], b[1], b[2], b[3], b[4],
                       b[5], 12345u);
    MX_HIP(hipGetLastError());
    // candidate traversal configurations
    for (int width : {4, 2})
      for (int bits : {32, 8})
      for (int stack : {8, 16, 0})
        for (int ww : {1, 0})
          for (int bpc : {8, 16, 24})
            for (int fetch : {0, 32, 48}) {
              if (fetch > 0 && ww == 0) continue;  // the fetch kernels always walk while-while
              if (width == 2 && bits == 8) continue;  // quantized boxes exist for 4-wide nodes
              muonx_tuning c = base;
              c.bvh_width = width;
              c.node_bits = bits;
              c.stack_entries = stack;
              c.while_while = ww;
              c.blocks_per_cu = bpc;
              c.ray_fetch = fetch;
              ctx->tuning = c;
              float best = 1e30f;
              for (int rep = 0; rep < 3 && rc == MUONX_OK; ++rep) {
                MX_HIP(hipEventRecord(e0, stream));
                rc = LaunchQuery(ctx, scene->view, false, false, o, d, nullptr, n_rays, prim.as<int32_t>(),
                                 t.as<float>(), nullptr, nullptr, stream);
                MX_HIP(hipEventRecord(e1, stream));
                MX_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                MX_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) best = std::min(best, ms);  // the first launch warms caches
              }
              if (rc != MUONX_OK) break;
              cands.push_back({c, best});
            }
    if (rc == MUONX_OK) {
      const auto it = std::min_element(cands.begin(), cands.end(),
                                       [](const Candidate &x, const Candidate &y) { return x.ms < y.ms; });
      ctx->tuning = it->t;
    } else {
      ctx->tuning = base;
    }
  }
  float render_ms = 0.0f;
  if (rc == MUONX_OK && job) {
    // path-queue capacity on the job itself: the capacity matters when the
    // job has more samples than paths in flight, so a small proxy job would
    // not show it; spp is only reduced to keep it under 2^26 samples
    muonx_render_params p = *job;
    const uint64_t npix = uint64_t(std::max(1, p.width)) * uint64_t(std::max(1, p.height));
    p.spp = int(std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(std::max(1, p.spp)), (uint64_t(1) << 26) / npix)));
    DevBuf accum;
    {
      std::lock_guard<std::mutex> lock(ctx->mu);
      MX_HIP(accum.Alloc(size_t(p.width) * size_t(p.height) * 3 * sizeof(float)));
    }
    float best = 1e30f;
    const muonx_tuning tuned = ctx->tuning;
    muonx_tuning best_t = tuned;
    // the engine's launch settings on the job: node precision, fetch
    // threshold, grid, capacity
    for (int bits : {32, 8})
    for (int fetch : {0, 48})
      for (int bpc : {16, 24})
        for (int mb : {2048, 8192}) {
          muonx_tuning c = tuned;
          if (c.bvh_width == 4) c.node_bits = bits;
          else if (bits == 8) continue;
          c.ray_fetch = fetch;
          c.blocks_per_cu = bpc;
          c.path_budget_mb = mb;
          {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->tuning = c;
          }
          muonx_trace_stats st;
          if ((rc = muonx_render(scene, &p, accum.as<float>(), &st, stream)) != MUONX_OK) break;  // engine + graphs
          const auto t0 = std::chrono::steady_clock::now();
          if ((rc = muonx_render(scene, &p, accum.as<float>(), &st, stream)) != MUONX_OK) break;
          MX_HIP(hipStreamSynchronize(stream));
          const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
          if (ms < best) {
            best = ms;
            best_t = c;
          }
        }
    render_ms = best;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tuning = rc == MUONX_OK ? best_t : base;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (stream) (void)hipStreamDestroy(stream);
  if (rc != MUONX_OK) return rc;
  if (report) {
    report->tuning = ctx->tuning;
    report->candidates = uint32_t(cands.size());
    const auto best = std::min_element(cands.begin(), cands.end(),
                                       [](const Candidate &x, const Candidate &y) { return x.ms < y.ms; });
    const auto worst = std::max_element(cands.begin(), cands.end(),
                                        [](const Candidate &x, const Candidate &y) { return x.ms < y.ms; });
    float base_ms = 0.0f;
    for (const Candidate &c : cands)
      if (c.t.bvh_width == base.bvh_width && c.t.node_bits == base.node_bits && c.t.stack_entries == base.stack_entries &&
          c.t.while_while == base.while_while && c.t.blocks_per_cu == base.blocks_per_cu &&
          c.t.ray_fetch == base.ray_fetch)
        base_ms = c.ms;
    report->sample_rays = n_rays;
    report->best_ms = best->ms;
    report->worst_ms = worst->ms;
    report->previous_ms = base_ms;
    report->render_ms = render_ms;
  }
  return MUONX_OK;
}
