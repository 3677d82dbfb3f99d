// muonx_render: camera ray generation + integrators on the GPU.
// Deterministic integrators (normals / albedo / depth, muon/integration.cc:47-76)
// run as one fused kernel with the exact traversal mode (the reference's
// node visits, so ties and edge hits resolve as in muon); the path tracer runs as a wavefront pipeline
// (pathtracer.hip).
#include <chrono>

#include "hitrec.h"
#include "render.h"
#include "rng.h"
#include "runtime.h"

namespace muonx {

void DestroyWorkspace(RenderWorkspace *ws) {
  if (!ws) return;  // the context never rendered
  DestroyContextScheduler(ws);
  if (ws->ev_in) (void)hipEventDestroy(ws->ev_in);
  if (ws->ev_out) (void)hipEventDestroy(ws->ev_out);
  delete ws;
}

RenderWorkspace *GetWorkspace(muonx_ctx *ctx) {
  if (!ctx->ws) ctx->ws = new RenderWorkspace();
  return ctx->ws;
}

CamParams MakeCam(const muonx_render_params &p) {
  CamParams c;
  c.eye = V3{p.eye[0], p.eye[1], p.eye[2]};
  c.u = V3{p.cam_u[0], p.cam_u[1], p.cam_u[2]};
  c.v = V3{p.cam_v[0], p.cam_v[1], p.cam_v[2]};
  c.w = V3{p.cam_w[0], p.cam_w[1], p.cam_w[2]};
  c.tan_fov_x = p.tan_fov_x;
  c.tan_fov_y = p.tan_fov_y;
  c.width = p.width;
  c.height = p.height;
  return c;
}

template <int STACK>
__global__ void __launch_bounds__(kSimpleBlock) k_simple(DevSceneView s, CamParams cam, int integrator, int spp,
                                                         int depth0_ok, uint32_t seed, float *__restrict__ accum,
                                                         unsigned long long *__restrict__ rays) {
  __shared__ uint2 stack[STACK * kSimpleBlock];
  const uint32_t pixel = blockIdx.x * kSimpleBlock + threadIdx.x;
  const uint32_t npix = uint32_t(cam.width) * uint32_t(cam.height);
  uint32_t traced = 0;
  if (pixel < npix) {
    const int px = int(pixel % uint32_t(cam.width)), py = int(pixel / uint32_t(cam.width));
    V3 acc{0.0f, 0.0f, 0.0f};
    for (int sidx = 0; sidx < spp && depth0_ok; ++sidx) {
      float x, y;
      sample_position(px, py, sidx, path_key(seed, 0, pixel, uint32_t(sidx)), x, y);
      V3 o, d;
      cast_ray(cam, x, y, o, d);
      const TraceRay r = make_ray(o, d);
      float best = __builtin_inff();
      TraceCounters cnt;
      ++traced;
      const int32_t p = trace_bvh<STACK, false, false, true>(s, r, best, stack + threadIdx.x, kSimpleBlock, cnt);
      V3 c{0.0f, 0.0f, 0.0f};
      if (p >= 0) {
        if (integrator == MUONX_INTEGRATOR_ALBEDO) {
          const float *m = s.materials + size_t(s.prim_mat[p]) * kMatStride;
          c = V3{m[3], m[4], m[5]} + V3{m[6], m[7], m[8]};
        } else {
          const HitRecord h = hit_record(s, uint32_t(p), o, d);
          if (integrator == MUONX_INTEGRATOR_NORMALS) {
            c = h.normal * 0.5f + 0.5f;
          } else {
            const float z = 1.0f / (1.0f + h.dist);
            c = V3{z, z, z};
          }
        }
      }
      acc = acc + c;
    }
    accum[3 * size_t(pixel)] = acc.x;
    accum[3 * size_t(pixel) + 1] = acc.y;
    accum[3 * size_t(pixel) + 2] = acc.z;
  }
  unsigned long long a = traced;
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off);
  if ((threadIdx.x & 63) == 0 && a) atomicAdd(rays, a);
}

int RenderSimple(muonx_scene *scene, const muonx_render_params &p, float *accum, unsigned long long *rays,
                 hipStream_t stream) {
  const DevSceneView &s = scene->view;
  const CamParams cam = MakeCam(p);
  // muon/integration.cc:29-33 applied at depth 0
  const int cutoff = p.nee != MUONX_NEE_OFF ? p.max_depth - 1 : p.max_depth;
  const int depth0_ok = (p.max_depth == -1 || 0 <= cutoff) ? 1 : 0;
  const uint32_t npix = uint32_t(p.width) * uint32_t(p.height);
  const dim3 grid((npix + kSimpleBlock - 1) / kSimpleBlock);
  if (s.bvh_depth <= 24) {
    hipLaunchKernelGGL(k_simple<24>, grid, dim3(kSimpleBlock), 0, stream, s, cam, p.integrator, p.spp, depth0_ok,
                       p.seed, accum, rays);
  } else if (s.bvh_depth <= 48) {
    hipLaunchKernelGGL(k_simple<48>, grid, dim3(kSimpleBlock), 0, stream, s, cam, p.integrator, p.spp, depth0_ok,
                       p.seed, accum, rays);
  } else if (s.bvh_depth <= 96) {
    hipLaunchKernelGGL(k_simple<96>, grid, dim3(kSimpleBlock), 0, stream, s, cam, p.integrator, p.spp, depth0_ok,
                       p.seed, accum, rays);
  } else {
    SetError("BVH too deep for the traversal stack");
    return MUONX_ERR_CAPACITY;
  }
  MX_HIP(hipGetLastError());
  return MUONX_OK;
}

}  // namespace muonx

using namespace muonx;

extern "C" int muonx_render(muonx_scene *scene, const muonx_render_params *params, float *accum,
                            muonx_trace_stats *stats, void *stream_) {
  if (!scene || !params || !accum) {
    SetError("muonx_render: null argument");
    return MUONX_ERR_INVALID;
  }
  const muonx_render_params &p = *params;
  if (p.width <= 0 || p.height <= 0 || p.spp <= 0) {
    SetError("muonx_render: width, height and spp must be positive");
    return MUONX_ERR_INVALID;
  }
  if (!p.has_camera) {
    SetError("muonx_render: scene has no camera");  // the reference dereferences a null camera
    return MUONX_ERR_INVALID;
  }
  muonx_ctx *ctx = scene->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  MX_HIP(hipSetDevice(ctx->device));
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  RenderWorkspace *ws = GetWorkspace(ctx);
  MX_HIP(ws->counters.Reserve(sizeof(unsigned long long) * kNumCounters));
  MX_HIP(hipMemsetAsync(ws->counters.ptr, 0, sizeof(unsigned long long) * kNumCounters, stream));
  hipEvent_t e0, e1;
  MX_HIP(hipEventCreate(&e0));
  MX_HIP(hipEventCreate(&e1));
  MX_HIP(hipEventRecord(e0, stream));
  int rc;
  unsigned long long *ctr = ws->counters.as<unsigned long long>();
  bool host_rays = false;
  unsigned long long rays[3] = {0, 0, 0};
  switch (p.integrator) {
    case MUONX_INTEGRATOR_NORMALS:
    case MUONX_INTEGRATOR_ALBEDO:
    case MUONX_INTEGRATOR_DEPTH:
      rc = RenderSimple(scene, p, accum, ctr + kCtrPrimary, stream);
      break;
    case MUONX_INTEGRATOR_PATHTRACER:
      rc = RenderPathTracer(scene, p, accum, stream, rays);
      host_rays = true;
      break;
    default:
      SetError("muonx_render: integrator " + std::to_string(p.integrator) + " is not on the GPU render path");
      rc = MUONX_ERR_INVALID;
  }
  if (rc != MUONX_OK) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }
  MX_HIP(hipEventRecord(e1, stream));
  if (stats) {
    unsigned long long h[kNumCounters];
    MX_HIP(hipMemcpyAsync(h, ws->counters.ptr, sizeof(h), hipMemcpyDeviceToHost, stream));
    MX_HIP(hipStreamSynchronize(stream));
    float ms = 0.0f;
    MX_HIP(hipEventElapsedTime(&ms, e0, e1));
    stats->primary_rays = host_rays ? rays[0] : h[kCtrPrimary];
    stats->secondary_rays = host_rays ? rays[1] : h[kCtrSecondary];
    stats->shadow_rays = host_rays ? rays[2] : h[kCtrShadow];
    stats->render_ns = uint64_t(double(ms) * 1e6);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return MUONX_OK;
}

extern "C" int muonx_render_host(muonx_scene *scene, const muonx_render_params *params, float *accum_host,
                                 muonx_trace_stats *stats) {
  if (!scene || !params || !accum_host) {
    SetError("muonx_render_host: null argument");
    return MUONX_ERR_INVALID;
  }
  if (params->width <= 0 || params->height <= 0) {
    SetError("muonx_render_host: width and height must be positive");
    return MUONX_ERR_INVALID;
  }
  const size_t bytes = size_t(params->width) * size_t(params->height) * 3 * sizeof(float);
  float *dev = nullptr;
  {
    std::lock_guard<std::mutex> lock(scene->ctx->mu);
    MX_HIP(hipSetDevice(scene->ctx->device));
    RenderWorkspace *ws = GetWorkspace(scene->ctx);
    MX_HIP(ws->film.Reserve(bytes));
    dev = ws->film.as<float>();
  }
  const int rc = muonx_render(scene, params, dev, stats, nullptr);
  if (rc != MUONX_OK) return rc;
  MX_HIP(hipMemcpy(accum_host, dev, bytes, hipMemcpyDeviceToHost));
  return MUONX_OK;
}
