// muonx_render: camera ray generation + integrators on the GPU.
// Deterministic integrators (normals / albedo / depth, muon/integration.cc:47-76)
// run as one fused kernel with the exact traversal mode (the reference's
// node visits, so ties and edge hits resolve as in muon); the path tracer runs as a wavefront pipeline
// (pathtracer.hip).
#include <chrono>

#include <climits>

#include "hitrec.h"
#include "render.h"
#include "wave.h"
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
                                                         unsigned long long *__restrict__ ctr) {
  __shared__ uint2 stack[STACK * kSimpleBlock];
  const uint32_t pixel = blockIdx.x * kSimpleBlock + threadIdx.x;
  const uint32_t npix = uint32_t(cam.width) * uint32_t(cam.height);
  uint32_t traced = 0;
  TraceCounters cnt;
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
      ++traced;
      const int32_t p = trace_bvh<STACK, false, true, true>(s, r, best, stack + threadIdx.x, kSimpleBlock, cnt);
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
  wave_count(traced, &ctr[kCtrPrimary]);
  wave_count(cnt.prims, &ctr[kCtrObjectTests]);
  wave_count(cnt.ohits, &ctr[kCtrObjectHits]);
  wave_count(cnt.btests, &ctr[kCtrBoundsTests]);
  wave_count(cnt.bhits, &ctr[kCtrBoundsHits]);
}

// Raytracer (muon/integration.cc:78-120) and AnalyticDirect
// (muon/integration.cc:126-199), deterministic, with the exact traversal
// mode. The raytracer's reflection recursion, color_d = local_d + ks_d *
// color_{d+1}, is evaluated back to front from the stored per-bounce terms so
// that its float operations are the reference's.
constexpr int kWhittedMaxBounces = 32;  // bound for max_depth = -1

// Light::ShadingInfoAt (muon/lighting.cc:7-72); a quad light leaves the
// direction and distance zero, as the reference's restatement does.
__device__ __forceinline__ void shading_info(const float *L, V3 p, V3 &color, V3 &dir, float &dist) {
  const int kind = int(L[0]);
  color = V3{L[1], L[2], L[3]};
  if (kind == MUONX_LIGHT_DIRECTIONAL) {
    dir = V3{L[4], L[5], L[6]};
    dist = __builtin_inff();
  } else if (kind == MUONX_LIGHT_POINT) {
    const V3 to_light = V3{L[4], L[5], L[6]} - p;
    dir = normalize(to_light);
    const float r = length(to_light);
    const float att = L[7] + L[8] * r + L[9] * r * r;
    color = color / att;
    dist = r;
  } else {
    dir = V3{0.0f, 0.0f, 0.0f};
    dist = 0.0f;
  }
}

template <int STACK>
__global__ void __launch_bounds__(kSimpleBlock) k_whitted(DevSceneView s, CamParams cam, int integrator, int spp,
                                                          int cutoff, uint32_t seed, float *__restrict__ accum,
                                                          unsigned long long *__restrict__ rays) {
  __shared__ uint2 stack[STACK * kSimpleBlock];
  const uint32_t pixel = blockIdx.x * kSimpleBlock + threadIdx.x;
  const uint32_t npix = uint32_t(cam.width) * uint32_t(cam.height);
  uint32_t n_primary = 0, n_secondary = 0, n_shadow = 0;
  TraceCounters cnt;
  if (pixel < npix) {
    const int px = int(pixel % uint32_t(cam.width)), py = int(pixel / uint32_t(cam.width));
    V3 acc{0.0f, 0.0f, 0.0f};
    V3 local[kWhittedMaxBounces], ks[kWhittedMaxBounces];
    for (int sidx = 0; sidx < spp; ++sidx) {
      float x, y;
      sample_position(px, py, sidx, path_key(seed, 0, pixel, uint32_t(sidx)), x, y);
      V3 o, d;
      cast_ray(cam, x, y, o, d);
      int nb = 0;       // bounces with a hit
      bool more = true;
      for (int depth = 0; more && depth < kWhittedMaxBounces; ++depth) {
        more = false;
        // Integrator::Trace returns 0 past the cutoff (muon/integration.cc:29-33)
        if (cutoff != INT_MAX && depth > cutoff) break;
        if (depth == 0) ++n_primary;
        else ++n_secondary;
        const TraceRay r = make_ray(o, d);
        float best = __builtin_inff();
        const int32_t p = trace_bvh<STACK, false, true, true>(s, r, best, stack + threadIdx.x, kSimpleBlock, cnt);
        if (p < 0) break;
        const HitRecord h = hit_record(s, uint32_t(p), o, d);
        const float *m = s.materials + size_t(s.prim_mat[p]) * kMatStride;
        const V3 amb{m[0], m[1], m[2]}, kd{m[3], m[4], m[5]}, spec{m[6], m[7], m[8]}, em{m[9], m[10], m[11]};
        const V3 shift = h.pos + kEps * h.normal;
        V3 color;
        if (integrator == MUONX_INTEGRATOR_RAYTRACER) {
          color = amb + em;
          for (uint32_t l = 0; l < s.num_lights; ++l) {
            V3 lc, ldir;
            float ldist;
            shading_info(s.lights + size_t(l) * kLightStride, h.pos, lc, ldir, ldist);
            ++n_shadow;
            const TraceRay sr = make_ray(shift, ldir);
            float sbest = ldist;
            const int32_t occ = trace_bvh<STACK, true, true, true>(s, sr, sbest, stack + threadIdx.x, kSimpleBlock,
                                                                   cnt);
            if (occ >= 0) continue;
            const V3 diffuse = kd * gmax(dot(h.normal, ldir), 0.0f);
            const V3 half = normalize(ldir - d);
            const V3 sp = spec * powf(gmax(dot(h.normal, half), 0.0f), m[12]);
            color = color + lc * (diffuse + sp);
          }
          local[nb] = color;
          ks[nb] = spec;
          ++nb;
          if (spec.x > 0.0f || spec.y > 0.0f || spec.z > 0.0f) {
            o = shift;
            d = reflect(d, h.normal);
            more = true;
          }
        } else {
          color = em;
          for (uint32_t l = 0; l < s.num_lights; ++l) {
            const float *L = s.lights + size_t(l) * kLightStride;
            if (int(L[0]) != MUONX_LIGHT_QUAD) continue;
            const V3 c0{L[4], L[5], L[6]}, e0{L[7], L[8], L[9]}, e1{L[10], L[11], L[12]};
            const V3 verts[4] = {c0, c0 + e0, (c0 + e0) + e1, c0 + e1};
            const V3 brdf = kd * kOneOverPi;
            const V3 radiance{L[1], L[2], L[3]};
            V3 irr{0.0f, 0.0f, 0.0f};
            for (int i = 0; i < 4; ++i) {
              const int ni = (i + 1) % 4;
              const V3 ui = normalize(verts[i] - h.pos);
              const V3 un = normalize(verts[ni] - h.pos);
              const float theta = acosf(dot(ui, un));
              const V3 gamma = normalize(cross(ui, un));
              irr = irr + theta * gamma;
            }
            irr = irr * 0.5f;
            color = color + (brdf * radiance) * dot(irr, h.normal);
          }
          local[nb] = color;
          ks[nb] = V3{0.0f, 0.0f, 0.0f};
          ++nb;
        }
      }
      V3 c{0.0f, 0.0f, 0.0f};
      if (nb > 0) {
        // color_d = local_d + ks_d * color_{d+1}, innermost first; a bounce
        // whose reflected ray found nothing adds ks * 0
        c = local[nb - 1];
        if (ks[nb - 1].x > 0.0f || ks[nb - 1].y > 0.0f || ks[nb - 1].z > 0.0f) c = c + ks[nb - 1] * V3{0.0f, 0.0f, 0.0f};
        for (int b = nb - 2; b >= 0; --b) c = local[b] + ks[b] * c;
      }
      acc = acc + c;
    }
    accum[3 * size_t(pixel)] = acc.x;
    accum[3 * size_t(pixel) + 1] = acc.y;
    accum[3 * size_t(pixel) + 2] = acc.z;
  }
  wave_count(n_primary, &rays[kCtrPrimary]);
  wave_count(n_secondary, &rays[kCtrSecondary]);
  wave_count(n_shadow, &rays[kCtrShadow]);
  wave_count(cnt.prims, &rays[kCtrObjectTests]);
  wave_count(cnt.ohits, &rays[kCtrObjectHits]);
  wave_count(cnt.btests, &rays[kCtrBoundsTests]);
  wave_count(cnt.bhits, &rays[kCtrBoundsHits]);
}

int RenderWhitted(muonx_scene *scene, const muonx_render_params &p, float *accum, unsigned long long *rays,
                  hipStream_t stream) {
  const DevSceneView &s = scene->view;
  const CamParams cam = MakeCam(p);
  const int cutoff =
      p.max_depth == -1 ? INT_MAX : (p.nee != MUONX_NEE_OFF ? p.max_depth - 1 : p.max_depth);
  const uint32_t npix = uint32_t(p.width) * uint32_t(p.height);
  const dim3 grid((npix + kSimpleBlock - 1) / kSimpleBlock);
#define MX_W(S)                                                                                                     \
  hipLaunchKernelGGL(k_whitted<S>, grid, dim3(kSimpleBlock), 0, stream, s, cam, p.integrator, p.spp, cutoff, p.seed, \
                     accum, rays)
  if (s.bvh_depth <= 24) MX_W(24);
  else if (s.bvh_depth <= 48) MX_W(48);
  else if (s.bvh_depth <= 96) MX_W(96);
  else {
    SetError("BVH too deep for the traversal stack");
    return MUONX_ERR_CAPACITY;
  }
#undef MX_W
  MX_HIP(hipGetLastError());
  return MUONX_OK;
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
  if (p.integrator != MUONX_INTEGRATOR_PATHTRACER && !scene->reference_tree) {
    SetError("muonx_render: this integrator traverses the reference's SAH tree; the scene's BVH was built on "
             "the device");
    return MUONX_ERR_STATE;
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
  unsigned long long rays[kEngineCounters] = {};
  switch (p.integrator) {
    case MUONX_INTEGRATOR_NORMALS:
    case MUONX_INTEGRATOR_ALBEDO:
    case MUONX_INTEGRATOR_DEPTH:
      rc = RenderSimple(scene, p, accum, ctr, stream);
      break;
    case MUONX_INTEGRATOR_RAYTRACER:
    case MUONX_INTEGRATOR_ANALYTICDIRECT:
      rc = RenderWhitted(scene, p, accum, ctr, stream);
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
    stats->culled_rays = host_rays ? rays[3] : 0;
    // the path tracer's: the work of its traversals in the device tree
    stats->object_tests = host_rays ? rays[4] : h[kCtrObjectTests];
    stats->object_hits = host_rays ? rays[5] : h[kCtrObjectHits];
    stats->bounds_tests = host_rays ? rays[6] : h[kCtrBoundsTests];
    stats->bounds_hits = host_rays ? rays[7] : h[kCtrBoundsHits];
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
