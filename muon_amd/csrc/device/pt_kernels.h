// Kernel arguments of the wavefront path tracer.
#ifndef MUONX_DEVICE_PT_KERNELS_H_
#define MUONX_DEVICE_PT_KERNELS_H_

#include "dev_scene.h"
#include "render.h"

namespace muonx {

constexpr int kPtBlock = 128;
// Slots of the u32 queue counters.
constexpr int kQCur = 0;   // ray queue 0
constexpr int kQNext = 1;  // ray queue 1
constexpr int kQMis = 2;   // MIS ray queue

// One generation: samples [s0, s0+nsamp) of pixels [p0, p0+npix); path slot
// p maps to pixel p0 + p % npix and sample s0 + p / npix.
struct PtGen {
  uint32_t p0, npix, s0, nsamp, npaths;
};

struct PtArgs {
  DevSceneView scene;
  CamParams cam;
  PtGen gen;
  uint32_t seed, job;
  int min_depth, cutoff, nee, is_mode, rr, light_samples, strata, per_section, shadow_per_vertex;
  // path state (SoA over path slots)
  float *rad, *thr, *m_fac, *m_add;
  uint32_t *ctr;
  // queues
  float4 *q_ray[2];  // 2 float4 per entry: (o.xyz, path), (d.xyz, 0)
  float4 *m_ray;     // MIS rays, same layout
  int32_t *hit_prim; // per entry of the queue being shaded
  // shadow slots (path * K + k)
  float4 *s_org, *s_dir;  // (o.xyz, 0), (d.xyz, tmax)
  float *s_contrib;
  uint8_t *s_vis, *s_valid;
  uint32_t *counts;
  unsigned long long *ray_ctr;
};

}  // namespace muonx

#endif
