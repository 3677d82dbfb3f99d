// Wavefront path tracer (PathTracer, muon/integration.cc:205-630) for gfx950.
//
// The reference recurses per camera sample. Here every (pixel, sample) pair
// of a generation is a path slot in struct-of-arrays state, and each bounce
// runs five kernels over compacted queues:
//   extend  closest hit of the path rays                   (Integrator::Trace)
//   shade   emission, NEE light samples -> shadow queue, MIS BRDF ray -> MIS
//           queue, BRDF-sampled continuation (+ Russian roulette) -> next
//           ray queue                                      (PathTracer::Shade)
//   shadow  any-hit of the shadow queue -> visibility flags (HasIntersection)
//   mis     closest hit of the MIS queue; light hits add the MIS-weighted
//           emission                      (ShadeDirectImportanceSampling)
//   gather  adds each path's visible NEE and MIS terms in a fixed order
// and a resolve kernel adds every sample's radiance into the film in sample
// order, so results do not depend on queue order. Queues are appended with
// one atomic per wavefront (ballot + popcount prefix).
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "hitrec.h"
#include "pt_kernels.h"
#include "render.h"
#include "shading.h"
#include "traverse_launch.h"

namespace muonx {

__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter) {
  const uint64_t active = __ballot(1);
  const uint64_t mask = __ballot(want);
  const uint32_t lane = __lane_id();
  const uint32_t leader = uint32_t(__ffsll((unsigned long long)active) - 1);
  uint32_t base = 0;
  if (lane == leader && mask) base = atomicAdd(counter, uint32_t(__popcll(mask)));
  base = uint32_t(__shfl(int(base), int(leader)));
  return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

__device__ __forceinline__ void wave_count(uint32_t n, unsigned long long *ctr) {
  unsigned long long a = n;
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off);
  if ((__lane_id()) == 0 && a) atomicAdd(ctr, a);
}

__device__ __forceinline__ uint32_t path_pixel(const PtGen &g, uint32_t p) { return g.p0 + p % g.npix; }
__device__ __forceinline__ uint32_t path_sample(const PtGen &g, uint32_t p) { return g.s0 + p / g.npix; }

// ---------------------------------------------------------------- raygen
__global__ void __launch_bounds__(kPtBlock) k_raygen(PtArgs a, PtGen g) {
  const uint32_t n = g.npaths;
  for (uint32_t p = blockIdx.x * kPtBlock + threadIdx.x; p < n; p += gridDim.x * kPtBlock) {
    const uint32_t pixel = path_pixel(g, p), sample = path_sample(g, p);
    const int px = int(pixel % uint32_t(a.cam.width)), py = int(pixel / uint32_t(a.cam.width));
    float x, y;
    const uint32_t key = path_key(a.seed, a.job, pixel, sample);
    sample_position(px, py, int(sample), key, x, y);
    V3 o, d;
    cast_ray(a.cam, x, y, o, d);
    a.q_ray[0][2 * p] = make_float4(o.x, o.y, o.z, __uint_as_float(p));
    a.q_ray[0][2 * p + 1] = make_float4(d.x, d.y, d.z, 0.0f);
    a.rad[3 * p] = 0.0f;
    a.rad[3 * p + 1] = 0.0f;
    a.rad[3 * p + 2] = 0.0f;
    a.thr[3 * p] = 1.0f;
    a.thr[3 * p + 1] = 1.0f;
    a.thr[3 * p + 2] = 1.0f;
    a.ctr[p] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.counts[kQCur] = n;
}

// ---------------------------------------------------------------- extend
template <int KIND, bool WW>
__global__ void __launch_bounds__(kPtBlock) k_extend(PtArgs a, StackArgs sa, int cur, int depth) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kPtBlock];
  const uint32_t n = a.counts[cur == 0 ? kQCur : kQNext];
  const float4 *q = a.q_ray[cur];
  uint32_t traced = 0;
  for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
    const uint32_t i = base + threadIdx.x;
    if (i < n) {
      const float4 qo = q[2 * i], qd = q[2 * i + 1];
      const TraceRay r = make_ray(V3{qo.x, qo.y, qo.z}, V3{qd.x, qd.y, qd.z});
      float best = __builtin_inff();
      TraceCounters c;
      auto st = make_stack<KIND>(lds, kPtBlock, sa);
      a.hit_prim[i] = trace_bvh_s<false, false, false, WW>(a.scene, r, best, st, c);
      ++traced;
    }
  }
  wave_count(traced, &a.ray_ctr[depth == 0 ? kCtrPrimary : kCtrSecondary]);
}

// ---------------------------------------------------------------- shade
__global__ void __launch_bounds__(kPtBlock) k_shade(PtArgs a, int cur, int depth) {
  const uint32_t n = a.counts[cur == 0 ? kQCur : kQNext];
  uint32_t *next_count = &a.counts[cur == 0 ? kQNext : kQCur];
  const float4 *q = a.q_ray[cur];
  float4 *qn = a.q_ray[cur ^ 1];
  const DevSceneView &s = a.scene;
  const int K = a.shadow_per_vertex;
  for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
    const uint32_t i = base + threadIdx.x;
    bool cont = false;
    bool want_mis = false;
    uint32_t p = 0;
    V3 shift{0, 0, 0}, next_dir{0, 0, 0}, mis_dir{0, 0, 0};
    if (i < n) {
      const float4 qo = q[2 * i], qd = q[2 * i + 1];
      p = __float_as_uint(qo.w);
      const int32_t prim = a.hit_prim[i];
      // clear this bounce's NEE / MIS terms of the path
      for (int k = 0; k < K; ++k) {
        a.s_vis[size_t(p) * K + k] = 0;
        a.s_valid[size_t(p) * K + k] = 0;
      }
      a.m_add[3 * p] = 0.0f;
      a.m_add[3 * p + 1] = 0.0f;
      a.m_add[3 * p + 2] = 0.0f;
      if (prim >= 0) {
        const V3 ro{qo.x, qo.y, qo.z}, rd{qd.x, qd.y, qd.z};
        const HitRecord h = hit_record(s, uint32_t(prim), ro, rd);
        const Mat m = load_mat(s.materials + size_t(s.prim_mat[prim]) * kMatStride);
        const uint32_t pixel = path_pixel(a.gen, p), sample = path_sample(a.gen, p);
        const uint32_t key = path_key(a.seed, a.job, pixel, sample);
        uint32_t ctr = a.ctr[p];
        V3 T{a.thr[3 * p], a.thr[3 * p + 1], a.thr[3 * p + 2]};
        V3 rad{a.rad[3 * p], a.rad[3 * p + 1], a.rad[3 * p + 2]};
        shift = h.pos + kEps * h.normal;
        if (depth >= a.min_depth) {
          // emission (muon/integration.cc:257-264)
          if (!((a.nee != MUONX_NEE_OFF && depth > 0) || dot(h.normal, -rd) < 0.0f)) rad = rad + T * m.emission;
          if (a.nee != MUONX_NEE_OFF) {
            // ShadeDirectNEE (muon/integration.cc:340-466): one shadow ray per light sample.
            const bool mis = a.nee == MUONX_NEE_MIS;
            int k = 0;
            for (uint32_t l = 0; l < s.num_lights; ++l) {
              const float *L = s.lights + size_t(l) * kLightStride;
              const int kind = int(L[0]);
              if (kind != MUONX_LIGHT_QUAD) {
                V3 ldir, col{L[1], L[2], L[3]};
                float ldist;
                if (kind == MUONX_LIGHT_POINT) {
                  const V3 to_light = V3{L[4], L[5], L[6]} - h.pos;
                  ldir = normalize(to_light);
                  ldist = length(to_light);
                  const float att = L[7] + L[8] * ldist + L[9] * ldist * ldist;
                  col = col / att;
                } else {
                  ldir = V3{L[4], L[5], L[6]};
                  ldist = __builtin_inff();
                }
                V3 irr = col;
                if (ldist < __builtin_inff()) irr = irr / (ldist * ldist);
                const float cosi = gmax(dot(h.normal, ldir), 0.0f);
                const V3 contrib = T * (irr * cosi * brdf_eval(m, ldir, rd, h.normal));
                const size_t slot = size_t(p) * K + k;
                a.s_contrib[3 * slot] = contrib.x;
                a.s_contrib[3 * slot + 1] = contrib.y;
                a.s_contrib[3 * slot + 2] = contrib.z;
                a.s_dir[slot] = make_float4(ldir.x, ldir.y, ldir.z, ldist);
                a.s_org[slot] = make_float4(shift.x, shift.y, shift.z, 0.0f);
                a.s_valid[slot] = 1;
                ++k;
                continue;
              }
              const QuadLight ql = load_quad(L);
              const V3 rn = normalize(cross(ql.e0, ql.e1));
              const float scale = ql.area / float(a.light_samples);
              for (int si = 0; si < a.strata; ++si) {
                for (int sj = 0; sj < a.strata; ++sj) {
                  for (int sk = 0; sk < a.per_section; ++sk) {
                    const float u = rng_uniform(key, ctr);
                    const float v = rng_uniform(key, ctr);
                    V3 lp = (ql.corner + (float(si) + u) / float(a.strata) * ql.e0) +
                            (float(sj) + v) / float(a.strata) * ql.e1;
                    lp = lp - kEps * rn;
                    const V3 to_light = lp - shift;
                    const V3 ldir = normalize(to_light);
                    const float ldist = length(to_light);
                    const float cosi = gmax(dot(h.normal, ldir), 0.0f);
                    const float cosl = gmax(dot(rn, ldir), 0.0f);
                    const float g = cosi * cosl / (ldist * ldist);
                    V3 c = g * brdf_eval(m, ldir, rd, h.normal);
                    if (mis) {
                      const float lpdf = nee_pdf(s, shift, ldir);
                      const float bpdf = is_pdf(a.is_mode, m, ldir, rd, h.normal);
                      c = c * power_heuristic(lpdf, bpdf);
                    }
                    const V3 contrib = T * (ql.color * (c * scale));
                    const size_t slot = size_t(p) * K + k;
                    a.s_contrib[3 * slot] = contrib.x;
                    a.s_contrib[3 * slot + 1] = contrib.y;
                    a.s_contrib[3 * slot + 2] = contrib.z;
                    a.s_dir[slot] = make_float4(ldir.x, ldir.y, ldir.z, ldist);
                    a.s_org[slot] = make_float4(shift.x, shift.y, shift.z, 0.0f);
                    // a zero contribution cannot change the image: skip its ray
                    a.s_valid[slot] = (contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f) ? 1 : 0;
                    ++k;
                  }
                }
              }
            }
            if (mis) {
              // ShadeDirectImportanceSampling (muon/integration.cc:468-501)
              V3 dir, nT;
              float bpdf;
              if (sample_reflection(a.is_mode, m, rd, h.normal, T, key, ctr, dir, bpdf, nT)) {
                const float lpdf = nee_pdf(s, shift, dir);
                const V3 f = nT * power_heuristic(bpdf, lpdf);
                a.m_fac[3 * p] = f.x;
                a.m_fac[3 * p + 1] = f.y;
                a.m_fac[3 * p + 2] = f.z;
                want_mis = true;
                mis_dir = dir;
              }
            }
          }
        }
        // ShadeIndirect (muon/integration.cc:294-328)
        V3 dir, nT;
        float pdf;
        if (sample_reflection(a.is_mode, m, rd, h.normal, T, key, ctr, dir, pdf, nT)) {
          bool alive = true;
          if (a.rr) {
            const float qc = gmin(gmax(gmax(nT.x, nT.y), nT.z), 1.0f);
            if (qc < rng_uniform(key, ctr)) alive = false;
            else nT = nT / qc;
          }
          // Trace() returns 0 past the depth cutoff (muon/integration.cc:29-33).
          if (alive && depth + 1 <= a.cutoff) {
            cont = true;
            next_dir = dir;
            T = nT;
          }
        }
        a.ctr[p] = ctr;
        a.thr[3 * p] = T.x;
        a.thr[3 * p + 1] = T.y;
        a.thr[3 * p + 2] = T.z;
        a.rad[3 * p] = rad.x;
        a.rad[3 * p + 1] = rad.y;
        a.rad[3 * p + 2] = rad.z;
      }
    }
    // compacted appends (all lanes of the wavefront participate)
    const uint32_t ni = wave_append(cont, next_count);
    if (cont) {
      qn[2 * ni] = make_float4(shift.x, shift.y, shift.z, __uint_as_float(p));
      qn[2 * ni + 1] = make_float4(next_dir.x, next_dir.y, next_dir.z, 0.0f);
    }
    const uint32_t mi = wave_append(want_mis, &a.counts[kQMis]);
    if (want_mis) {
      a.m_ray[2 * mi] = make_float4(shift.x, shift.y, shift.z, __uint_as_float(p));
      a.m_ray[2 * mi + 1] = make_float4(mis_dir.x, mis_dir.y, mis_dir.z, 0.0f);
    }
  }
}

// ---------------------------------------------------------------- shadow (any hit)
// Visits the K shadow slots of each shaded path.
template <int KIND, bool WW>
__global__ void __launch_bounds__(kPtBlock) k_shadow(PtArgs a, StackArgs sa, int cur) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kPtBlock];
  const uint32_t npaths_shaded = a.counts[cur == 0 ? kQCur : kQNext];
  const uint32_t K = uint32_t(a.shadow_per_vertex);
  const uint32_t n = npaths_shaded * K;
  const float4 *q = a.q_ray[cur];
  uint32_t traced = 0;
  for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
    const uint32_t i = base + threadIdx.x;
    if (i < n) {
      const uint32_t p = __float_as_uint(q[2 * (i / K)].w);
      const size_t slot = size_t(p) * K + (i % K);
      if (a.hit_prim[i / K] >= 0 && a.s_valid[slot]) {
        const float4 so = a.s_org[slot], sd = a.s_dir[slot];
        const TraceRay r = make_ray(V3{so.x, so.y, so.z}, V3{sd.x, sd.y, sd.z});
        float best = sd.w;
        TraceCounters c;
        auto st = make_stack<KIND>(lds, kPtBlock, sa);
        const int32_t hit = trace_bvh_s<true, false, false, WW>(a.scene, r, best, st, c);
        a.s_vis[slot] = hit < 0 ? 1 : 0;
        ++traced;
      }
    }
  }
  wave_count(traced, &a.ray_ctr[kCtrShadow]);
}

// ---------------------------------------------------------------- MIS extend
template <int KIND, bool WW>
__global__ void __launch_bounds__(kPtBlock) k_mis(PtArgs a, StackArgs sa) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kPtBlock];
  const uint32_t n = a.counts[kQMis];
  uint32_t traced = 0;
  for (uint32_t base = blockIdx.x * kPtBlock; base < n; base += gridDim.x * kPtBlock) {
    const uint32_t i = base + threadIdx.x;
    if (i < n) {
      const float4 mo = a.m_ray[2 * i], md = a.m_ray[2 * i + 1];
      const uint32_t p = __float_as_uint(mo.w);
      const TraceRay r = make_ray(V3{mo.x, mo.y, mo.z}, V3{md.x, md.y, md.z});
      float best = __builtin_inff();
      TraceCounters c;
      auto st = make_stack<KIND>(lds, kPtBlock, sa);
      const int32_t hit = trace_bvh_s<false, false, false, WW>(a.scene, r, best, st, c);
      ++traced;
      if (hit >= 0) {
        const int32_t light = a.scene.prim_light[hit];
        if (light >= 0) {
          const float *L = a.scene.lights + size_t(light) * kLightStride;
          a.m_add[3 * p] = a.m_fac[3 * p] * L[1];
          a.m_add[3 * p + 1] = a.m_fac[3 * p + 1] * L[2];
          a.m_add[3 * p + 2] = a.m_fac[3 * p + 2] * L[3];
        }
      }
    }
  }
  wave_count(traced, &a.ray_ctr[kCtrSecondary]);
}

// ---------------------------------------------------------------- gather
__global__ void __launch_bounds__(kPtBlock) k_gather(PtArgs a, int cur) {
  const uint32_t n = a.counts[cur == 0 ? kQCur : kQNext];
  const float4 *q = a.q_ray[cur];
  const int K = a.shadow_per_vertex;
  for (uint32_t i = blockIdx.x * kPtBlock + threadIdx.x; i < n; i += gridDim.x * kPtBlock) {
    if (a.hit_prim[i] < 0) continue;
    const uint32_t p = __float_as_uint(q[2 * i].w);
    V3 direct{0.0f, 0.0f, 0.0f};
    for (int k = 0; k < K; ++k) {
      const size_t slot = size_t(p) * K + k;
      if (a.s_vis[slot]) direct = direct + V3{a.s_contrib[3 * slot], a.s_contrib[3 * slot + 1], a.s_contrib[3 * slot + 2]};
    }
    direct = direct + V3{a.m_add[3 * p], a.m_add[3 * p + 1], a.m_add[3 * p + 2]};
    a.rad[3 * p] += direct.x;
    a.rad[3 * p + 1] += direct.y;
    a.rad[3 * p + 2] += direct.z;
  }
}

// ---------------------------------------------------------------- counters / film
__global__ void k_swap_counts(uint32_t *counts, int cur) {
  // the queue just consumed becomes empty; reset the MIS queue
  if (threadIdx.x == 0) {
    counts[cur == 0 ? kQCur : kQNext] = 0;
    counts[kQMis] = 0;
  }
}

__global__ void __launch_bounds__(kPtBlock) k_resolve(PtArgs a, PtGen g, float *__restrict__ accum) {
  for (uint32_t j = blockIdx.x * kPtBlock + threadIdx.x; j < g.npix; j += gridDim.x * kPtBlock) {
    const uint32_t pixel = g.p0 + j;
    float *px = accum + 3 * size_t(pixel);
    float r = px[0], gg = px[1], b = px[2];
    for (uint32_t s = 0; s < g.nsamp; ++s) {
      const uint32_t p = s * g.npix + j;
      r = r + a.rad[3 * p];
      gg = gg + a.rad[3 * p + 1];
      b = b + a.rad[3 * p + 2];
    }
    px[0] = r;
    px[1] = gg;
    px[2] = b;
  }
}

// ---------------------------------------------------------------- host driver
namespace {

int ShadowPerVertex(const muonx_scene *scene, int strata, int per) {
  int K = 0;
  const std::vector<float> &L = scene->host_lights;
  for (uint32_t l = 0; l < scene->view.num_lights; ++l) {
    if (int(L[size_t(l) * kLightStride]) == MUONX_LIGHT_QUAD) K += strata * strata * per;
    else K += 1;
  }
  return K;
}

struct BounceLaunch {
  const PtArgs &a;
  const TraverseCfg &cfg;
  int cur, depth;
  bool any_nee, mis;
  hipStream_t st;
  template <int KIND, bool WW>
  void run() const {
    const StackArgs sa{cfg.spill, cfg.gcap};
    const dim3 grid(cfg.grid), block(kPtBlock);
    hipLaunchKernelGGL((k_extend<KIND, WW>), grid, block, 0, st, a, sa, cur, depth);
    hipLaunchKernelGGL(k_shade, grid, block, 0, st, a, cur, depth);
    if (any_nee) hipLaunchKernelGGL((k_shadow<KIND, WW>), grid, block, 0, st, a, sa, cur);
    if (mis) hipLaunchKernelGGL((k_mis<KIND, WW>), grid, block, 0, st, a, sa);
    if (any_nee) hipLaunchKernelGGL(k_gather, grid, block, 0, st, a, cur);
    hipLaunchKernelGGL(k_swap_counts, dim3(1), dim3(64), 0, st, a.counts, cur);
  }
};

}  // namespace

int RenderPathTracer(muonx_scene *scene, const muonx_render_params &p, float *accum, RenderWorkspace *ws,
                     hipStream_t stream) {
  const DevSceneView &s = scene->view;
  PtArgs a{};
  a.scene = s;
  a.cam = MakeCam(p);
  a.seed = p.seed;
  a.job = 0;
  a.min_depth = p.min_depth;
  a.cutoff = p.max_depth == -1 ? INT_MAX : (p.nee != MUONX_NEE_OFF ? p.max_depth - 1 : p.max_depth);
  a.nee = p.nee;
  a.is_mode = p.importance_sampling;
  a.rr = p.russian_roulette;
  a.light_samples = std::max(1, p.light_samples);
  // glm::sqrt(int) is a double square root (muon/integration.cc:393)
  a.strata = p.light_stratify ? int(std::sqrt(double(p.light_samples))) : 1;
  a.per_section = p.light_stratify ? 1 : p.light_samples;
  if (a.strata < 0) a.strata = 0;
  if (a.per_section < 0) a.per_section = 0;
  const bool any_nee = p.nee != MUONX_NEE_OFF && s.num_lights > 0;
  a.shadow_per_vertex = any_nee ? std::max(1, ShadowPerVertex(scene, a.strata, a.per_section)) : 1;
  const bool mis = p.nee == MUONX_NEE_MIS;

  const uint32_t npix = uint32_t(p.width) * uint32_t(p.height);
  const uint64_t total = uint64_t(npix) * uint64_t(p.spp);
  // Path capacity: bounded by a memory budget and by the job size.
  const uint64_t per_path = 4 * (3 + 3 + 1) + 2 * 2 * 16 + 4 + 3 * 4 * 2 + 2 * 16 +
                            uint64_t(a.shadow_per_vertex) * (3 * 4 + 2 * 16 + 2);
  const uint64_t budget = uint64_t(std::max(16, scene->ctx->tuning.path_budget_mb)) << 20;
  uint64_t cap = std::max<uint64_t>(1, budget / per_path);
  cap = std::min<uint64_t>(cap, total);
  cap = std::min<uint64_t>(cap, uint64_t(1) << 26);
  // Generations cover whole sample ranges of a pixel range.
  uint32_t gen_pix = uint32_t(std::min<uint64_t>(npix, cap));
  uint32_t gen_samp = uint32_t(std::max<uint64_t>(1, cap / gen_pix));
  gen_samp = std::min<uint32_t>(gen_samp, uint32_t(p.spp));
  const uint64_t P = uint64_t(gen_pix) * gen_samp;

  // workspace layout
  const size_t K = size_t(a.shadow_per_vertex);
  const size_t bytes_paths = P * (3 + 3 + 3 + 3) * 4 + P * 4;          // rad thr m_fac m_add ctr
  const size_t bytes_queues = 2 * P * 32 + P * 4 + P * 32;             // 2 ray queues, hit, mis queue
  const size_t bytes_shadow = P * K * (12 + 16 + 16 + 1 + 1);
  MX_HIP(ws->paths.Reserve(bytes_paths + 256));
  MX_HIP(ws->queues.Reserve(bytes_queues + 256));
  MX_HIP(ws->misc.Reserve(bytes_shadow + 1024));
  char *pp = ws->paths.as<char>();
  a.rad = reinterpret_cast<float *>(pp);
  a.thr = a.rad + 3 * P;
  a.m_fac = a.thr + 3 * P;
  a.m_add = a.m_fac + 3 * P;
  a.ctr = reinterpret_cast<uint32_t *>(a.m_add + 3 * P);
  char *qp = ws->queues.as<char>();
  a.q_ray[0] = reinterpret_cast<float4 *>(qp);
  a.q_ray[1] = a.q_ray[0] + 2 * P;
  a.m_ray = a.q_ray[1] + 2 * P;
  a.hit_prim = reinterpret_cast<int32_t *>(a.m_ray + 2 * P);
  char *sp = ws->misc.as<char>();
  a.s_org = reinterpret_cast<float4 *>(sp);
  a.s_dir = a.s_org + P * K;
  a.s_contrib = reinterpret_cast<float *>(a.s_dir + P * K);
  a.s_vis = reinterpret_cast<uint8_t *>(a.s_contrib + 3 * P * K);
  a.s_valid = a.s_vis + P * K;
  unsigned long long *ctr_base = ws->counters.as<unsigned long long>();
  a.ray_ctr = ctr_base;
  a.counts = reinterpret_cast<uint32_t *>(ctr_base + 4);  // [kQCur, kQNext, kQMis, spare]

  MX_HIP(hipMemsetAsync(accum, 0, size_t(npix) * 3 * sizeof(float), stream));
  if (!any_nee) MX_HIP(hipMemsetAsync(a.s_valid, 0, P * K, stream));

  TraverseCfg cfg;
  int rc = PlanTraverse(scene->ctx, s.bvh_depth, P * std::max<uint64_t>(1, K), kPtBlock, false, cfg);
  if (rc != MUONX_OK) return rc;
  if ((rc = BeginSpill(scene->ctx, cfg, stream)) != MUONX_OK) return rc;
  const int grid = cfg.grid;
  uint32_t *h_count = nullptr;
  if (p.max_depth == -1) MX_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_count), sizeof(uint32_t)));
  for (uint32_t p0 = 0; p0 < npix && rc == MUONX_OK; p0 += gen_pix) {
    const uint32_t np = std::min(gen_pix, npix - p0);
    for (uint32_t s0 = 0; s0 < uint32_t(p.spp); s0 += gen_samp) {
      PtGen g;
      g.p0 = p0;
      g.npix = np;
      g.s0 = s0;
      g.nsamp = std::min<uint32_t>(gen_samp, uint32_t(p.spp) - s0);
      g.npaths = g.npix * g.nsamp;
      a.gen = g;
      MX_HIP(hipMemsetAsync(a.counts, 0, 4 * sizeof(uint32_t), stream));
      hipLaunchKernelGGL(k_raygen, dim3(grid), dim3(kPtBlock), 0, stream, a, g);
      // muon/integration.cc:29-33 at depth 0
      if (a.cutoff >= 0) {
        int cur = 0;
        const int max_bounces = p.max_depth == -1 ? 4096 : a.cutoff + 1;
        for (int depth = 0; depth < max_bounces; ++depth) {
          DispatchTraverse(cfg, BounceLaunch{a, cfg, cur, depth, any_nee, mis, stream});
          MX_HIP(hipGetLastError());
          cur ^= 1;
          if (p.max_depth == -1 && (depth % 4 == 3)) {
            MX_HIP(hipMemcpyAsync(h_count, &a.counts[cur == 0 ? kQCur : kQNext], sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, stream));
            MX_HIP(hipStreamSynchronize(stream));
            if (*h_count == 0) break;
          }
        }
      }
      hipLaunchKernelGGL(k_resolve, dim3(grid), dim3(kPtBlock), 0, stream, a, g, accum);
      MX_HIP(hipGetLastError());
    }
  }
  if (rc == MUONX_OK) rc = EndSpill(scene->ctx, cfg, stream);
  if (h_count) (void)hipHostFree(h_count);
  return rc;
}

}  // namespace muonx
