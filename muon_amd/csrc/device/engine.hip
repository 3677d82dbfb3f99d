// Persistent wavefront engine (see engine.h): kernels and their host-side control.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "engine.h"
#include "engine_plan.h"
#include "hitrec.h"
#include "wave.h"
#include "shading.h"

namespace muonx {

__device__ __forceinline__ unsigned long long to_fixed(float v) {
  if (!(v > 0.0f)) return 0ull;  // zero, negative or NaN
  const double x = double(v) * kFilmScale + 0.5;
  return x < 4.0e18 ? (unsigned long long)x : 4000000000000000000ull;
}

// Adds a wavefront's count to this block's slot of counter `id`.
__device__ __forceinline__ void ray_count(EngCounters *c, int id, uint32_t n) {
  unsigned long long v = n;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (__lane_id() == 0 && v) atomicAdd(&c->ray_slots[id][blockIdx.x % kRaySlots].v, v);
}

// Adds a wavefront's traversal work (counted per lane) to the counters.
__device__ __forceinline__ void work_count(EngCounters *c, const TraceCounters &t) {
  ray_count(c, kCtrObjTests, t.prims);
  ray_count(c, kCtrObjHits, t.ohits);
  ray_count(c, kCtrBoxTests, t.btests);
  ray_count(c, kCtrBoxHits, t.bhits);
}

// ---------------------------------------------------------------- extend
template <int KIND, bool WW, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(kEngBlock) k_eng_extend(EngArgs a, StackArgs sa, int cur) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kEngBlock];
  const uint32_t n = a.c->qcount[cur].v;
  const float4 *q = a.q_ray[cur];
  const PathState &ps = a.st[cur];
  uint32_t prim_rays = 0, sec_rays = 0;
  TraceCounters c;
  for (uint32_t base = blockIdx.x * kEngBlock; base < n; base += gridDim.x * kEngBlock) {
    const uint32_t i = base + threadIdx.x;
    if (i < n) {
      const float4 qo = q[2 * i], qd = q[2 * i + 1];
      const uint32_t jw = ps.job[i], js = PathJobSlot(jw);
      const int depth = ps.depth[i];
      int32_t hit = -1;
      // Integrator::Trace returns 0 past the depth cutoff (muon/integration.cc:29-33)
      if (depth <= a.jobs[js].cutoff && depth < kMaxBounces) {
        // a register copy: the traversal's global spill stores must not force
        // re-reads of the scene's pointers
        const DevSceneView sv = a.scenes[PathSceneSlot(jw)];
        const TraceRay r = make_ray(V3{qo.x, qo.y, qo.z}, V3{qd.x, qd.y, qd.z});
        float best = __builtin_inff();
        auto st = make_stack<KIND>(lds, kEngBlock, sa);
        hit = trace_fast<WIDTH, false, COUNT, WW>(sv, r, best, st, c);
        if (depth == 0) ++prim_rays;
        else ++sec_rays;
      }
      a.hit_prim[i] = hit;
    }
  }
  ray_count(a.c, 0, prim_rays);
  ray_count(a.c, 1, sec_rays);
  if (COUNT) work_count(a.c, c);
}

// ---------------------------------------------------------------- shade
// The job word and hit primitive of a thread's next path in a grid-stride
// loop, loaded one pass ahead: the chain of dependent fetches of the next
// path (job record and scene view, then primitive, transform, material)
// starts from registers instead of waiting for these two loads first.
struct PathAhead {
  uint32_t jw = 0;
  int32_t prim = -1;
  __device__ __forceinline__ PathAhead(const PathState &ps, const int32_t *hit_prim, uint32_t i, uint32_t n) {
    next(ps, hit_prim, i, n);
  }
  __device__ __forceinline__ void next(const PathState &ps, const int32_t *hit_prim, uint32_t i, uint32_t n) {
    if (i < n) {
      jw = ps.job[i];
      prim = hit_prim[i];
    }
  }
};

// PathTracer::Shade (muon/integration.cc:205-271) in two kernels, split where
// the random draws of the light samples end, so each holds fewer registers:
//   shade_light  hit record, emission, NEE shadow rays (ShadeDirectNEE)
//   shade_bsdf   MIS BRDF ray (ShadeDirectImportanceSampling) and the
//                continuation with Russian roulette (ShadeIndirect)
__global__ void __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_eng_shade_light(EngArgs a, int cur) {
  __shared__ uint32_t wsum[kShadeWaves + 1];
  const uint32_t n = a.c->qcount[cur].v;
  const float4 *q = a.q_ray[cur];
  const PathState &ps = a.st[cur];
  const size_t P = a.capacity;
  const uint32_t stride = gridDim.x * kShadeBlock;
  PathAhead ahead(ps, a.hit_prim, blockIdx.x * kShadeBlock + threadIdx.x, n);
  for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t jw = ahead.jw;
    const int32_t prim = ahead.prim;
    ahead.next(ps, a.hit_prim, i + stride, n);
    uint32_t nshadow = 0;
    uint64_t vmask = 0;  // valid shadow rays k < 64
    int K = 0;
    if (i < n) {
      const JobDev job = a.jobs[PathJobSlot(jw)];
      const int depth = ps.depth[i];
      K = job.K;
      for (int k = 0; k < K; ++k) {
        a.s_vis[k * P + i] = 0;
        a.s_valid[k * P + i] = 0;
      }
      a.m_hit[i] = -1;
      if (prim >= 0) {
        const DevSceneView s = a.scenes[PathSceneSlot(jw)];
        const float4 qo = q[2 * i], qd = q[2 * i + 1];
        const V3 ro{qo.x, qo.y, qo.z}, rd{qd.x, qd.y, qd.z};
        const HitRecord h = hit_record(s, uint32_t(prim), ro, rd);
        const int32_t mi = ldg(s.prim_mat + prim);
        const V3 shift = h.pos + kEps * h.normal;
        // the hit for shade_bsdf: shadow/MIS origin + material, normal
        a.shift[i] = make_float4(shift.x, shift.y, shift.z, __int_as_float(mi));
        a.hnorm[i] = make_float4(h.normal.x, h.normal.y, h.normal.z, 0.0f);
        if (depth >= job.min_depth) {
          const Mat m = load_mat(s.materials + size_t(mi) * kMatStride);
          const V3 T{ps.thr[3 * i], ps.thr[3 * i + 1], ps.thr[3 * i + 2]};
          V3 rad{ps.rad[3 * i], ps.rad[3 * i + 1], ps.rad[3 * i + 2]};
          // emission (muon/integration.cc:257-264)
          if (!((job.nee != MUONX_NEE_OFF && depth > 0) || dot(h.normal, -rd) < 0.0f)) rad = rad + T * m.emission;
          ps.rad[3 * i] = rad.x;
          ps.rad[3 * i + 1] = rad.y;
          ps.rad[3 * i + 2] = rad.z;
          if (job.nee != MUONX_NEE_OFF) {
            // ShadeDirectNEE (muon/integration.cc:340-466): one shadow ray per light sample
            const PathKey key{ps.key_a[i], ps.key_b[i]};
            uint32_t ctr = ps.ctr[i];
            const bool mis = job.nee == MUONX_NEE_MIS;
            int k = 0;
            for (uint32_t l = 0; l < s.num_lights; ++l) {
              const float *L = s.lights + size_t(l) * kLightStride;
              const int kind = int(ldg(L));
              if (kind != MUONX_LIGHT_QUAD) {
                V3 ldir, col{ldg(L + 1), ldg(L + 2), ldg(L + 3)};
                float ldist;
                if (kind == MUONX_LIGHT_POINT) {
                  const V3 to_light = V3{ldg(L + 4), ldg(L + 5), ldg(L + 6)} - h.pos;
                  ldir = est_normalize(to_light);
                  ldist = est_length(to_light);
                  const float att = ldg(L + 7) + ldg(L + 8) * ldist + ldg(L + 9) * ldist * ldist;
                  col = est_div(col, att);
                } else {
                  ldir = V3{ldg(L + 4), ldg(L + 5), ldg(L + 6)};
                  ldist = __builtin_inff();
                }
                V3 irr = col;
                if (ldist < __builtin_inff()) irr = est_div(irr, ldist * ldist);
                const float cosi = gmax(dot(h.normal, ldir), 0.0f);
                const V3 contrib = T * (irr * cosi * brdf_eval(m, ldir, rd, h.normal));
                const size_t sl = k * P + i;
                a.s_contrib[3 * sl] = contrib.x;
                a.s_contrib[3 * sl + 1] = contrib.y;
                a.s_contrib[3 * sl + 2] = contrib.z;
                a.s_dir[sl] = make_float4(ldir.x, ldir.y, ldir.z, ldist);
                // a zero contribution cannot change the image: its ray is skipped
                const uint8_t valid = (contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f) ? 1 : 0;
                a.s_valid[sl] = valid;
                nshadow += valid;
                if (valid && k < 64) vmask |= 1ull << k;
                ++k;
                continue;
              }
              const QuadLight ql = load_quad(L);
              const V3 rn = est_normalize(cross(ql.e0, ql.e1));
              const float scale = est_div(ql.area, float(job.light_samples));
              const float cell = est_rcp(float(job.strata));
              for (int si = 0; si < job.strata; ++si) {
                for (int sj = 0; sj < job.strata; ++sj) {
                  for (int sk = 0; sk < job.per_section; ++sk) {
                    const float u = rng_uniform(key, ctr);
                    const float v = rng_uniform(key, ctr);
                    V3 lp = (ql.corner + ((float(si) + u) * cell) * ql.e0) + ((float(sj) + v) * cell) * ql.e1;
                    lp = lp - kEps * rn;
                    const V3 to_light = lp - shift;
                    const V3 ldir = est_normalize(to_light);
                    const float ldist = est_length(to_light);
                    const float cosi = gmax(dot(h.normal, ldir), 0.0f);
                    const float cosl = gmax(dot(rn, ldir), 0.0f);
                    const float g = est_div(cosi * cosl, ldist * ldist);
                    V3 c = g * brdf_eval(m, ldir, rd, h.normal);
                    if (mis) {
                      const float lpdf = nee_pdf(s, shift, ldir);
                      const float bpdf = is_pdf(job.is_mode, m, ldir, rd, h.normal);
                      c = c * power_heuristic(lpdf, bpdf);
                    }
                    const V3 contrib = T * (ql.color * (c * scale));
                    const size_t sl = k * P + i;
                    a.s_contrib[3 * sl] = contrib.x;
                    a.s_contrib[3 * sl + 1] = contrib.y;
                    a.s_contrib[3 * sl + 2] = contrib.z;
                    a.s_dir[sl] = make_float4(ldir.x, ldir.y, ldir.z, ldist);
                    const uint8_t valid = (contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f) ? 1 : 0;
                    a.s_valid[sl] = valid;
                    nshadow += valid;
                    if (valid && k < 64) vmask |= 1ull << k;
                    ++k;
                  }
                }
              }
            }
            ps.ctr[i] = ctr;
          }
        }
      }
    }
    // the shadow rays to trace, compacted (zero contributions are skipped)
    uint32_t si = block_append_n<kShadeWaves>(nshadow, &a.c->shadow_count.v, wsum);
    for (uint64_t m = vmask; m; m &= m - 1)
      a.s_queue[si++] = ShadowSlot(uint32_t(__ffsll((unsigned long long)m) - 1), a.capacity, i);
    for (int k = 64; nshadow && k < K; ++k) {
      const uint32_t sl = ShadowSlot(uint32_t(k), a.capacity, i);
      if (a.s_valid[sl]) a.s_queue[si++] = sl;
    }
  }
}

__global__ void __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_eng_shade_bsdf(EngArgs a, int cur) {
  __shared__ uint32_t wsum[kShadeWaves + 1];
  uint32_t culled_n = 0;
  const uint32_t n = a.c->qcount[cur].v;
  const float4 *q = a.q_ray[cur];
  float4 *qn = a.q_ray[cur ^ 1];
  const PathState &ps = a.st[cur];
  const PathState &pn = a.st[cur ^ 1];
  const uint32_t stride = gridDim.x * kShadeBlock;
  PathAhead ahead(ps, a.hit_prim, blockIdx.x * kShadeBlock + threadIdx.x, n);
  for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t jw = ahead.jw;
    const int32_t prim = ahead.prim;
    ahead.next(ps, a.hit_prim, i + stride, n);
    bool cont = false, want_mis = false, culled = false;
    uint32_t js = 0, pixel = 0, ctr = 0;
    PathKey key{0u, 0u};
    int depth = 0;
    float mis_tmax = __builtin_inff();
    V3 shift{0, 0, 0}, next_dir{0, 0, 0}, mis_dir{0, 0, 0}, T{0, 0, 0};
    if (i < n && prim >= 0) {
      js = jw;  // the job word: a continuing path carries it on
      const JobDev job = a.jobs[PathJobSlot(js)];
      const DevSceneView s = a.scenes[PathSceneSlot(js)];
      depth = ps.depth[i];
      const float4 sh = a.shift[i], hn = a.hnorm[i];
      shift = V3{sh.x, sh.y, sh.z};
      const V3 normal{hn.x, hn.y, hn.z};
      const Mat m = load_mat(s.materials + size_t(__float_as_int(sh.w)) * kMatStride);
      const float4 qd = q[2 * i + 1];
      const V3 rd{qd.x, qd.y, qd.z};
      pixel = ps.pixel[i];
      key = PathKey{ps.key_a[i], ps.key_b[i]};
      ctr = ps.ctr[i];
      T = V3{ps.thr[3 * i], ps.thr[3 * i + 1], ps.thr[3 * i + 2]};
      if (depth >= job.min_depth && job.nee == MUONX_NEE_MIS) {
        // ShadeDirectImportanceSampling (muon/integration.cc:468-501)
        V3 dir, nT;
        float bpdf;
        if (sample_reflection(job.is_mode, m, rd, normal, T, key, ctr, dir, bpdf, nT)) {
          const float lpdf = nee_pdf(s, shift, dir);
          const V3 f = nT * power_heuristic(bpdf, lpdf);
          a.m_fac[3 * i] = f.x;
          a.m_fac[3 * i + 1] = f.y;
          a.m_fac[3 * i + 2] = f.z;
          want_mis = true;
          mis_dir = dir;
          if (s.cull_mis) {
            // The term is nonzero only when the ray's closest hit is an
            // emitting primitive. Test the few emitters with the traversal's
            // own primitive test: a ray that hits none is settled here (no
            // light hit, m_hit stays -1); otherwise the traversal only needs
            // to look up to the nearest emitter hit.
            const TraceRay mr = make_ray(shift, dir);
            ObjCache oc;
            oc.tid = -1;
            float dmin = __builtin_inff(), blo;
            const float *lb = s.light_box;
            if (box_hit(mr, lb[0], lb[3], lb[1], lb[4], lb[2], lb[5], __builtin_inff(), blo)) {
              for (uint32_t k = 0; k < s.num_light_prims; ++k) {
                float dist;
                uint32_t flags;
                if (prim_hit(s, ldg(s.light_prims + k), mr, oc, dist, flags) && dist > 0.0f && dist < dmin) dmin = dist;
              }
            }
            if (dmin == __builtin_inff()) {
              want_mis = false;
              culled = true;
            } else {
              mis_tmax = dmin * kWidenHi;
            }
          }
        }
      }
      // ShadeIndirect (muon/integration.cc:294-328)
      V3 dir, nT;
      float pdf;
      if (sample_reflection(job.is_mode, m, rd, normal, T, key, ctr, dir, pdf, nT)) {
        bool alive = true;
        if (job.rr) {
          const float qc = gmin(gmax(gmax(nT.x, nT.y), nT.z), 1.0f);
          if (qc < rng_uniform(key, ctr)) alive = false;
          else nT = est_div(nT, qc);
        }
        // Trace() returns 0 past the depth cutoff (muon/integration.cc:29-33)
        if (alive && depth + 1 <= job.cutoff && depth + 1 < kMaxBounces) {
          cont = true;
          next_dir = dir;
          T = nT;
        }
      }
    }
    // a continuing path moves to the next queue with its state; its radiance
    // follows in retire, after the NEE / MIS terms are added
    const uint32_t ni = block_append<kShadeWaves>(cont, &a.c->qcount[cur ^ 1].v, wsum);
    if (i < n) a.nidx[i] = cont ? ni : 0xffffffffu;
    if (cont) {
      qn[2 * ni] = make_float4(shift.x, shift.y, shift.z, 0.0f);
      qn[2 * ni + 1] = make_float4(next_dir.x, next_dir.y, next_dir.z, 0.0f);
      pn.job[ni] = js;
      pn.pixel[ni] = pixel;
      pn.ctr[ni] = ctr;
      pn.key_a[ni] = key.a;
      pn.key_b[ni] = key.b;
      pn.depth[ni] = depth + 1;
      pn.thr[3 * ni] = T.x;
      pn.thr[3 * ni + 1] = T.y;
      pn.thr[3 * ni + 2] = T.z;
    }
    const uint32_t mi = block_append<kShadeWaves>(want_mis, &a.c->mis_count.v, wsum);
    if (want_mis) {
      a.m_ray[2 * mi] = make_float4(shift.x, shift.y, shift.z, __uint_as_float(i));
      a.m_ray[2 * mi + 1] = make_float4(mis_dir.x, mis_dir.y, mis_dir.z, mis_tmax);
    }
    culled_n += culled ? 1u : 0u;
  }
  // settled MIS rays count as traced secondary rays (the reference traces them)
  ray_count(a.c, 1, culled_n);
  ray_count(a.c, 3, culled_n);
}

// PathTracer::Shade without next-event estimation (muon/integration.cc:205-271
// with NEE off: emission, then ShadeIndirect with Russian roulette) in one
// kernel, used for an iteration in which no admitted job has NEE on. The
// split kernels above do the same arithmetic and draw the same random numbers
// for such a path (shade_light draws none without NEE, shade_bsdf skips its
// MIS sample), so a path's result does not depend on which pair ran; this
// one skips the hit-record hand-off through memory and the shadow and MIS
// traversals, which would have no rays.
__global__ void __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_eng_shade_plain(EngArgs a, int cur) {
  __shared__ uint32_t wsum[kShadeWaves + 1];
  const uint32_t n = a.c->qcount[cur].v;
  const float4 *q = a.q_ray[cur];
  float4 *qn = a.q_ray[cur ^ 1];
  const PathState &ps = a.st[cur];
  const PathState &pn = a.st[cur ^ 1];
  const uint32_t stride = gridDim.x * kShadeBlock;
  PathAhead ahead(ps, a.hit_prim, blockIdx.x * kShadeBlock + threadIdx.x, n);
  for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t jw = ahead.jw;
    const int32_t prim = ahead.prim;
    ahead.next(ps, a.hit_prim, i + stride, n);
    bool cont = false;
    uint32_t js = 0, pixel = 0, ctr = 0;
    PathKey key{0u, 0u};
    int depth = 0;
    V3 shift{0, 0, 0}, next_dir{0, 0, 0}, T{0, 0, 0};
    if (i < n) {
      a.m_hit[i] = -1;  // retire: no MIS term
      if (prim >= 0) {
        js = jw;  // the job word: a continuing path carries it on
        const JobDev job = a.jobs[PathJobSlot(js)];
        const DevSceneView s = a.scenes[PathSceneSlot(js)];
        depth = ps.depth[i];
        const float4 qo = q[2 * i], qd = q[2 * i + 1];
        const V3 ro{qo.x, qo.y, qo.z}, rd{qd.x, qd.y, qd.z};
        const HitRecord h = hit_record(s, uint32_t(prim), ro, rd);
        const int32_t mi = ldg(s.prim_mat + prim);
        shift = h.pos + kEps * h.normal;
        const Mat m = load_mat(s.materials + size_t(mi) * kMatStride);
        T = V3{ps.thr[3 * i], ps.thr[3 * i + 1], ps.thr[3 * i + 2]};
        if (depth >= job.min_depth && !(dot(h.normal, -rd) < 0.0f)) {
          // emission (muon/integration.cc:257-264)
          const V3 rad = V3{ps.rad[3 * i], ps.rad[3 * i + 1], ps.rad[3 * i + 2]} + T * m.emission;
          ps.rad[3 * i] = rad.x;
          ps.rad[3 * i + 1] = rad.y;
          ps.rad[3 * i + 2] = rad.z;
        }
        pixel = ps.pixel[i];
        key = PathKey{ps.key_a[i], ps.key_b[i]};
        ctr = ps.ctr[i];
        // ShadeIndirect (muon/integration.cc:294-328)
        V3 dir, nT;
        float pdf;
        if (sample_reflection(job.is_mode, m, rd, h.normal, T, key, ctr, dir, pdf, nT)) {
          bool alive = true;
          if (job.rr) {
            const float qc = gmin(gmax(gmax(nT.x, nT.y), nT.z), 1.0f);
            if (qc < rng_uniform(key, ctr)) alive = false;
            else nT = est_div(nT, qc);
          }
          // Trace() returns 0 past the depth cutoff (muon/integration.cc:29-33)
          if (alive && depth + 1 <= job.cutoff && depth + 1 < kMaxBounces) {
            cont = true;
            next_dir = dir;
            T = nT;
          }
        }
      }
    }
    // a continuing path moves to the next queue with its state; its radiance
    // follows in retire
    const uint32_t ni = block_append<kShadeWaves>(cont, &a.c->qcount[cur ^ 1].v, wsum);
    if (i < n) a.nidx[i] = cont ? ni : 0xffffffffu;
    if (cont) {
      qn[2 * ni] = make_float4(shift.x, shift.y, shift.z, 0.0f);
      qn[2 * ni + 1] = make_float4(next_dir.x, next_dir.y, next_dir.z, 0.0f);
      pn.job[ni] = js;
      pn.pixel[ni] = pixel;
      pn.ctr[ni] = ctr;
      pn.key_a[ni] = key.a;
      pn.key_b[ni] = key.b;
      pn.depth[ni] = depth + 1;
      pn.thr[3 * ni] = T.x;
      pn.thr[3 * ni + 1] = T.y;
      pn.thr[3 * ni + 2] = T.z;
    }
  }
}

// ---------------------------------------------------------------- shadow (any hit)
template <int KIND, bool WW, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(kEngBlock) k_eng_shadow(EngArgs a, StackArgs sa, int cur) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kEngBlock];
  const uint32_t P = a.capacity;
  const uint32_t n = a.c->shadow_count.v;
  const PathState &ps = a.st[cur];
  uint32_t traced = 0;
  TraceCounters c;
  for (uint32_t base = blockIdx.x * kEngBlock; base < n; base += gridDim.x * kEngBlock) {
    const uint32_t j = base + threadIdx.x;
    if (j < n) {
      const uint32_t sl = a.s_queue[j];
      const uint32_t i = sl % P;
      const DevSceneView sv = a.scenes[PathSceneSlot(ps.job[i])];
      const float4 so = a.shift[i], sd = a.s_dir[sl];
      const TraceRay r = make_ray(V3{so.x, so.y, so.z}, V3{sd.x, sd.y, sd.z});
      float best = sd.w;
      auto st = make_stack<KIND>(lds, kEngBlock, sa);
      const int32_t hit = trace_fast<WIDTH, true, COUNT, WW>(sv, r, best, st, c);
      a.s_vis[sl] = hit < 0 ? 1 : 0;
      ++traced;
    }
  }
  ray_count(a.c, 2, traced);
  if (COUNT) work_count(a.c, c);
}

// ---------------------------------------------------------------- MIS extend
template <int KIND, bool WW, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(kEngBlock) k_eng_mis(EngArgs a, StackArgs sa, int cur) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kEngBlock];
  const uint32_t n = a.c->mis_count.v;
  const PathState &ps = a.st[cur];
  uint32_t traced = 0;
  TraceCounters c;
  for (uint32_t base = blockIdx.x * kEngBlock; base < n; base += gridDim.x * kEngBlock) {
    const uint32_t j = base + threadIdx.x;
    if (j < n) {
      const float4 mo = a.m_ray[2 * j], md = a.m_ray[2 * j + 1];
      const uint32_t i = __float_as_uint(mo.w);
      const DevSceneView s = a.scenes[PathSceneSlot(ps.job[i])];
      const TraceRay r = make_ray(V3{mo.x, mo.y, mo.z}, V3{md.x, md.y, md.z});
      float best = md.w;  // the nearest emitter hit, or +inf
      auto st = make_stack<KIND>(lds, kEngBlock, sa);
      a.m_hit[i] = trace_fast<WIDTH, false, COUNT, WW>(s, r, best, st, c);
      ++traced;
    }
  }
  ray_count(a.c, 1, traced);
  if (COUNT) work_count(a.c, c);
}

// ---------------------------------------------------------------- per-lane fetch variants
// The three traversal kernels with per-lane work claiming (WaveFetch): a
// lane that finishes its ray takes the next one once `fetch_min` lanes of its
// wavefront are idle. Results are those of the kernels above.
template <int KIND, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(kEngBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) k_eng_extend_dyn(EngArgs a, StackArgs sa, int cur, int fetch_min) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kEngBlock];
  const uint32_t n = a.c->qcount[cur].v;
  const float4 *q = a.q_ray[cur];
  const PathState &ps = a.st[cur];
  auto st = make_stack<KIND>(lds, kEngBlock, sa);
  LaneTrace<WIDTH, false, COUNT> lt;
  DevSceneView sv;
  TraceCounters c;
  WaveFetch wf(n);
  uint32_t i = 0, prim_rays = 0, sec_rays = 0;
  bool have = false, drained = false;
  while (true) {
    const bool need = !have && !drained;
    const uint64_t needm = __ballot(need);
    if (needm && (__popcll(needm) >= fetch_min || __ballot(have) == 0)) {
      const uint32_t k = wf.claim(needm, &a.c->claim[0].v);
      if (need) {
        if (k < n) {
          i = k;
          const uint32_t jw = ps.job[i], js = PathJobSlot(jw);
          const int depth = ps.depth[i];
          // the ray and scene view are fetched with the job's cutoff, not
          // after it: every queued path has both, and starting the lane's
          // walk only sets registers, so it happens whether or not the
          // cutoff lets the ray be traced
          const float4 qo = q[2 * i], qd = q[2 * i + 1];
          sv = a.scenes[PathSceneSlot(jw)];
          lt.start(make_ray(V3{qo.x, qo.y, qo.z}, V3{qd.x, qd.y, qd.z}), __builtin_inff(), st);
          // Integrator::Trace returns 0 past the depth cutoff (muon/integration.cc:29-33)
          if (depth <= a.jobs[js].cutoff && depth < kMaxBounces) {
            have = true;
            if (depth == 0) ++prim_rays;
            else ++sec_rays;
          } else {
            a.hit_prim[i] = -1;
          }
        } else {
          drained = true;
        }
      }
    }
    if (__ballot(have) == 0) break;
    if (have && lt.round(sv, st, c)) {
      a.hit_prim[i] = lt.hit;
      have = false;
    }
  }
  ray_count(a.c, 0, prim_rays);
  ray_count(a.c, 1, sec_rays);
  if (COUNT) work_count(a.c, c);
}

template <int KIND, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(kEngBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) k_eng_shadow_dyn(EngArgs a, StackArgs sa, int cur, int fetch_min) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kEngBlock];
  const uint32_t P = a.capacity;
  const uint32_t n = a.c->shadow_count.v;
  const PathState &ps = a.st[cur];
  auto st = make_stack<KIND>(lds, kEngBlock, sa);
  LaneTrace<WIDTH, true, COUNT> lt;
  DevSceneView sv;
  TraceCounters c;
  WaveFetch wf(n);
  uint32_t sl = 0, traced = 0;
  bool have = false, drained = false;
  while (true) {
    const bool need = !have && !drained;
    const uint64_t needm = __ballot(need);
    if (needm && (__popcll(needm) >= fetch_min || __ballot(have) == 0)) {
      const uint32_t k = wf.claim(needm, &a.c->claim[1].v);
      if (need) {
        if (k < n) {
          sl = a.s_queue[k];
          const uint32_t i = sl % P;
          sv = a.scenes[PathSceneSlot(ps.job[i])];
          const float4 so = a.shift[i], sd = a.s_dir[sl];
          lt.start(make_ray(V3{so.x, so.y, so.z}, V3{sd.x, sd.y, sd.z}), sd.w, st);
          have = true;
          ++traced;
        } else {
          drained = true;
        }
      }
    }
    if (__ballot(have) == 0) break;
    if (have && lt.round(sv, st, c)) {
      a.s_vis[sl] = lt.hit < 0 ? 1 : 0;
      have = false;
    }
  }
  ray_count(a.c, 2, traced);
  if (COUNT) work_count(a.c, c);
}

template <int KIND, int WIDTH, bool COUNT>
__global__ void __launch_bounds__(kEngBlock) __attribute__((amdgpu_waves_per_eu(7, 8))) k_eng_mis_dyn(EngArgs a, StackArgs sa, int cur, int fetch_min) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kEngBlock];
  const uint32_t n = a.c->mis_count.v;
  const PathState &ps = a.st[cur];
  auto st = make_stack<KIND>(lds, kEngBlock, sa);
  LaneTrace<WIDTH, false, COUNT> lt;
  DevSceneView sv;
  TraceCounters c;
  WaveFetch wf(n);
  uint32_t i = 0, traced = 0;
  bool have = false, drained = false;
  while (true) {
    const bool need = !have && !drained;
    const uint64_t needm = __ballot(need);
    if (needm && (__popcll(needm) >= fetch_min || __ballot(have) == 0)) {
      const uint32_t k = wf.claim(needm, &a.c->claim[2].v);
      if (need) {
        if (k < n) {
          const float4 mo = a.m_ray[2 * k], md = a.m_ray[2 * k + 1];
          i = __float_as_uint(mo.w);
          sv = a.scenes[PathSceneSlot(ps.job[i])];
          lt.start(make_ray(V3{mo.x, mo.y, mo.z}, V3{md.x, md.y, md.z}), md.w, st);
          have = true;
          ++traced;
        } else {
          drained = true;
        }
      }
    }
    if (__ballot(have) == 0) break;
    if (have && lt.round(sv, st, c)) {
      a.m_hit[i] = lt.hit;  // retire adds the weighted emission when it is a light
      have = false;
    }
  }
  ray_count(a.c, 1, traced);
  if (COUNT) work_count(a.c, c);
}

// ---------------------------------------------------------------- retire
constexpr uint32_t kDoneTable = 64;  // per-block done-count entries (a power of two)

// Adds each shaded path's NEE and MIS terms in a fixed order. A continuing
// path's radiance moves to its place in the next queue; a finished path adds
// it to the job's film and counts towards the job's done counter.
__global__ void __launch_bounds__(kEngBlock) k_eng_retire(EngArgs a, int cur) {
  const uint32_t n = a.c->qcount[cur].v;
  const PathState &ps = a.st[cur];
  const PathState &pn = a.st[cur ^ 1];
  const size_t P = a.capacity;
  __shared__ uint32_t tjob[kDoneTable];
  __shared__ unsigned long long tcnt[kDoneTable];
  for (uint32_t k = threadIdx.x; k < kDoneTable; k += blockDim.x) {
    tjob[k] = 0xffffffffu;
    tcnt[k] = 0;
  }
  __syncthreads();
  for (uint32_t base = blockIdx.x * kEngBlock; base < n; base += gridDim.x * kEngBlock) {
    const uint32_t i = base + threadIdx.x;
    bool fin = false;
    uint32_t js = 0xffffffffu;
    if (i < n) {
      const uint32_t jw = ps.job[i];
      js = PathJobSlot(jw);
      V3 rad{ps.rad[3 * i], ps.rad[3 * i + 1], ps.rad[3 * i + 2]};
      if (a.hit_prim[i] >= 0) {
        const int K = a.jobs[js].K;
        V3 direct{0.0f, 0.0f, 0.0f};
        for (int k = 0; k < K; ++k) {
          const size_t sl = k * P + i;
          if (a.s_vis[sl]) direct = direct + V3{a.s_contrib[3 * sl], a.s_contrib[3 * sl + 1], a.s_contrib[3 * sl + 2]};
        }
        // the MIS BRDF ray: a light hit adds next_T * color * w
        // (muon/integration.cc:490-499), zero otherwise
        V3 madd{0.0f, 0.0f, 0.0f};
        const int32_t mh = a.m_hit[i];
        if (mh >= 0) {
          const DevSceneView &s = a.scenes[PathSceneSlot(jw)];
          const int32_t light = ldg(s.prim_light + mh);
          if (light >= 0) {
            const float *L = s.lights + size_t(light) * kLightStride;
            madd = V3{a.m_fac[3 * i] * ldg(L + 1), a.m_fac[3 * i + 1] * ldg(L + 2), a.m_fac[3 * i + 2] * ldg(L + 3)};
          }
        }
        direct = direct + madd;
        rad = rad + direct;
      }
      const uint32_t ni = a.nidx[i];
      if (ni != 0xffffffffu) {
        pn.rad[3 * ni] = rad.x;
        pn.rad[3 * ni + 1] = rad.y;
        pn.rad[3 * ni + 2] = rad.z;
      } else {
        unsigned long long *px = a.jobs[js].film + 3 * size_t(ps.pixel[i]);
        const unsigned long long fx = to_fixed(rad.x), fy = to_fixed(rad.y), fz = to_fixed(rad.z);
        if (fx) atomic_add_global(px, fx);
        if (fy) atomic_add_global(px + 1, fy);
        if (fz) atomic_add_global(px + 2, fz);
        fin = true;
      }
    }
    // done counters: per block, counts gather in a small LDS table keyed by
    // job slot and reach the job's counter with one atomic per (block, job)
    // at the end (the queue holds the paths of several jobs in block-sized
    // runs; one atomic per wavefront iteration on a handful of addresses
    // would serialise, and the host reads the counters after the kernel)
    uint64_t pending = __ballot(fin);
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const uint32_t jl = uint32_t(__shfl(int(js), leader));
      const uint64_t same = __ballot(fin && js == jl);
      if (__lane_id() == 0) {
        const uint32_t h = jl & (kDoneTable - 1);
        const uint32_t prev = atomicCAS(&tjob[h], 0xffffffffu, jl);
        if (prev == 0xffffffffu || prev == jl) atomicAdd(&tcnt[h], (unsigned long long)__popcll(same));
        else atomic_add_global(a.jobs[jl].done, (unsigned long long)__popcll(same));  // another job holds the entry
      }
      pending &= ~same;
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kDoneTable; k += blockDim.x)
    if (tcnt[k]) atomic_add_global(a.jobs[tjob[k]].done, tcnt[k]);
}

// ---------------------------------------------------------------- plan / fill
// After retire: the queue just consumed is empty; new samples go behind the
// continuing paths of the next queue.
__global__ void k_eng_plan(EngArgs a, int cur) {
  if (blockIdx.x != 0) return;
  EngCounters *c = a.c;
  // the ray-count slots folded by the wavefront's lanes, one slot each (the
  // slots sit on separate cache lines; a single thread walking them one
  // after another was the kernel's whole cost)
  static_assert(kRaySlots <= 64, "one lane per slot");
  const int lane = int(threadIdx.x);
  unsigned long long v[kRayCounters], tot[kRayCounters];
#pragma unroll
  for (int id = 0; id < kRayCounters; ++id) {  // every load first, then the sums
    v[id] = lane < kRaySlots ? c->ray_slots[id][lane].v : 0ull;
    tot[id] = lane == 0 ? c->rays[id] : 0ull;
  }
#pragma unroll
  for (int id = 0; id < kRayCounters; ++id) {
    if (lane < kRaySlots) c->ray_slots[id][lane].v = 0;
    for (int off = 32; off > 0; off >>= 1) v[id] += __shfl_down(v[id], off);
    if (lane == 0) c->rays[id] = tot[id] + v[id];
  }
  if (lane != 0) return;
  c->qcount[cur].v = 0;
  c->mis_count.v = 0;
  c->shadow_count.v = 0;
  c->claim[0].v = c->claim[1].v = c->claim[2].v = 0;
  const uint32_t next = c->qcount[cur ^ 1].v;
  const uint32_t nfill = FillCount(c->total_items, c->cursor, a.capacity, next);
  c->plan_n = nfill;
  c->plan_base = next;
  c->plan_cursor = c->cursor;
  c->cursor += nfill;
  c->qcount[cur ^ 1].v = next + nfill;
}

__global__ void __launch_bounds__(kEngBlock) k_eng_fill(EngArgs a, int cur) {
  const EngCounters *c = a.c;
  const uint32_t n = c->plan_n;
  const uint32_t qbase = c->plan_base;
  const unsigned long long cursor = c->plan_cursor;
  const unsigned long long lo = c->seq_lo, hi = c->seq_hi;
  float4 *qn = a.q_ray[cur ^ 1];
  const PathState &pn = a.st[cur ^ 1];
  for (uint32_t j = blockIdx.x * kEngBlock + threadIdx.x; j < n; j += gridDim.x * kEngBlock) {
    const unsigned long long item = cursor + j;
    // last admitted job whose first item is <= item
    const uint32_t js = uint32_t(FindJobSeq(a.job_first, a.max_jobs, lo, hi, item) % a.max_jobs);
    const JobDev &job = a.jobs[js];
    const ItemPos ip = ItemPosition(item - job.first_item, job.npix, job.sample_base);
    const uint32_t sample = ip.sample, pixel = ip.pixel;
    const int px = int(pixel % uint32_t(job.cam.width)), py = int(pixel / uint32_t(job.cam.width));
    float x, y;
    const PathKey key = path_key(job.seed, job.key, pixel, sample);
    sample_position(px, py, int(sample), key, x, y);
    V3 o, d;
    cast_ray(job.cam, x, y, o, d);
    const uint32_t ni = qbase + j;
    pn.key_a[ni] = key.a;
    pn.key_b[ni] = key.b;
    pn.job[ni] = PackPathJob(js, job.scene);
    pn.pixel[ni] = pixel;
    pn.ctr[ni] = 0;
    pn.depth[ni] = 0;
    pn.thr[3 * ni] = 1.0f;
    pn.thr[3 * ni + 1] = 1.0f;
    pn.thr[3 * ni + 2] = 1.0f;
    pn.rad[3 * ni] = 0.0f;
    pn.rad[3 * ni + 1] = 0.0f;
    pn.rad[3 * ni + 2] = 0.0f;
    qn[2 * ni] = make_float4(o.x, o.y, o.z, 0.0f);
    qn[2 * ni + 1] = make_float4(d.x, d.y, d.z, 0.0f);
  }
}

// Zeroes the done counter and, when the job asks for it, the film of each
// job admitted since the last launch: blockIdx.y picks the job (slots from
// first_slot on, wrapping at max_jobs), blockIdx.x strides over its film.
__global__ void k_admit_clear(const JobDev *jobs, uint32_t first_slot, uint32_t max_jobs) {
  const JobDev &j = jobs[(first_slot + blockIdx.y) % max_jobs];
  if (blockIdx.x == 0 && threadIdx.x == 0) *j.done = 0ull;
  if (!j.clear_film) return;
  const size_t n = size_t(j.npix) * 3;
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += size_t(gridDim.x) * blockDim.x)
    j.film[k] = 0ull;
}

// Fixed-point films to float sums for up to kResolveBatch finished renders:
// blockIdx.y picks the render, blockIdx.x strides over its film.
__global__ void k_film_resolve(FilmResolveBatch b) {
  const FilmResolve &r = b.r[blockIdx.y];
  const uint32_t n = r.npix * 3u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    r.out[i] = float(double(r.film[i]) / kFilmScale);
}

// ---------------------------------------------------------------- host side
namespace {

// Sub-allocates `count` T at address `p` (256-byte aligned) and advances
// `p`; with a zero base it only measures.
template <typename T>
T *Carve(uintptr_t &p, size_t count) {
  T *r = reinterpret_cast<T *>(p);
  p += (count * sizeof(T) + 255) & ~size_t(255);
  return r;
}

// Assigns the engine's arrays from `base` in the order and sizes of
// kPathArrays / kShadowArrays (engine_plan.h); returns the buffer's bytes.
// A null base only measures. Returns 0 if an array's element type does not
// have the size the layout gives it.
class Carver {
 public:
  Carver(uintptr_t base, const ArraySpec *spec, int count, size_t rows) : base_(base), spec_(spec), count_(count) {
    bytes_ = LayoutOffsets(spec, count, rows, off_);
  }
  template <typename T>
  void operator()(T *&ptr) {
    if (k_ >= count_ || spec_[k_].elem_bytes != sizeof(T)) ok_ = false;
    ptr = reinterpret_cast<T *>(base_ + (k_ < count_ ? off_[k_] : 0));
    ++k_;
  }
  size_t bytes() const { return ok_ && k_ == count_ ? bytes_ : 0; }

 private:
  uintptr_t base_;
  const ArraySpec *spec_;
  int count_, k_ = 0;
  bool ok_ = true;
  size_t off_[kNumPathArrays > kNumShadowArrays ? kNumPathArrays : kNumShadowArrays] = {};
  size_t bytes_ = 0;
};

size_t CarvePaths(uintptr_t base, size_t P, EngArgs &a) {
  Carver c(base, kPathArrays, kNumPathArrays, P);
  for (auto &st : a.st) {
    c(st.job);
    c(st.pixel);
    c(st.ctr);
    c(st.depth);
    c(st.rad);
    c(st.thr);
    c(st.key_a);
    c(st.key_b);
  }
  c(a.q_ray[0]);
  c(a.q_ray[1]);
  c(a.m_ray);
  c(a.hit_prim);
  c(a.nidx);
  c(a.m_fac);
  c(a.m_hit);
  c(a.shift);
  c(a.hnorm);
  return c.bytes();
}

size_t CarveShadow(uintptr_t base, size_t P, size_t K, EngArgs &a) {
  Carver c(base, kShadowArrays, kNumShadowArrays, P * K);
  c(a.s_dir);
  c(a.s_contrib);
  c(a.s_vis);
  c(a.s_valid);
  c(a.s_queue);
  return c.bytes();
}

struct IterLaunch {
  const EngArgs &a;
  const TraverseCfg &cfg;
  int cur;
  int fetch;   // per-lane fetch threshold, 0 = grid-stride kernels
  bool count;  // traversal kernels count their box and primitive tests
  bool plain;  // no admitted job has NEE on: one shade kernel, no shadow / MIS rays
  hipStream_t st;
  template <int KIND, bool WW, int WIDTH>
  void run() const {
    if (count) launch<KIND, WW, WIDTH, true>();
    else launch<KIND, WW, WIDTH, false>();
  }
  template <int KIND, bool WW, int WIDTH, bool COUNT>
  void launch() const {
    const StackArgs sa{cfg.spill, cfg.gcap};
    const dim3 grid(cfg.grid), block(kEngBlock);
    const dim3 sgrid(std::max(1, cfg.grid * kEngBlock / kShadeBlock)), sblock(kShadeBlock);
    if (plain) {
      if (fetch > 0)
        hipLaunchKernelGGL((k_eng_extend_dyn<KIND, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur, fetch);
      else
        hipLaunchKernelGGL((k_eng_extend<KIND, WW, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur);
      hipLaunchKernelGGL(k_eng_shade_plain, sgrid, sblock, 0, st, a, cur);
    } else if (fetch > 0) {
      hipLaunchKernelGGL((k_eng_extend_dyn<KIND, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur, fetch);
      hipLaunchKernelGGL(k_eng_shade_light, sgrid, sblock, 0, st, a, cur);
      hipLaunchKernelGGL(k_eng_shade_bsdf, sgrid, sblock, 0, st, a, cur);
      hipLaunchKernelGGL((k_eng_shadow_dyn<KIND, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur, fetch);
      hipLaunchKernelGGL((k_eng_mis_dyn<KIND, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur, fetch);
    } else {
      hipLaunchKernelGGL((k_eng_extend<KIND, WW, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur);
      hipLaunchKernelGGL(k_eng_shade_light, sgrid, sblock, 0, st, a, cur);
      hipLaunchKernelGGL(k_eng_shade_bsdf, sgrid, sblock, 0, st, a, cur);
      hipLaunchKernelGGL((k_eng_shadow<KIND, WW, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur);
      hipLaunchKernelGGL((k_eng_mis<KIND, WW, WIDTH, COUNT>), grid, block, 0, st, a, sa, cur);
    }
    hipLaunchKernelGGL(k_eng_retire, grid, block, 0, st, a, cur);
    hipLaunchKernelGGL(k_eng_plan, dim3(1), dim3(64), 0, st, a, cur);
    hipLaunchKernelGGL(k_eng_fill, grid, block, 0, st, a, cur);
  }
};

}  // namespace

int Engine::Create(muonx_ctx *ctx, const muonx_tuning &tuning, uint32_t capacity, uint32_t max_jobs, int kmax,
                   bool use_graph, Engine **out) {
  *out = nullptr;
  if (capacity == 0 || max_jobs == 0 || kmax < 1) {
    SetError("engine: capacity, max_jobs and kmax must be positive");
    return MUONX_ERR_INVALID;
  }
  if (max_jobs > kMaxPathSlotIndex) {  // a path's job word holds the job slot in 16 bits
    SetError("engine: max_jobs must be at most 65536");
    return MUONX_ERR_INVALID;
  }
  if (capacity > (1u << 30) || uint64_t(capacity) * uint64_t(kmax) > kMaxShadowSlots) {
    SetError("engine: paths in flight x shadow rays per vertex must stay below 2^31");
    return MUONX_ERR_INVALID;
  }
  Engine *e = new Engine();
  e->ctx_ = ctx;
  e->tuning_ = tuning;
  e->num_cus_ = ctx->num_cus;
  e->capacity_ = capacity;
  e->max_jobs_ = max_jobs;
  e->kmax_ = kmax;
  e->use_graph_ = use_graph;
  auto fail = [&](int code) {
    delete e;
    return code;
  };
  if (hipStreamCreateWithFlags(&e->stream_, hipStreamNonBlocking) != hipSuccess) {
    SetError("engine: stream creation failed");
    return fail(MUONX_ERR_HIP);
  }
  // The buffers are carved twice: from a null base to size them, then for real.
  EngArgs &a = e->a_;
  size_t path_bytes = 0, shadow_bytes = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const uintptr_t b0 = pass ? reinterpret_cast<uintptr_t>(e->paths_.ptr) : 0;
    const uintptr_t b1 = pass ? reinterpret_cast<uintptr_t>(e->shadow_.ptr) : 0;
    path_bytes = CarvePaths(b0, capacity, a);
    shadow_bytes = CarveShadow(b1, capacity, size_t(kmax), a);
    if (path_bytes == 0 || shadow_bytes == 0) {
      SetError("engine: buffer layout does not match the path arrays");
      return fail(MUONX_ERR_STATE);
    }
    if (pass == 0 &&
        (e->paths_.Alloc(path_bytes) != hipSuccess || e->shadow_.Alloc(shadow_bytes) != hipSuccess ||
         e->counters_.Alloc(sizeof(EngCounters)) != hipSuccess ||
         e->jobs_.Alloc(sizeof(JobDev) * max_jobs) != hipSuccess ||
         e->job_first_.Alloc(sizeof(unsigned long long) * max_jobs) != hipSuccess ||
         e->done_.Alloc(sizeof(unsigned long long) * max_jobs) != hipSuccess)) {
      SetError("engine: out of device memory for " + std::to_string(capacity) + " paths");
      return fail(MUONX_ERR_NOMEM);
    }
  }
  e->device_bytes_ = path_bytes + shadow_bytes;
  if (hipHostMalloc(reinterpret_cast<void **>(&e->h_jobs_), sizeof(JobDev) * max_jobs) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&e->h_first_), sizeof(unsigned long long) * max_jobs) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&e->h_window_), sizeof(unsigned long long) * 3 * 4) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&e->h_done_[0]), sizeof(unsigned long long) * (max_jobs + kRayCounters)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&e->h_done_[1]), sizeof(unsigned long long) * (max_jobs + kRayCounters)) != hipSuccess) {
    SetError("engine: pinned host allocation failed");
    return fail(MUONX_ERR_NOMEM);
  }
  for (auto &ev : e->window_ev_)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(MUONX_ERR_HIP);
  for (auto &ev : e->done_ev_)
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(MUONX_ERR_HIP);
  a.c = e->counters_.as<EngCounters>();
  a.jobs = e->jobs_.as<JobDev>();
  a.job_first = e->job_first_.as<unsigned long long>();
  a.max_jobs = max_jobs;
  a.capacity = capacity;
  a.kmax = kmax;
  a.scenes = nullptr;
  if (hipMemsetAsync(a.c, 0, sizeof(EngCounters), e->stream_) != hipSuccess ||
      hipMemsetAsync(e->done_.ptr, 0, sizeof(unsigned long long) * max_jobs, e->stream_) != hipSuccess)
    return fail(MUONX_ERR_HIP);
  *out = e;
  return MUONX_OK;
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  DropGraphs();
  for (auto &ev : window_ev_)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : done_ev_)
    if (ev) (void)hipEventDestroy(ev);
  if (h_jobs_) (void)hipHostFree(h_jobs_);
  if (h_first_) (void)hipHostFree(h_first_);
  if (h_window_) (void)hipHostFree(h_window_);
  for (auto *h : h_done_)
    if (h) (void)hipHostFree(h);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::AddScene(muonx_scene *scene, uint32_t *slot) {
  // entries are keyed by scene id (ids are never reused)
  for (size_t i = 0; i < scene_ids_.size(); ++i) {
    if (scene_ids_[i] == scene->id) {
      *slot = uint32_t(i);
      return MUONX_OK;
    }
  }
  // the scene table and spill region may be rewritten or reallocated below
  MX_HIP(hipStreamSynchronize(stream_));
  size_t at = scene_ids_.size();
  for (size_t i = 0; i < scene_ids_.size(); ++i) {
    if (!SceneLive(scene_ids_[i])) {  // a destroyed scene's entry: no job can refer to it
      at = i;
      break;
    }
  }
  if (at == scene_ids_.size()) {
    if (at >= kMaxPathSlotIndex) {  // a path's job word holds the scene slot in 16 bits
      SetError("engine: 65536 live scenes in use");
      return MUONX_ERR_CAPACITY;
    }
    scene_ids_.push_back(scene->id);
    host_scenes_.push_back(scene->view);
    scene_ptrs_.push_back(scene);
  } else {
    scene_ids_[at] = scene->id;
    host_scenes_[at] = scene->view;
    scene_ptrs_[at] = scene;
  }
  MX_HIP(scenes_.Reserve(sizeof(DevSceneView) * std::max<size_t>(16, host_scenes_.capacity())));
  MX_HIP(hipMemcpyAsync(scenes_.ptr, host_scenes_.data(), sizeof(DevSceneView) * host_scenes_.size(),
                        hipMemcpyHostToDevice, stream_));
  MX_HIP(hipStreamSynchronize(stream_));  // host_scenes_ may reallocate later
  a_.scenes = scenes_.as<DevSceneView>();
  *slot = uint32_t(at);
  // the traversal plan covers the deepest resident BVH
  const uint32_t depth = std::max(max_depth_, scene->view.bvh_depth);
  const uint32_t wstack = std::max(max_wstack_, scene->view.wstack);
  if (depth != max_depth_ || wstack != max_wstack_ || !cfg_.grid) {
    max_depth_ = depth;
    max_wstack_ = wstack;
    TraverseCfg cfg;
    const int rc = PlanTraverseInto(tuning_, num_cus_, max_depth_, max_wstack_, uint64_t(capacity_) * kmax_,
                                    kEngBlock, spill_, cfg, false);
    if (rc != MUONX_OK) return rc;
    cfg_ = cfg;
  }
  DropGraphs();  // a new plan, or the scene table pointer may have moved
  return MUONX_OK;
}

void Engine::DropGraphs() {
  for (auto &set : graph_)
    for (auto &g : set) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

namespace {
// Quad-light sampling grid of ShadeDirectNEE (muon/integration.cc:390-400):
// strata x strata cells with `per` samples each. glm::sqrt(int) is a double
// square root; non-positive sample counts draw nothing.
void LightGrid(const muonx_render_params &p, int *strata, int *per) {
  if (p.light_samples <= 0) {
    *strata = p.light_stratify ? 0 : 1;
    *per = p.light_stratify ? 1 : 0;
    return;
  }
  *strata = p.light_stratify ? int(std::sqrt(double(p.light_samples))) : 1;
  *per = p.light_stratify ? 1 : p.light_samples;
}
}  // namespace

int Engine::ShadowRaysPerVertex(const muonx_scene *scene, const muonx_render_params &p) {
  if (p.nee == MUONX_NEE_OFF || scene->view.num_lights == 0) return 1;
  int strata, per;
  LightGrid(p, &strata, &per);
  int64_t K = 0;
  for (uint32_t l = 0; l < scene->view.num_lights; ++l) {
    if (int(scene->host_lights[size_t(l) * kLightStride]) == MUONX_LIGHT_QUAD) K += int64_t(strata) * strata * per;
    else K += 1;
  }
  return int(std::min<int64_t>(std::max<int64_t>(1, K), INT_MAX));
}

int Engine::Admit(const JobSpec &js, uint64_t *seq) {
  if (seq_hi_ - seq_lo_ >= max_jobs_) {
    SetError("engine: all job slots are in use");
    return MUONX_ERR_CAPACITY;
  }
  const muonx_render_params &p = js.params;
  const muonx_scene *scene = scene_ptrs_.at(js.scene_slot);
  const int K = ShadowRaysPerVertex(scene, p);
  if (K > kmax_) {
    SetError("engine: job needs " + std::to_string(K) + " shadow rays per vertex, engine has " +
             std::to_string(kmax_));
    return MUONX_ERR_CAPACITY;
  }
  const uint32_t slot = uint32_t(seq_hi_ % max_jobs_);
  JobDev &j = h_jobs_[slot];
  j.cam = MakeCam(p);
  j.scene = js.scene_slot;
  j.npix = uint32_t(p.width) * uint32_t(p.height);
  j.seed = p.seed;
  j.key = js.key;
  j.min_depth = p.min_depth;
  j.cutoff = p.max_depth == -1 ? INT_MAX : (p.nee != MUONX_NEE_OFF ? p.max_depth - 1 : p.max_depth);
  j.nee = p.nee;
  j.is_mode = p.importance_sampling;
  j.rr = p.russian_roulette;
  j.light_samples = std::max(1, p.light_samples);
  LightGrid(p, &j.strata, &j.per_section);
  j.K = (p.nee == MUONX_NEE_OFF || scene->view.num_lights == 0) ? 0 : K;
  if (js.sample_count == 0 || uint64_t(js.sample_begin) + js.sample_count > uint64_t(p.spp)) {
    SetError("engine: sample slice outside [0, spp)");
    return MUONX_ERR_INVALID;
  }
  j.first_item = total_items_;
  j.nitems = uint64_t(j.npix) * uint64_t(js.sample_count);
  j.sample_base = js.sample_begin;
  j.film = js.film;
  j.done = done_.as<unsigned long long>() + slot;
  j.clear_film = js.clear_film ? 1u : 0u;
  h_first_[slot] = j.first_item;
  // the record reaches the device, and the film and done counter are
  // cleared, with the other jobs admitted before the next launch (FlushAdmitted)
  *seq = seq_hi_;
  ++seq_hi_;
  total_items_ += j.nitems;
  return MUONX_OK;
}

int Engine::FlushAdmitted() {
  if (flushed_hi_ == seq_hi_) return MUONX_OK;
  const uint64_t lo = flushed_hi_, hi = seq_hi_;
  SlotRun runs[2];
  const int nr = SlotRuns(lo, hi, max_jobs_, runs);  // the new slots: at most two runs of the ring
  for (int r = 0; r < nr; ++r) {
    const uint32_t s0 = runs[r].first, run = runs[r].count;
    MX_HIP(hipMemcpyAsync(jobs_.as<JobDev>() + s0, h_jobs_ + s0, sizeof(JobDev) * run, hipMemcpyHostToDevice,
                          stream_));
    MX_HIP(hipMemcpyAsync(job_first_.as<unsigned long long>() + s0, h_first_ + s0,
                          sizeof(unsigned long long) * run, hipMemcpyHostToDevice, stream_));
  }
  constexpr uint64_t kMaxGridY = 65535;  // the grid's y extent holds one job each
  for (uint64_t q = lo; q < hi; q += kMaxGridY) {
    const uint32_t m = uint32_t(std::min<uint64_t>(hi - q, kMaxGridY));
    hipLaunchKernelGGL(k_admit_clear, dim3(kClearBlocksPerJob, m), dim3(256), 0, stream_, jobs_.as<JobDev>(),
                       uint32_t(q % max_jobs_), max_jobs_);
  }
  MX_HIP(hipGetLastError());
  flushed_hi_ = hi;
  return MUONX_OK;
}

int Engine::SyncWindow() {
  const int frc = FlushAdmitted();
  if (frc != MUONX_OK) return frc;
  if (window_written_[0] == seq_lo_ && window_written_[1] == seq_hi_ && window_written_[2] == total_items_)
    return MUONX_OK;
  const int k = window_slot_;
  window_slot_ = (window_slot_ + 1) & 3;
  MX_HIP(hipEventSynchronize(window_ev_[k]));  // the copy that last used this staging slot
  unsigned long long *w = h_window_ + 3 * k;
  w[0] = total_items_;
  w[1] = seq_lo_;
  w[2] = seq_hi_;
  // total_items, seq_lo, seq_hi are consecutive in EngCounters
  MX_HIP(hipMemcpyAsync(&a_.c->total_items, w, 3 * sizeof(unsigned long long), hipMemcpyHostToDevice, stream_));
  MX_HIP(hipEventRecord(window_ev_[k], stream_));
  window_written_[0] = seq_lo_;
  window_written_[1] = seq_hi_;
  window_written_[2] = total_items_;
  return MUONX_OK;
}

bool Engine::PlainWindow() const {
  for (uint64_t q = seq_lo_; q < seq_hi_; ++q)
    if (h_jobs_[q % max_jobs_].nee != MUONX_NEE_OFF) return false;
  return true;
}

int Engine::LaunchIteration(int cur, bool plain) {
  DispatchTraverse(cfg_, IterLaunch{a_, cfg_, cur, tuning_.ray_fetch, tuning_.count_work != 0, plain, stream_});
  MX_HIP(hipGetLastError());
  return MUONX_OK;
}

int Engine::Prime() {
  int rc = SyncWindow();
  if (rc != MUONX_OK) return rc;
  hipLaunchKernelGGL(k_eng_plan, dim3(1), dim3(64), 0, stream_, a_, cur_);
  hipLaunchKernelGGL(k_eng_fill, dim3(cfg_.grid), dim3(kEngBlock), 0, stream_, a_, cur_);
  MX_HIP(hipGetLastError());
  cur_ ^= 1;
  return MUONX_OK;
}

int Engine::Iterate() {
  if (!a_.scenes) {
    SetError("engine: no scene registered");
    return MUONX_ERR_STATE;
  }
  int rc = SyncWindow();
  if (rc != MUONX_OK) return rc;
  // Jobs below seq_lo have no path in flight, so with no NEE job in the
  // window no path needs the NEE / MIS kernels.
  const bool plain = PlainWindow();
  hipGraphExec_t &graph = graph_[plain ? 1 : 0][cur_];
  if (use_graph_) {
    if (!graph) {
      hipGraph_t g;
      MX_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      rc = LaunchIteration(cur_, plain);
      hipGraph_t captured = nullptr;
      const hipError_t ec = hipStreamEndCapture(stream_, &captured);
      if (rc != MUONX_OK) return rc;
      if (ec != hipSuccess) {
        SetError(std::string("engine: graph capture failed: ") + hipGetErrorString(ec));
        return MUONX_ERR_HIP;
      }
      g = captured;
      MX_HIP(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
      MX_HIP(hipGraphDestroy(g));
    }
    MX_HIP(hipGraphLaunch(graph, stream_));
  } else {
    rc = LaunchIteration(cur_, plain);
    if (rc != MUONX_OK) return rc;
  }
  cur_ ^= 1;
  return MUONX_OK;
}

int Engine::SnapshotDone(int b) {
  MX_HIP(hipMemcpyAsync(h_done_[b], done_.ptr, sizeof(unsigned long long) * max_jobs_, hipMemcpyDeviceToHost,
                        stream_));
  MX_HIP(hipMemcpyAsync(h_done_[b] + max_jobs_, &a_.c->rays[0], kRayCounters * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, stream_));
  MX_HIP(hipEventRecord(done_ev_[b], stream_));
  return MUONX_OK;
}

int Engine::WaitDone(int b) {
  MX_HIP(hipEventSynchronize(done_ev_[b]));
  return MUONX_OK;
}

int Engine::ResolveFilms(const FilmResolve *films, size_t n) {
  for (size_t k = 0; k < n; k += kResolveBatch) {
    FilmResolveBatch b{};
    const uint32_t m = uint32_t(std::min<size_t>(kResolveBatch, n - k));
    uint32_t most = 0;
    for (uint32_t q = 0; q < m; ++q) {
      b.r[q] = films[k + q];
      most = std::max(most, films[k + q].npix * 3u);
    }
    hipLaunchKernelGGL(k_film_resolve, dim3(std::min<uint32_t>(1024, (most + 255) / 256), m), dim3(256), 0,
                       stream_, b);
    MX_HIP(hipGetLastError());
  }
  return MUONX_OK;
}

int Engine::ResetRayCounts() {
  MX_HIP(hipMemsetAsync(&a_.c->rays[0], 0, kRayCounters * sizeof(unsigned long long), stream_));
  return MUONX_OK;
}

int Engine::RayCounts(unsigned long long out[kRayCounters]) {
  MX_HIP(hipMemcpyAsync(out, &a_.c->rays[0], kRayCounters * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        stream_));
  MX_HIP(hipStreamSynchronize(stream_));
  return MUONX_OK;
}

}  // namespace muonx
