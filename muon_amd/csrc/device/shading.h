// Device BRDFs, hemisphere sampling, lights and MIS helpers, restating
// muon/materials.cc, muon/hemisphere_sampling.cc, muon/lighting.cc and the
// sampling helpers of muon/integration.cc with the same formulas. Where the
// reference converts a cosine or tangent to an angle only to take its cosine,
// sine or tangent again (the samplers' polar angle, GGX's D and G1) the
// algebraic identity is used instead: cos(acos x) = x,
// sin(acos x) = sqrt(1 - x^2), tan^2(acos x) = (1 - x^2) / x^2 and
// (cos, sin)(atan(a / b)) = (b, a) / sqrt(a^2 + b^2). Those stages and the
// device math library's transcendentals make these match the reference
// statistically (same distributions and estimators), not bit for bit.
//
// The estimator's own arithmetic (sampling, BRDF values, pdfs, MIS weights,
// light-sample weights) runs on the hardware's single-instruction reciprocal,
// square root, reciprocal square root, sine / cosine and exp2 / log2 (about
// 1 ulp, sine / cosine about 2^-20 absolute) instead of the correctly
// rounded library sequences: the error is many orders below the estimator's
// noise and carries no sign, so it cannot shift an expected value. Geometry
// (hit records, intersection tests, ray origins) stays correctly rounded.
#ifndef MUONX_DEVICE_SHADING_H_
#define MUONX_DEVICE_SHADING_H_

#include "../common/vecmath.h"
#include "dev_scene.h"
#include "gload.h"
#include "rng.h"

namespace muonx {

// Estimator arithmetic (see the file comment).
__device__ __forceinline__ float est_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float est_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float est_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float est_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ V3 est_div(V3 a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ V3 est_normalize(V3 v) { return v * __builtin_amdgcn_rsqf(dot(v, v)); }
__device__ __forceinline__ float est_length(V3 v) { return __builtin_amdgcn_sqrtf(dot(v, v)); }
// x^y for x >= 0 (powf's values at x = 0: 1 for y = 0, else 0 for y > 0)
__device__ __forceinline__ float est_pow(float x, float y) {
  if (!(x > 0.0f)) return y == 0.0f ? 1.0f : 0.0f;
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

// muon/transform.cc:5-19
__device__ __forceinline__ V3 rotate_to_frame(V3 s, V3 w) {
  const V3 a = fabsf(w.y) <= 0.9f ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
  const V3 u = est_normalize(cross(a, w));
  const V3 v = est_normalize(cross(w, u));
  return (s.x * u + s.y * v) + s.z * w;
}

// (cos phi sin theta, sin phi sin theta, cos theta) from the polar angle's
// cosine and sine, with phi = 2 pi * turns (the hardware's sine and cosine
// take the angle in turns).
__device__ __forceinline__ V3 spherical(float ct, float st, float turns) {
  const float sp = __builtin_amdgcn_sinf(turns), cp = __builtin_amdgcn_cosf(turns);
  return V3{cp * st, sp * st, ct};
}

// muon/hemisphere_sampling.cc:8-22 (theta = acos(r1), phi = 2 pi r2)
__device__ __forceinline__ V3 sample_hemisphere(V3 n, PathKey key, uint32_t &ctr) {
  const float r1 = rng_uniform(key, ctr);
  const float r2 = rng_uniform(key, ctr);
  return rotate_to_frame(spherical(r1, est_sqrt(gmax(1.0f - r1 * r1, 0.0f)), r2), n);
}

// muon/hemisphere_sampling.cc:24-38 (theta = acos(sqrt(r1)), phi = 2 pi r2)
__device__ __forceinline__ V3 sample_cosine(V3 n, PathKey key, uint32_t &ctr) {
  const float r1 = rng_uniform(key, ctr);
  const float r2 = rng_uniform(key, ctr);
  return rotate_to_frame(spherical(est_sqrt(r1), est_sqrt(1.0f - r1), r2), n);
}

struct Mat {
  V3 diffuse, specular, emission;
  float shininess, roughness, refl;
  int brdf;
};

__device__ __forceinline__ Mat load_mat(const float *m) {
  Mat r;
  r.diffuse = V3{ldg(m + 3), ldg(m + 4), ldg(m + 5)};
  r.specular = V3{ldg(m + 6), ldg(m + 7), ldg(m + 8)};
  r.emission = V3{ldg(m + 9), ldg(m + 10), ldg(m + 11)};
  r.shininess = ldg(m + 12);
  r.roughness = ldg(m + 13);
  r.brdf = int(ldg(m + 14));
  r.refl = ldg(m + 15);
  return r;
}

// muon/materials.cc:131-141: a^2 / (pi cos^4 (a^2 + tan^2)^2) with cos = h.n
// clamped to [0, 1], written as a^2 / (pi (cos^2 (a^2 - 1) + 1)^2).
__device__ __forceinline__ float ggx_d(V3 h, V3 n, float rough) {
  const float c = gclamp(dot(h, n), 0.0f, 1.0f);
  const float a2 = rough * rough;
  const float q = c * c * (a2 - 1.0f) + 1.0f;
  const float denom = kPi * (q * q);
  return denom > 0.0f ? est_div(a2, denom) : 0.0f;
}

// muon/materials.cc:148-158: 2 / (1 + sqrt(1 + a^2 tan^2)), tan^2 from v.n.
__device__ __forceinline__ float smith_g1(V3 v, V3 n, float rough) {
  const float vn = dot(v, n);
  if (vn <= 0.0f) return 0.0f;
  const float c = gmin(vn, 1.0f);
  const float a2 = rough * rough;
  const float t2 = est_div(1.0f - c * c, c * c);
  return est_div(2.0f, 1.0f + est_sqrt(1.0f + a2 * t2));
}

// BRDF::Sample: Lambertian muon/materials.cc:17-24, Phong :49-71, GGX :180-205.
__device__ __forceinline__ V3 brdf_sample(const Mat &m, V3 rd, V3 n, PathKey key, uint32_t &ctr) {
  if (m.brdf == MUONX_BRDF_LAMBERTIAN) return sample_cosine(n, key, ctr);
  const float r0 = rng_uniform(key, ctr);
  if (r0 > m.refl) return sample_cosine(n, key, ctr);
  const float r1 = rng_uniform(key, ctr);
  const float r2 = rng_uniform(key, ctr);
  if (m.brdf == MUONX_BRDF_PHONG) {
    // theta = acos(r1^(1 / (s + 1))), phi = 2 pi r2
    const float ct = est_pow(r1, est_rcp(m.shininess + 1.0f));
    return rotate_to_frame(spherical(ct, est_sqrt(gmax(1.0f - ct * ct, 0.0f)), r2), reflect(rd, n));
  }
  // theta = atan(a sqrt(r1) / sqrt(1 - r1)), phi = 2 pi r2; 1 - r1 > 0 (r1 < 1)
  const float b = est_sqrt(1.0f - r1), a = m.roughness * est_sqrt(r1);
  const float inv = est_rsqrt(b * b + a * a);
  const V3 h = rotate_to_frame(spherical(b * inv, a * inv, r2), n);
  return reflect(rd, h);
}

// BRDF::PDF: muon/materials.cc:26-34, :73-91, :207-225.
__device__ __forceinline__ float brdf_pdf(const Mat &m, V3 wi, V3 rd, V3 n) {
  const float cos_term = gmax(dot(n, wi), 0.0f) * kOneOverPi;
  if (m.brdf == MUONX_BRDF_LAMBERTIAN) return cos_term;
  const float t = m.refl;
  if (m.brdf == MUONX_BRDF_PHONG) {
    return (1.0f - t) * cos_term +
           t * (m.shininess + 1.0f) * 0.5f * kOneOverPi * est_pow(gmax(dot(reflect(rd, n), wi), 0.0f), m.shininess);
  }
  const V3 h = est_normalize(wi - rd);
  return (1.0f - t) * cos_term + est_div(t * ggx_d(h, n, m.roughness) * gmax(dot(n, h), 0.0f), 4.0f * dot(h, wi));
}

// BRDF::Eval: muon/materials.cc:36-43, :93-111, :227-266.
__device__ __forceinline__ V3 brdf_eval(const Mat &m, V3 wi, V3 rd, V3 n) {
  const V3 kd = m.diffuse * kOneOverPi;
  if (m.brdf == MUONX_BRDF_LAMBERTIAN) return kd;
  if (m.brdf == MUONX_BRDF_PHONG) {
    return kd + m.specular * (m.shininess + 2.0f) * (0.5f * kOneOverPi) *
                    est_pow(gmax(dot(reflect(rd, n), wi), 0.0f), m.shininess);
  }
  const V3 h = est_normalize(wi - rd);
  const float one_m = 1.0f - dot(wi, h);
  const float one_m2 = one_m * one_m;
  const float f5 = one_m2 * one_m2 * one_m;
  const V3 fres = m.specular + (1.0f - m.specular) * f5;
  const float g = smith_g1(wi, n, m.roughness) * smith_g1(-rd, n, m.roughness);
  const float d = ggx_d(h, n, m.roughness);
  return kd + est_div(fres * g * d, 4.0f * dot(wi, n) * dot(-rd, n));
}

// Light record accessors (muonx_scene_desc light layout).
struct QuadLight {
  V3 color, corner, e0, e1;
  float area;
};
__device__ __forceinline__ QuadLight load_quad(const float *L) {
  QuadLight q;
  q.color = V3{ldg(L + 1), ldg(L + 2), ldg(L + 3)};
  q.corner = V3{ldg(L + 4), ldg(L + 5), ldg(L + 6)};
  q.e0 = V3{ldg(L + 7), ldg(L + 8), ldg(L + 9)};
  q.e1 = V3{ldg(L + 10), ldg(L + 11), ldg(L + 12)};
  q.area = ldg(L + 15);
  return q;
}

// Quad::Intersect (muon/lighting.cc:7-36); returns the distance from the
// origin to the hit point, or -1.
__device__ __forceinline__ float quad_hit(const QuadLight &q, V3 o, V3 d) {
  const V3 nrm = est_normalize(cross(q.e1, q.e0));
  const float rdn = dot(d, nrm);
  if (rdn == 0.0f) return -1.0f;
  const float t = est_div(dot(q.corner - o, nrm), rdn);
  if (t < 0.0f) return -1.0f;
  const V3 p = o + d * t;
  const V3 pe = p - q.corner;
  const float l0 = est_length(q.e0), l1 = est_length(q.e1);
  const float u = est_div(dot(pe, q.e0), l0);
  const float v = est_div(dot(pe, q.e1), l1);
  if (u > 0.0f && u < l0 && v > 0.0f && v < l1) return est_length(p - o);
  return -1.0f;
}

// PathTracer::NEEPDF (muon/integration.cc:564-608): area-measure pdf of the
// ray's direction summed over the quad lights it hits, averaged over all
// quad lights.
__device__ __forceinline__ float nee_pdf(const DevSceneView &s, V3 o, V3 d) {
  float pdf = 0.0f;
  int affecting = 0;
  for (uint32_t l = 0; l < s.num_lights; ++l) {
    const float *L = s.lights + size_t(l) * kLightStride;
    if (int(ldg(L)) != MUONX_LIGHT_QUAD) continue;
    ++affecting;
    const QuadLight q = load_quad(L);
    const float dist = quad_hit(q, o, d);
    if (dist < 0.0f) continue;
    const V3 rn = est_normalize(cross(q.e0, q.e1));
    const float cosl = fabsf(dot(rn, d));
    const float denom = q.area * cosl;
    if (denom == 0.0f) continue;
    pdf += est_div(dist * dist, denom);
  }
  if (affecting == 0) return 0.0f;
  return est_div(pdf, float(affecting));
}

// muon/integration.cc:332-338
__device__ __forceinline__ float power_heuristic(float p0, float p1) {
  const float denom = p0 * p0 + p1 * p1;
  if (denom == 0.0f) return 0.0f;
  return est_div(p0 * p0, denom);
}

// PathTracer::ImportanceSamplingPDF (muon/integration.cc:610-630)
__device__ __forceinline__ float is_pdf(int mode, const Mat &m, V3 wi, V3 rd, V3 n) {
  if (mode == MUONX_IS_HEMISPHERE) return 0.5f * kOneOverPi;
  if (mode == MUONX_IS_COSINE) return gmax(dot(n, wi), 0.0f) * kOneOverPi;
  return brdf_pdf(m, wi, rd, n);
}

// PathTracer::SampleReflection (muon/integration.cc:503-562).
__device__ __forceinline__ bool sample_reflection(int mode, const Mat &m, V3 rd, V3 n, V3 T, PathKey key,
                                                  uint32_t &ctr, V3 &dir, float &pdf, V3 &next_T) {
  if (mode == MUONX_IS_HEMISPHERE) dir = sample_hemisphere(n, key, ctr);
  else if (mode == MUONX_IS_COSINE) dir = sample_cosine(n, key, ctr);
  else dir = brdf_sample(m, rd, n, key, ctr);
  if (dot(dir, n) <= 0.0f) return false;
  const V3 f = brdf_eval(m, dir, rd, n);
  pdf = is_pdf(mode, m, dir, rd, n);
  if (mode == MUONX_IS_HEMISPHERE) next_T = T * 2.0f * kPi * f * gmax(dot(n, dir), 0.0f);
  else if (mode == MUONX_IS_COSINE) next_T = T * kPi * f;
  else next_T = est_div(T * f, pdf) * gmax(dot(n, dir), 0.0f);
  return true;
}

}  // namespace muonx

#endif
