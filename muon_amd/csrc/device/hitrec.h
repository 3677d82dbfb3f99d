// Full intersection record of a known primitive (muon/objects.cc:29-46 with
// the Tri / Sphere object-space tests): world position, world normal
// normalize(M^-T n) and world distance, recomputed for the closest hit found
// by the traversal. Every step is the reference's arithmetic, so the record
// carries the bits muon's Intersection would hold.
#ifndef MUONX_DEVICE_HITREC_H_
#define MUONX_DEVICE_HITREC_H_

#include "trace.h"

namespace muonx {

struct HitRecord {
  V3 pos, normal;
  float dist;
};

__device__ __forceinline__ HitRecord hit_record(const DevSceneView &s, uint32_t i, V3 ro, V3 rd) {
  const float4 q0 = ldg(s.prims + 3 * i);
  const float4 q1 = ldg(s.prims + 3 * i + 1);
  const uint32_t tw = __float_as_uint(q0.w);
  const float *xf = s.xforms + size_t(tw & 0x7fffffffu) * 48;
  V3 o, d;
  to_object(xf, ro, rd, o, d);
  V3 p, n;
  if (tw & kSphereBit) {
    const V3 c{q0.x, q0.y, q0.z};
    sphere_test(c, q1.x, o, d, p);
    n = normalize(p - c);
  } else {
    const float4 q2 = ldg(s.prims + 3 * i + 2);
    const V3 a{q0.x, q0.y, q0.z}, b{q1.x, q1.y, q1.z}, c{q2.x, q2.y, q2.z};
    const V3 fn = cross(b - a, c - a);
    const float dn = dot(fn, d);
    const float t = (dot(a, fn) - dot(o, fn)) / dn;
    p = o + d * t;
    if (s.prim_use_vn && ldg(s.prim_use_vn + i)) {
      const float n2 = dot(fn, fn);
      float v = dot(fn, cross(b - a, p - a));
      float w = dot(fn, cross(c - b, p - b));
      float u = dot(fn, cross(a - c, p - c));
      u = u / n2;
      v = v / n2;
      w = w / n2;
      const float *N = s.prim_vn + size_t(i) * 9;
      n = (w * V3{ldg(N), ldg(N + 1), ldg(N + 2)} + u * V3{ldg(N + 3), ldg(N + 4), ldg(N + 5)}) +
          v * V3{ldg(N + 6), ldg(N + 7), ldg(N + 8)};
    } else {
      n = normalize(fn);
    }
  }
  HitRecord h;
  float m[16], mt[16];
  load_xform(xf, m);
  load_xform(xf + 32, mt);
  h.pos = xform_point(m, p);
  h.normal = normalize(xform_vec(mt, n));
  h.dist = length(h.pos - ro);
  return h;
}

}  // namespace muonx

#endif
