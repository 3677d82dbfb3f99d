// Scalar FP32 vector math shared by the host ingest code and the gfx950
// kernels. Operation order follows the glm 0.9.9 routines muon calls, so
// that host and device produce the reference's bits for +,-,*,/,sqrt
// (everything is compiled with -ffp-contract=off and IEEE division/sqrt):
//   dot    ((x*x' + y*y') + z*z')        glm/detail/func_geometric.inl:48-55
//   cross  glm/detail/func_geometric.inl:68-79
//   normalize v * (1 / sqrt(dot(v,v)))    glm/detail/func_geometric.inl:82-90
//   min/max (y<x ? y : x), (x<y ? y : x)  glm/detail/func_common.inl:17-30
//   mat4*vec4 (c0*x + c1*y) + (c2*z + c3*w) glm/detail/type_mat4x4.inl:536-572
#ifndef MUONX_COMMON_VECMATH_H_
#define MUONX_COMMON_VECMATH_H_

#include <cmath>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define MX_HD __host__ __device__ __forceinline__
#else
#define MX_HD inline
#endif

namespace muonx {

struct V3 {
  float x, y, z;
};

MX_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
MX_HD V3 v3s(float s) { return V3{s, s, s}; }
MX_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MX_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MX_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
MX_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
MX_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
MX_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
MX_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
MX_HD V3 operator+(V3 a, float s) { return V3{a.x + s, a.y + s, a.z + s}; }
MX_HD V3 operator-(float s, V3 a) { return V3{s - a.x, s - a.y, s - a.z}; }

MX_HD float gmin(float x, float y) { return (y < x) ? y : x; }
MX_HD float gmax(float x, float y) { return (x < y) ? y : x; }
MX_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

MX_HD float dot(V3 a, V3 b) {
  float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
  return (px + py) + pz;
}
MX_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
MX_HD float length(V3 v) { return sqrtf(dot(v, v)); }
MX_HD V3 normalize(V3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
MX_HD float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// Column-major 4x4 affine-or-not matrix, m[col*4 + row].
// Position transform: xyz of mat4 * (p, 1).
MX_HD V3 xform_point(const float *m, V3 p) {
  float a0x = m[0] * p.x + m[4] * p.y, a0y = m[1] * p.x + m[5] * p.y, a0z = m[2] * p.x + m[6] * p.y;
  float a1x = m[8] * p.z + m[12] * 1.0f, a1y = m[9] * p.z + m[13] * 1.0f, a1z = m[10] * p.z + m[14] * 1.0f;
  return V3{a0x + a1x, a0y + a1y, a0z + a1z};
}
// Direction transform before normalisation: xyz of mat4 * (d, 0).
MX_HD V3 xform_vec(const float *m, V3 d) {
  float a0x = m[0] * d.x + m[4] * d.y, a0y = m[1] * d.x + m[5] * d.y, a0z = m[2] * d.x + m[6] * d.y;
  float a1x = m[8] * d.z + m[12] * 0.0f, a1y = m[9] * d.z + m[13] * 0.0f, a1z = m[10] * d.z + m[14] * 0.0f;
  return V3{a0x + a1x, a0y + a1y, a0z + a1z};
}
// muon/transform.h:26-28
MX_HD V3 reflect(V3 d, V3 n) { return d - ((2.0f * dot(d, n)) * n); }

constexpr float kPi = 3.14159265358979323846264338327950288f;
constexpr float kOneOverPi = 0.318309886183790671537767526745028724f;
constexpr float kEps = 0.0001f;  // muon/types.h:13

}  // namespace muonx

#endif
