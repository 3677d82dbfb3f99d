// ORACLE — test infrastructure only. Nothing in the muon_amd product links,
// imports or executes this code; only tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py use it, as the oracle results are checked against.
//
// Scalar FP32 vector / matrix math restated in the operation order of the
// glm 0.9.9 routines the reference calls (no SIMD, no FMA contraction; the
// oracle is compiled with -ffp-contract=off):
//   dot       third_party/glm/detail/func_geometric.inl:48-55  ((x*x'+y*y')+z*z')
//   cross     third_party/glm/detail/func_geometric.inl:68-79
//   length    third_party/glm/detail/func_geometric.inl:8-14   sqrt(dot(v,v))
//   normalize third_party/glm/detail/func_geometric.inl:82-90  v*(1/sqrt(dot))
//   min/max   third_party/glm/detail/func_common.inl:17-30     (y<x?y:x), (x<y?y:x)
//   mat4*vec4 third_party/glm/detail/type_mat4x4.inl:536-572   (c0*x+c1*y)+(c2*z+c3*w)
//   mat4*mat4 third_party/glm/detail/type_mat4x4.inl:630-645
//   inverse   third_party/glm/detail/func_matrix.inl:294-360 (cofactor form)
//   translate/rotate/scale third_party/glm/ext/matrix_transform.inl:18-86
#ifndef MUONX_ORACLE_OMATH_H_
#define MUONX_ORACLE_OMATH_H_

#include <cmath>

namespace oracle {

struct V3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  V3() = default;
  constexpr V3(float a, float b, float c) : x(a), y(b), z(c) {}
  explicit constexpr V3(float s) : x(s), y(s), z(s) {}
  float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3 operator+(V3 a, float s) { return V3(a.x + s, a.y + s, a.z + s); }
inline V3 operator-(float s, V3 a) { return V3(s - a.x, s - a.y, s - a.z); }
inline V3 &operator+=(V3 &a, V3 b) { a = a + b; return a; }
inline V3 &operator-=(V3 &a, V3 b) { a = a - b; return a; }
inline V3 &operator*=(V3 &a, float s) { a = a * s; return a; }
inline V3 &operator/=(V3 &a, float s) { a = a / s; return a; }

inline float fmin_glm(float x, float y) { return (y < x) ? y : x; }
inline float fmax_glm(float x, float y) { return (x < y) ? y : x; }
inline float clamp_glm(float x, float lo, float hi) { return fmin_glm(fmax_glm(x, lo), hi); }

inline float dot(V3 a, V3 b) {
  V3 t = a * b;
  return t.x + t.y + t.z;
}
inline V3 cross(V3 x, V3 y) {
  return V3(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
inline float length(V3 v) { return std::sqrt(dot(v, v)); }
inline float length2(V3 v) { return dot(v, v); }
inline V3 normalize(V3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline float comp_add(V3 v) { return (0.0f + v.x + v.y) + v.z; }
inline float comp_max(V3 v) { return fmax_glm(fmax_glm(v.x, v.y), v.z); }

struct V4 {
  float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
  V4() = default;
  constexpr V4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
  float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
};
inline V4 operator+(V4 a, V4 b) { return V4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
inline V4 operator-(V4 a, V4 b) { return V4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
inline V4 operator*(V4 a, V4 b) { return V4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
inline V4 operator*(V4 a, float s) { return V4(a.x * s, a.y * s, a.z * s, a.w * s); }

// Column-major 4x4 matrix; c[i] is column i.
struct M4 {
  V4 c[4];
  M4() : M4(0.0f) {}
  explicit M4(float d) {
    c[0] = V4(d, 0, 0, 0);
    c[1] = V4(0, d, 0, 0);
    c[2] = V4(0, 0, d, 0);
    c[3] = V4(0, 0, 0, d);
  }
  V4 &operator[](int i) { return c[i]; }
  const V4 &operator[](int i) const { return c[i]; }
};

inline V4 mul(const M4 &m, V4 v) {
  V4 a0 = m[0] * V4(v.x, v.x, v.x, v.x) + m[1] * V4(v.y, v.y, v.y, v.y);
  V4 a1 = m[2] * V4(v.z, v.z, v.z, v.z) + m[3] * V4(v.w, v.w, v.w, v.w);
  return a0 + a1;
}

inline M4 mul(const M4 &a, const M4 &b) {
  M4 r;
  for (int i = 0; i < 4; ++i) {
    r[i] = a[0] * b[i].x + a[1] * b[i].y + a[2] * b[i].z + a[3] * b[i].w;
  }
  return r;
}

inline M4 transpose(const M4 &m) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i][j] = m[j][i];
  return r;
}

inline M4 inverse(const M4 &m) {
  float c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
  float c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
  float c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
  float c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
  float c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3];
  float c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
  float c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
  float c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
  float c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  float c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
  float c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3];
  float c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
  float c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
  float c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
  float c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  float c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
  float c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1];
  float c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  V4 f0(c00, c00, c02, c03), f1(c04, c04, c06, c07), f2(c08, c08, c10, c11);
  V4 f3(c12, c12, c14, c15), f4(c16, c16, c18, c19), f5(c20, c20, c22, c23);
  V4 v0(m[1][0], m[0][0], m[0][0], m[0][0]);
  V4 v1(m[1][1], m[0][1], m[0][1], m[0][1]);
  V4 v2(m[1][2], m[0][2], m[0][2], m[0][2]);
  V4 v3(m[1][3], m[0][3], m[0][3], m[0][3]);
  V4 i0 = v1 * f0 - v2 * f1 + v3 * f2;
  V4 i1 = v0 * f0 - v2 * f3 + v3 * f4;
  V4 i2 = v0 * f1 - v1 * f3 + v3 * f5;
  V4 i3 = v0 * f2 - v1 * f4 + v2 * f5;
  V4 sa(+1, -1, +1, -1), sb(-1, +1, -1, +1);
  M4 inv;
  inv[0] = i0 * sa;
  inv[1] = i1 * sb;
  inv[2] = i2 * sa;
  inv[3] = i3 * sb;
  V4 row0(inv[0][0], inv[1][0], inv[2][0], inv[3][0]);
  V4 d0 = m[0] * row0;
  float d1 = (d0.x + d0.y) + (d0.z + d0.w);
  float one_over_det = 1.0f / d1;
  for (int i = 0; i < 4; ++i) inv[i] = inv[i] * one_over_det;
  return inv;
}

inline M4 translate(V3 v) {
  M4 m(1.0f);
  M4 r = m;
  r[3] = m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3];
  return r;
}

inline M4 rotate(float angle, V3 v) {
  M4 m(1.0f);
  float c = std::cos(angle);
  float s = std::sin(angle);
  V3 axis = normalize(v);
  V3 temp = (1.0f - c) * axis;
  M4 rot;
  rot[0][0] = c + temp[0] * axis[0];
  rot[0][1] = temp[0] * axis[1] + s * axis[2];
  rot[0][2] = temp[0] * axis[2] - s * axis[1];
  rot[1][0] = temp[1] * axis[0] - s * axis[2];
  rot[1][1] = c + temp[1] * axis[1];
  rot[1][2] = temp[1] * axis[2] + s * axis[0];
  rot[2][0] = temp[2] * axis[0] + s * axis[1];
  rot[2][1] = temp[2] * axis[1] - s * axis[0];
  rot[2][2] = c + temp[2] * axis[2];
  M4 r;
  r[0] = m[0] * rot[0][0] + m[1] * rot[0][1] + m[2] * rot[0][2];
  r[1] = m[0] * rot[1][0] + m[1] * rot[1][1] + m[2] * rot[1][2];
  r[2] = m[0] * rot[2][0] + m[1] * rot[2][1] + m[2] * rot[2][2];
  r[3] = m[3];
  return r;
}

inline M4 scale(V3 v) {
  M4 m(1.0f);
  M4 r;
  r[0] = m[0] * v.x;
  r[1] = m[1] * v.y;
  r[2] = m[2] * v.z;
  r[3] = m[3];
  return r;
}

// glm::radians: degrees * float(0.01745329251994329576923690768489).
inline float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

// muon/transform.h:8-18
inline V3 transform_position(const M4 &m, V3 p) {
  V4 r = mul(m, V4(p.x, p.y, p.z, 1.0f));
  return V3(r.x, r.y, r.z);
}
inline V3 transform_direction(const M4 &m, V3 d) {
  V4 r = mul(m, V4(d.x, d.y, d.z, 0.0f));
  return normalize(V3(r.x, r.y, r.z));
}
// muon/transform.h:26-28
inline V3 reflect(V3 d, V3 n) { return d - ((2.0f * dot(d, n)) * n); }

constexpr float kPi = 3.14159265358979323846264338327950288f;
constexpr float kOneOverPi = 0.318309886183790671537767526745028724f;
constexpr float kEpsilon = 0.0001f;  // muon/types.h:13

}  // namespace oracle

#endif
