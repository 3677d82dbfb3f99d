// Host-side 4x4 matrix builders for the parser's transform stack, with the
// operation order of the glm 0.9.9 routines muon calls, so the transforms
// uploaded to the device carry the reference's bits:
//   mat4*mat4  glm/detail/type_mat4x4.inl:630-645
//   inverse    glm/detail/func_matrix.inl:294-360 (cofactor expansion)
//   translate / rotate / scale  glm/ext/matrix_transform.inl:18-86
// Matrices are column-major float[16] (m[col*4 + row]).
#ifndef MUONX_HOST_XFORM_H_
#define MUONX_HOST_XFORM_H_

#include <array>
#include <cmath>

#include "../common/vecmath.h"

namespace muonx {
namespace host {

using M4 = std::array<float, 16>;

inline M4 Identity() {
  M4 m{};
  m[0] = m[5] = m[10] = m[15] = 1.0f;
  return m;
}

// r[:, i] = ((a0*b[0][i] + a1*b[1][i]) + a2*b[2][i]) + a3*b[3][i]
inline M4 Mul(const M4 &a, const M4 &b) {
  M4 r{};
  for (int i = 0; i < 4; ++i) {
    const float *bc = &b[i * 4];
    for (int row = 0; row < 4; ++row) {
      float acc = a[0 * 4 + row] * bc[0];
      acc = acc + a[1 * 4 + row] * bc[1];
      acc = acc + a[2 * 4 + row] * bc[2];
      acc = acc + a[3 * 4 + row] * bc[3];
      r[i * 4 + row] = acc;
    }
  }
  return r;
}

inline M4 Transpose(const M4 &m) {
  M4 r{};
  for (int c = 0; c < 4; ++c)
    for (int row = 0; row < 4; ++row) r[c * 4 + row] = m[row * 4 + c];
  return r;
}

inline M4 Inverse(const M4 &mm) {
  auto m = [&](int c, int r) { return mm[c * 4 + r]; };
  const float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
  const float c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
  const float c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
  const float c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
  const float c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
  const float c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
  const float c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
  const float c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
  const float c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
  const float c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
  const float c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
  const float c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
  const float c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
  const float c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
  const float c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
  const float c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
  const float c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
  const float c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
  const float f[6][4] = {{c00, c00, c02, c03}, {c04, c04, c06, c07}, {c08, c08, c10, c11},
                         {c12, c12, c14, c15}, {c16, c16, c18, c19}, {c20, c20, c22, c23}};
  const float v[4][4] = {{m(1, 0), m(0, 0), m(0, 0), m(0, 0)},
                         {m(1, 1), m(0, 1), m(0, 1), m(0, 1)},
                         {m(1, 2), m(0, 2), m(0, 2), m(0, 2)},
                         {m(1, 3), m(0, 3), m(0, 3), m(0, 3)}};
  // Column k of the adjugate: (va*fa - vb*fb) + vc*fc with the sign pattern
  // (+,-,+,-) for even columns and (-,+,-,+) for odd ones.
  const int terms[4][6] = {{1, 0, 2, 1, 3, 2}, {0, 0, 2, 3, 3, 4}, {0, 1, 1, 3, 3, 5}, {0, 2, 1, 4, 2, 5}};
  M4 inv{};
  for (int k = 0; k < 4; ++k) {
    const int *t = terms[k];
    for (int r = 0; r < 4; ++r) {
      float val = (v[t[0]][r] * f[t[1]][r] - v[t[2]][r] * f[t[3]][r]) + v[t[4]][r] * f[t[5]][r];
      float sign = ((k + r) % 2 == 0) ? 1.0f : -1.0f;
      inv[k * 4 + r] = val * sign;
    }
  }
  const float d0 = mm[0] * inv[0 * 4 + 0], d1 = mm[1] * inv[1 * 4 + 0];
  const float d2 = mm[2] * inv[2 * 4 + 0], d3 = mm[3] * inv[3 * 4 + 0];
  const float det = (d0 + d1) + (d2 + d3);
  const float one_over_det = 1.0f / det;
  for (float &x : inv) x = x * one_over_det;
  return inv;
}

inline M4 Translate(float x, float y, float z) {
  M4 m = Identity();
  M4 r = m;
  for (int row = 0; row < 4; ++row) {
    float acc = m[0 * 4 + row] * x;
    acc = acc + m[1 * 4 + row] * y;
    acc = acc + m[2 * 4 + row] * z;
    acc = acc + m[3 * 4 + row];
    r[12 + row] = acc;
  }
  return r;
}

inline M4 Rotate(float angle, float ax, float ay, float az) {
  const M4 m = Identity();
  const float c = std::cos(angle);
  const float s = std::sin(angle);
  V3 axis = normalize(V3{ax, ay, az});
  V3 temp = (1.0f - c) * axis;
  float rot[3][3];
  rot[0][0] = c + temp.x * axis.x;
  rot[0][1] = temp.x * axis.y + s * axis.z;
  rot[0][2] = temp.x * axis.z - s * axis.y;
  rot[1][0] = temp.y * axis.x - s * axis.z;
  rot[1][1] = c + temp.y * axis.y;
  rot[1][2] = temp.y * axis.z + s * axis.x;
  rot[2][0] = temp.z * axis.x + s * axis.y;
  rot[2][1] = temp.z * axis.y - s * axis.x;
  rot[2][2] = c + temp.z * axis.z;
  M4 r{};
  for (int i = 0; i < 3; ++i)
    for (int row = 0; row < 4; ++row) {
      float acc = m[0 * 4 + row] * rot[i][0];
      acc = acc + m[1 * 4 + row] * rot[i][1];
      acc = acc + m[2 * 4 + row] * rot[i][2];
      r[i * 4 + row] = acc;
    }
  for (int row = 0; row < 4; ++row) r[12 + row] = m[12 + row];
  return r;
}

inline M4 Scale(float x, float y, float z) {
  const M4 m = Identity();
  M4 r{};
  const float s[3] = {x, y, z};
  for (int i = 0; i < 3; ++i)
    for (int row = 0; row < 4; ++row) r[i * 4 + row] = m[i * 4 + row] * s[i];
  for (int row = 0; row < 4; ++row) r[12 + row] = m[12 + row];
  return r;
}

// glm::radians (glm/detail/func_trigonometric.inl:9-14)
inline float Radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

}  // namespace host
}  // namespace muonx

#endif
