// Host-side construction of the fast traversal mode's trees from the
// descriptor's binary BVH: the SAH-optimal collapse into 4-wide nodes and
// their 8-bit quantization (64 B nodes). No HIP types: the scene upload
// (scene.hip) and the CPU checks (tests/cpu/wide_bvh_check.cpp) include it.
#ifndef MUONX_DEVICE_WIDE_BVH_H_
#define MUONX_DEVICE_WIDE_BVH_H_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../../include/muonx.h"
#include "bvh_layout.h"

namespace muonx {

// 16-byte records with the layout of HIP's float4 / uint4 (the device reads
// the uploaded arrays as those types).
struct F4 {
  float x, y, z, w;
};
struct U4 {
  uint32_t x, y, z, w;
};

inline float BitsAsFloat(uint32_t u) {
  float f;
  std::memcpy(&f, &u, sizeof(f));
  return f;
}

// Wide trees: a binary subtree of at most this many primitives becomes one
// leaf. Measured on the level-7 icosphere (2^24 incoherent rays): 1 -> 4159,
// 2 -> 4451, 3 -> 4355, 4 -> 4199, 8 -> 3800 Mrays/s.
constexpr uint32_t kLeafMerge = 2;

inline float SurfaceArea(const float *b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Primitive range of every binary subtree (pre-order: children come after
// their parent): first primitive and count; count 0 when the subtree's leaves
// are not one contiguous range.
struct Span {
  uint32_t first, count;
};
inline std::vector<Span> SubtreeSpans(const muonx_scene_desc *d) {
  std::vector<Span> sp(d->num_nodes, Span{0, 0});
  for (int64_t i = int64_t(d->num_nodes) - 1; i >= 0; --i) {
    if (d->node_count[i] != 0) {
      sp[size_t(i)] = Span{uint32_t(d->node_start[i]), uint32_t(d->node_count[i])};
      continue;
    }
    const Span a = sp[size_t(d->node_left[i])], b = sp[size_t(d->node_right[i])];
    if (a.count && b.count && a.first + a.count == b.first) sp[size_t(i)] = Span{a.first, a.count + b.count};
    else if (a.count && b.count && b.first + b.count == a.first) sp[size_t(i)] = Span{b.first, a.count + b.count};
  }
  return sp;
}

// SAH weights of the collapse, per unit of surface area: one wide-node step
// and one primitive test, in the ratio of their instruction counts.
constexpr float kCollapseNodeCost = 1.0f, kCollapsePrimCost = 0.75f;

// Collapses the binary tree (pre-order arrays of the descriptor) into 4-ary
// nodes of least SAH cost over the binary tree's cuts (the dynamic program
// of Ylitie, Karras and Laine, "Efficient incoherent ray traversal on GPUs
// through compressed wide BVHs", HPG 2017, without their leaf formation):
// S(n, k) is the least cost of the subtree at n as at most k children of
// one wide node, either one child (a leaf, or a wide node of cost
// SA(n) * node cost + S(left, i) + S(right, 4 - i)) or the subtrees of its two
// children sharing the k places; a leaf costs SA * primitives * prim cost.
// Leaves are the binary leaves and, when `merge` is not 0, binary subtrees
// of at most `merge` primitives (one contiguous range), which are never split
// and whose references then carry their count (kLeafCountShift); with
// `merge` 0 leaf references are leaf bit | first primitive. Children keep the
// binary tree's left-to-right order. Returns the nodes (8 float4 each,
// pre-order) and in `max_stack` a bound on the stack entries a traversal can
// hold: 3 per wide node along the deepest root-to-leaf path (a step pushes at
// most 3 children; the quantized nodes' empty slots, which the traversal does
// not skip by count, can add pushes on nodes of fewer than 4 children).
inline std::vector<F4> CollapseWide(const muonx_scene_desc *d, const std::vector<Span> &span, uint32_t merge,
                                 uint32_t *max_stack) {
  auto is_leaf = [&](int32_t c) {
    return d->node_count[c] != 0 || (merge && span[size_t(c)].count && span[size_t(c)].count <= merge);
  };
  auto leaf_ref = [&](int32_t c) {
    if (!merge) return kLeafBit | uint32_t(d->node_start[c]);
    const Span &q = span[size_t(c)];
    return kLeafBit | ((q.count - 1u) << kLeafCountShift) | q.first;
  };
  // ---- costs, children after parents (pre-order), so bottom-up in reverse
  const size_t nn = d->num_nodes;
  std::vector<float> cost(nn * 4);  // cost[4 n + k - 1] = S(n, k)
  auto S = [&](int32_t n, int k) -> float & { return cost[size_t(n) * 4 + size_t(k - 1)]; };
  // best (least cost, first on ties) share i of k places for the left child
  auto best_split = [&](int32_t l, int32_t r, int k, float *c) {
    int bi = 1;
    float bc = S(l, 1) + S(r, k - 1);
    for (int i = 2; i < k; ++i) {
      const float ci = S(l, i) + S(r, k - i);
      if (ci < bc) {
        bc = ci;
        bi = i;
      }
    }
    *c = bc;
    return bi;
  };
  for (int64_t n = int64_t(nn) - 1; n >= 0; --n) {
    const int32_t c = int32_t(n);
    if (is_leaf(c)) {
      const uint32_t cnt = d->node_count[c] != 0 ? uint32_t(d->node_count[c]) : span[size_t(c)].count;
      const float lc = SurfaceArea(d->node_bounds + size_t(c) * 6) * float(cnt) * kCollapsePrimCost;
      for (int k = 1; k <= 4; ++k) S(c, k) = lc;
      continue;
    }
    const int32_t l = d->node_left[c], r = d->node_right[c];
    float wide;
    best_split(l, r, 4, &wide);
    S(c, 1) = SurfaceArea(d->node_bounds + size_t(c) * 6) * kCollapseNodeCost + wide;
    for (int k = 2; k <= 4; ++k) {
      float sc;
      best_split(l, r, k, &sc);
      S(c, k) = std::min(S(c, 1), sc);
    }
  }
  // the children of a wide node: the subtree at n in at most k places
  auto gather = [&](auto &&self, int32_t n, int k, std::vector<int32_t> &kids) -> void {
    if (k == 1 || is_leaf(n)) {
      kids.push_back(n);
      return;
    }
    float sc;
    const int i = best_split(d->node_left[n], d->node_right[n], k, &sc);
    if (!(sc < S(n, 1))) {
      kids.push_back(n);
      return;
    }
    self(self, d->node_left[n], i, kids);
    self(self, d->node_right[n], k - i, kids);
  };

  std::vector<F4> out;
  struct Item {
    int32_t bin;     // binary internal node
    uint32_t wide;   // its wide node index
    uint32_t pushes; // stack entries held above it
  };
  std::vector<Item> todo;
  out.resize(8);
  todo.push_back({0, 0, 0});
  *max_stack = 1;
  std::vector<int32_t> kids;
  while (!todo.empty()) {
    const Item it = todo.back();
    todo.pop_back();
    kids.clear();
    float unused;
    const int i = best_split(d->node_left[it.bin], d->node_right[it.bin], 4, &unused);
    gather(gather, d->node_left[it.bin], i, kids);
    gather(gather, d->node_right[it.bin], 4 - i, kids);
    const int n = int(kids.size());
    int32_t c[4] = {-1, -1, -1, -1};
    for (int k = 0; k < n; ++k) c[k] = kids[size_t(k)];
    float lo[3][4], hi[3][4];
    uint32_t ref[4];
    const uint32_t pushes = it.pushes + 3u;
    for (int k = 0; k < 4; ++k) {
      if (k >= n) {
        for (int a = 0; a < 3; ++a) lo[a][k] = hi[a][k] = 0.0f;
        ref[k] = 0;
        continue;
      }
      const float *b = d->node_bounds + size_t(c[k]) * 6;
      for (int a = 0; a < 3; ++a) {
        lo[a][k] = b[a];
        hi[a][k] = b[a + 3];
      }
      if (is_leaf(c[k])) {
        ref[k] = leaf_ref(c[k]);
        *max_stack = std::max(*max_stack, pushes);
      } else {
        const uint32_t w = uint32_t(out.size() / 8);
        out.resize(out.size() + 8);
        ref[k] = w;
        todo.push_back({c[k], w, pushes});
      }
    }
    F4 *q = &out[size_t(it.wide) * 8];
    for (int a = 0; a < 3; ++a) {
      q[2 * a] = F4{lo[a][0], lo[a][1], lo[a][2], lo[a][3]};
      q[2 * a + 1] = F4{hi[a][0], hi[a][1], hi[a][2], hi[a][3]};
    }
    q[6] = F4{BitsAsFloat(ref[0]), BitsAsFloat(ref[1]), BitsAsFloat(ref[2]), BitsAsFloat(ref[3])};
    q[7] = F4{BitsAsFloat(uint32_t(n)), 0.0f, 0.0f, 0.0f};
  }
  return out;
}

inline uint32_t FloatBits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, sizeof(u));
  return u;
}

// A quantized plane origin + q * 2^e as a real number. The double sum is
// exact: with q < 256, the terms span at most 53 bits whenever the node's
// extent is nonzero (the extent is then at least one float ulp of the origin,
// so 2^e >= ulp(origin) / 256), and q is 0 when the extent is 0.
inline double PlaneValue(float origin, uint32_t q, int e) { return double(origin) + std::ldexp(double(q), e); }

// Quantizes the 4-wide nodes (8 float4 each) to 64 B nodes (4 uint4 each):
//   n0 = (origin.xyz, e.x + 127 | (e.y + 127) << 8 | (e.z + 127) << 16 | count << 24)
//   n1 = (lo.x[4], lo.y[4], lo.z[4], hi.x[4]) as bytes (child k in byte k)
//   n2 = (hi.y[4], hi.z[4], 0, 0)
//   n3 = refs[4]
// with child plane = origin + q * 2^e. Every quantized box contains the
// child's float box in exact arithmetic (so also after rounding the plane to
// float, which is monotonic); the device's slab test is conservative against
// these real-valued planes, so a traversal over the quantized nodes visits a
// superset of the boxes. An empty slot (k >= count) gets the inverted box
// lo = 255, hi = 0 on every axis, which the slab test rejects unless the
// node is smaller than about 2^-18 of its distance from the ray origin, and
// the reference `empty_ref`, a leaf of the sentinel primitive that never
// hits, so that a ray which does pass it only tests that primitive; the
// traversal therefore needs no per-child count check.
inline std::vector<U4> QuantizeWide(const std::vector<F4> &w, uint32_t empty_ref) {
  const size_t nn = w.size() / 8;
  std::vector<U4> out(nn * 4);
  for (size_t i = 0; i < nn; ++i) {
    const F4 *q = &w[i * 8];
    const uint32_t n = FloatBits(q[7].x) <= 4 ? FloatBits(q[7].x) : 4;
    float lo[3][4], hi[3][4];
    for (int k = 0; k < 4; ++k) {
      const float *l0 = &q[0].x, *h0 = &q[1].x, *l1 = &q[2].x, *h1 = &q[3].x, *l2 = &q[4].x, *h2 = &q[5].x;
      lo[0][k] = l0[k];
      hi[0][k] = h0[k];
      lo[1][k] = l1[k];
      hi[1][k] = h1[k];
      lo[2][k] = l2[k];
      hi[2][k] = h2[k];
    }
    float origin[3];
    int ex[3];
    uint32_t qlo[3][4] = {}, qhi[3][4] = {};
    for (int a = 0; a < 3; ++a) {
      float mn = lo[a][0], mx = hi[a][0];
      for (uint32_t k = 1; k < n; ++k) {
        mn = std::min(mn, lo[a][k]);
        mx = std::max(mx, hi[a][k]);
      }
      origin[a] = mn;
      const double extent = double(mx) - double(mn);
      int e = -126;
      if (extent > 0.0) {
        int ee;
        std::frexp(extent / 255.0, &ee);  // extent / 255 < 2^ee
        e = std::max(-126, ee);
      }
      for (;; ++e) {
        bool ok = e <= 127;
        for (uint32_t k = 0; ok && k < n; ++k) {
          const double sc = std::ldexp(1.0, e);
          int64_t l = int64_t(std::floor((double(lo[a][k]) - double(mn)) / sc));
          int64_t h = int64_t(std::ceil((double(hi[a][k]) - double(mn)) / sc));
          l = std::max<int64_t>(0, std::min<int64_t>(255, l));
          while (l > 0 && PlaneValue(mn, uint32_t(l), e) > double(lo[a][k])) --l;
          h = std::max<int64_t>(0, h);
          while (h <= 255 && PlaneValue(mn, uint32_t(h), e) < double(hi[a][k])) ++h;
          if (h > 255) {
            ok = false;
            break;
          }
          qlo[a][k] = uint32_t(l);
          qhi[a][k] = uint32_t(h);
        }
        if (ok || e >= 127) break;
      }
      ex[a] = e;
      for (uint32_t k = n; k < 4; ++k) {
        qlo[a][k] = 255u;
        qhi[a][k] = 0u;
      }
    }
    uint32_t refs[4] = {FloatBits(q[6].x), FloatBits(q[6].y), FloatBits(q[6].z), FloatBits(q[6].w)};
    for (uint32_t k = n; k < 4; ++k) refs[k] = empty_ref;
    auto pack = [](const uint32_t v[4]) { return v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24); };
    U4 *o = &out[i * 4];
    o[0] = U4{FloatBits(origin[0]), FloatBits(origin[1]), FloatBits(origin[2]),
                      uint32_t(ex[0] + 127) | (uint32_t(ex[1] + 127) << 8) | (uint32_t(ex[2] + 127) << 16) | (n << 24)};
    o[1] = U4{pack(qlo[0]), pack(qlo[1]), pack(qlo[2]), pack(qhi[0])};
    o[2] = U4{pack(qhi[1]), pack(qhi[2]), 0u, 0u};
    o[3] = U4{refs[0], refs[1], refs[2], refs[3]};
  }
  return out;
}


}  // namespace muonx

#endif
