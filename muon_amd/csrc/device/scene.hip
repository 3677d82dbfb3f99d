// Scene upload: converts the host descriptor (parse-order primitives + the
// binary SAH BVH in pre-order) into the device layout of dev_scene.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../common/vecmath.h"
#include "bvh_device.h"
#include "runtime.h"

namespace muonx {
namespace {

float BitsAsFloat(uint32_t u) {
  float f;
  std::memcpy(&f, &u, sizeof(f));
  return f;
}

// Reflectiveness of Phong (muon/materials.cc:113-123) and GGX (:270-293).
float Reflectiveness(const float *m) {
  const float s = ((0.0f + m[6]) + m[7] + m[8]) / 3.0f;
  const float d = ((0.0f + m[3]) + m[4] + m[5]) / 3.0f;
  const float denom = d + s;
  if (int(m[14]) == MUONX_BRDF_GGX) {
    if (denom == 0.0f) return 1.0f;
    return gmax(s / denom, 0.25f);
  }
  return denom > 0.0f ? s / denom : 0.0f;
}

// Wide trees: a binary subtree of at most this many primitives becomes one
// leaf. Measured on the level-7 icosphere (2^24 incoherent rays): 1 -> 4159,
// 2 -> 4451, 3 -> 4355, 4 -> 4199, 8 -> 3800 Mrays/s.
constexpr uint32_t kLeafMerge = 2;

float SurfaceArea(const float *b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Primitive range of every binary subtree (pre-order: children come after
// their parent): first primitive and count; count 0 when the subtree's leaves
// are not one contiguous range.
struct Span {
  uint32_t first, count;
};
std::vector<Span> SubtreeSpans(const muonx_scene_desc *d) {
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
std::vector<float4> CollapseWide(const muonx_scene_desc *d, const std::vector<Span> &span, uint32_t merge,
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

  std::vector<float4> out;
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
    float4 *q = &out[size_t(it.wide) * 8];
    for (int a = 0; a < 3; ++a) {
      q[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
      q[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    q[6] = make_float4(BitsAsFloat(ref[0]), BitsAsFloat(ref[1]), BitsAsFloat(ref[2]), BitsAsFloat(ref[3]));
    q[7] = make_float4(BitsAsFloat(uint32_t(n)), 0.0f, 0.0f, 0.0f);
  }
  return out;
}

uint32_t FloatBits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, sizeof(u));
  return u;
}

// A quantized plane origin + q * 2^e as a real number. The double sum is
// exact: with q < 256, the terms span at most 53 bits whenever the node's
// extent is nonzero (the extent is then at least one float ulp of the origin,
// so 2^e >= ulp(origin) / 256), and q is 0 when the extent is 0.
double PlaneValue(float origin, uint32_t q, int e) { return double(origin) + std::ldexp(double(q), e); }

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
std::vector<uint4> QuantizeWide(const std::vector<float4> &w, uint32_t empty_ref) {
  const size_t nn = w.size() / 8;
  std::vector<uint4> out(nn * 4);
  for (size_t i = 0; i < nn; ++i) {
    const float4 *q = &w[i * 8];
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
    uint4 *o = &out[i * 4];
    o[0] = make_uint4(FloatBits(origin[0]), FloatBits(origin[1]), FloatBits(origin[2]),
                      uint32_t(ex[0] + 127) | (uint32_t(ex[1] + 127) << 8) | (uint32_t(ex[2] + 127) << 16) | (n << 24));
    o[1] = make_uint4(pack(qlo[0]), pack(qlo[1]), pack(qlo[2]), pack(qhi[0]));
    o[2] = make_uint4(pack(qhi[1]), pack(qhi[2]), 0u, 0u);
    o[3] = make_uint4(refs[0], refs[1], refs[2], refs[3]);
  }
  return out;
}

template <typename T>
int Upload(DevBuf &buf, const std::vector<T> &host, uint64_t &total) {
  MX_HIP(buf.Alloc(host.size() * sizeof(T)));
  if (!host.empty()) MX_HIP(hipMemcpy(buf.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  total += buf.bytes;
  return MUONX_OK;
}

}  // namespace

namespace {
std::mutex g_scene_mu;
std::unordered_set<uint64_t> g_live_scenes;
uint64_t g_next_scene_id = 1;
}  // namespace

uint64_t RegisterScene() {
  std::lock_guard<std::mutex> lock(g_scene_mu);
  const uint64_t id = g_next_scene_id++;
  g_live_scenes.insert(id);
  return id;
}

void UnregisterScene(uint64_t id) {
  std::lock_guard<std::mutex> lock(g_scene_mu);
  g_live_scenes.erase(id);
}

bool SceneLive(uint64_t id) {
  std::lock_guard<std::mutex> lock(g_scene_mu);
  return g_live_scenes.count(id) != 0;
}

}  // namespace muonx

using namespace muonx;

extern "C" int muonx_scene_create(muonx_ctx *ctx, const muonx_scene_desc *din, muonx_scene **out) {
  if (!ctx || !din || !out) {
    SetError("muonx_scene_create: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = nullptr;
  const uint32_t np = din->num_prims;
  if (np == 0) {
    SetError("muonx_scene_create: scene has no primitives");  // the reference dereferences a null BVH root
    return MUONX_ERR_INVALID;
  }
  if (!din->prim_type || !din->prim_geom || !din->prim_transform || !din->prim_material || !din->transforms ||
      !din->materials) {
    SetError("muonx_scene_create: descriptor is missing arrays");
    return MUONX_ERR_INVALID;
  }
  for (uint32_t i = 0; i < np; ++i) {
    if (din->prim_transform[i] < 0 || uint32_t(din->prim_transform[i]) >= din->num_transforms) {
      SetError("muonx_scene_create: invalid transform index");
      return MUONX_ERR_INVALID;
    }
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  MX_HIP(hipSetDevice(ctx->device));
  // a descriptor without a tree: build one on the device
  DeviceBvh built;
  muonx_scene_desc dd = *din;
  const bool device_tree = din->num_nodes == 0;
  if (device_tree) {
    const int brc = BuildDeviceBvh(ctx, din, &built);
    if (brc != MUONX_OK) return brc;
    dd.num_nodes = uint32_t(built.left.size());
    dd.node_bounds = built.bounds.data();
    dd.node_left = built.left.data();
    dd.node_right = built.right.data();
    dd.node_axis = built.axis.data();
    dd.node_start = built.start.data();
    dd.node_count = built.count.data();
    dd.leaf_prims = built.leaf.data();
    dd.acceleration = 1;
  }
  const muonx_scene_desc *d = &dd;
  if (!d->node_bounds || !d->node_left || !d->node_right || !d->node_axis || !d->node_start || !d->node_count ||
      !d->leaf_prims) {
    SetError("muonx_scene_create: descriptor is missing arrays");
    return MUONX_ERR_INVALID;
  }

  // ---- primitives in leaf order (+ one sentinel that never hits)
  std::vector<float4> prims(size_t(np + 1) * 3, make_float4(0, 0, 0, 0));
  std::vector<int32_t> pid(np + 1, -1), pmat(np + 1, 0), plight(np + 1, -1);
  std::vector<float> pvn;
  std::vector<uint8_t> puvn;
  bool any_vn = false;
  if (d->prim_use_vnormal && d->prim_vnormal)
    for (uint32_t i = 0; i < np; ++i) any_vn |= d->prim_use_vnormal[i] != 0;
  if (any_vn) {
    pvn.assign(size_t(np + 1) * 9, 0.0f);
    puvn.assign(np + 1, 0);
  }
  for (uint32_t k = 0; k < np; ++k) {
    const int32_t id = d->leaf_prims[k];
    if (id < 0 || uint32_t(id) >= np) {
      SetError("muonx_scene_create: leaf_prims holds an invalid id");
      return MUONX_ERR_INVALID;
    }
    const float *g = d->prim_geom + size_t(id) * 12;
    const int32_t tid = d->prim_transform[id];
    if (tid < 0 || uint32_t(tid) >= d->num_transforms) {
      SetError("muonx_scene_create: invalid transform index");
      return MUONX_ERR_INVALID;
    }
    const int32_t mid = d->prim_material[id];
    if (mid < 0 || uint32_t(mid) >= d->num_materials) {
      SetError("muonx_scene_create: invalid material index");
      return MUONX_ERR_INVALID;
    }
    uint32_t tw = uint32_t(tid);
    float4 *q = &prims[size_t(k) * 3];
    if (d->prim_type[id] == MUONX_PRIM_SPHERE) {
      tw |= kSphereBit;
      q[0] = make_float4(g[0], g[1], g[2], BitsAsFloat(tw));
      q[1] = make_float4(g[3], 0, 0, 0);
    } else {
      q[0] = make_float4(g[0], g[1], g[2], BitsAsFloat(tw));
      q[1] = make_float4(g[3], g[4], g[5], 0);
      q[2] = make_float4(g[6], g[7], g[8], 0);
    }
    pid[k] = id;
    pmat[k] = mid;
    plight[k] = d->prim_light ? d->prim_light[id] : -1;
    if (any_vn) {
      puvn[k] = d->prim_use_vnormal[id];
      std::memcpy(&pvn[size_t(k) * 9], d->prim_vnormal + size_t(id) * 9, 9 * sizeof(float));
    }
  }
  // Sentinel: transform 0, all-zero triangle (zero normal: never hit), end of leaf.
  prims[size_t(np) * 3 + 1].w = BitsAsFloat(1u);

  // ---- BVH: child boxes in parent; internal binary nodes in pre-order.
  const uint32_t nn = d->num_nodes;
  std::vector<int32_t> dev_index(nn, -1);
  uint32_t n_internal = 0;
  uint32_t depth = 1;
  {
    // depth by a pre-order walk
    std::vector<uint32_t> dep(nn, 1);
    for (uint32_t i = 0; i < nn; ++i) {
      if (d->node_count[i] == 0) {
        const int32_t l = d->node_left[i], r = d->node_right[i];
        if (l <= int32_t(i) || r <= int32_t(i) || uint32_t(l) >= nn || uint32_t(r) >= nn) {
          SetError("muonx_scene_create: BVH is not in pre-order");
          return MUONX_ERR_INVALID;
        }
        dev_index[i] = int32_t(n_internal++);
        dep[l] = dep[r] = dep[i] + 1;
      } else {
        const int64_t s = d->node_start[i], c = d->node_count[i];
        if (s < 0 || c <= 0 || s + c > int64_t(np)) {
          SetError("muonx_scene_create: leaf range out of bounds");
          return MUONX_ERR_INVALID;
        }
        prims[size_t(s + c - 1) * 3 + 1].w = BitsAsFloat(1u);  // last primitive of the leaf
      }
      depth = std::max(depth, dep[i]);
    }
  }
  const bool root_leaf = d->node_count[0] != 0;
  const uint32_t n_dev = root_leaf ? 1 : n_internal;
  std::vector<float4> nodes(size_t(n_dev) * 4);
  auto child_ref = [&](int32_t c) -> uint32_t {
    if (d->node_count[c] != 0) return kLeafBit | uint32_t(d->node_start[c]);
    return uint32_t(dev_index[c]);
  };
  auto child_meta = [&](int32_t c) -> uint32_t { return d->node_count[c] != 0 ? kLeafBit : 0u; };
  auto fill = [&](uint32_t di, const float *b0, const float *b1, uint32_t r0, uint32_t r1, uint32_t m0, uint32_t m1,
                  uint32_t axis) {
    float4 *q = &nodes[size_t(di) * 4];
    q[0] = make_float4(b0[0], b0[3], b0[1], b0[4]);
    q[1] = make_float4(b1[0], b1[3], b1[1], b1[4]);
    q[2] = make_float4(b0[2], b0[5], b1[2], b1[5]);
    q[3] = make_float4(BitsAsFloat(r0), BitsAsFloat(r1), BitsAsFloat(m0 | (axis << 29)),
                       BitsAsFloat(m1));
  };
  if (root_leaf) {
    const float *b = d->node_bounds;
    fill(0, b, b, kLeafBit | uint32_t(d->node_start[0]), kLeafBit | np, kLeafBit, kLeafBit, 0);
    depth = 2;
  } else {
    for (uint32_t i = 0; i < nn; ++i) {
      if (d->node_count[i] != 0) continue;
      const int32_t l = d->node_left[i], r = d->node_right[i];
      fill(uint32_t(dev_index[i]), d->node_bounds + size_t(l) * 6, d->node_bounds + size_t(r) * 6, child_ref(l),
           child_ref(r), child_meta(l), child_meta(r), uint32_t(d->node_axis[i]) & 3u);
    }
  }

  // ---- wide BVH (fast traversal mode)
  // Leaf references carry counts (and small subtrees merge into one leaf)
  // when every count and first primitive fits the reference.
  uint32_t max_leaf = 0;
  for (uint32_t i = 0; i < nn; ++i) max_leaf = std::max<uint32_t>(max_leaf, uint32_t(std::max(0, d->node_count[i])));
  const bool leaf_counts = np < (1u << kLeafCountShift) && max_leaf <= kLeafMaxCount;
  uint32_t merge = 0;
  if (leaf_counts) {
    merge = std::min(kLeafMerge, kLeafMaxCount);
  }
  std::vector<float4> wnodes;
  uint32_t wstack = 1;
  if (root_leaf) {
    wnodes.assign(8, make_float4(0, 0, 0, 0));
    const float *b = d->node_bounds;
    wnodes[0] = make_float4(b[0], 0, 0, 0);
    wnodes[1] = make_float4(b[3], 0, 0, 0);
    wnodes[2] = make_float4(b[1], 0, 0, 0);
    wnodes[3] = make_float4(b[4], 0, 0, 0);
    wnodes[4] = make_float4(b[2], 0, 0, 0);
    wnodes[5] = make_float4(b[5], 0, 0, 0);
    const uint32_t ref = leaf_counts ? kLeafBit | ((uint32_t(d->node_count[0]) - 1u) << kLeafCountShift) |
                                           uint32_t(d->node_start[0])
                                     : kLeafBit | uint32_t(d->node_start[0]);
    wnodes[6] = make_float4(BitsAsFloat(ref), 0, 0, 0);
    wnodes[7] = make_float4(BitsAsFloat(1u), 0, 0, 0);
  } else {
    wnodes = CollapseWide(d, SubtreeSpans(d), merge, &wstack);
  }

  std::vector<float> mats(size_t(d->num_materials) * kMatStride);
  for (uint32_t i = 0; i < d->num_materials; ++i) {
    std::memcpy(&mats[size_t(i) * 16], d->materials + size_t(i) * 16, 16 * sizeof(float));
    mats[size_t(i) * 16 + 15] = Reflectiveness(&mats[size_t(i) * 16]);
  }
  std::vector<float> lights(size_t(std::max(1u, d->num_lights)) * kLightStride, 0.0f);
  uint32_t nquad = 0;
  for (uint32_t i = 0; i < d->num_lights; ++i) {
    std::memcpy(&lights[size_t(i) * 16], d->lights + size_t(i) * 16, 16 * sizeof(float));
    if (int(lights[size_t(i) * 16]) == MUONX_LIGHT_QUAD) {
      ++nquad;
      // slot 15: |e0 x e1| (muon/integration.cc:383-384)
      const float *L = &lights[size_t(i) * 16];
      V3 e0{L[7], L[8], L[9]}, e1{L[10], L[11], L[12]};
      lights[size_t(i) * 16 + 15] = length(cross(e0, e1));
    }
  }
  std::vector<float> xf(d->transforms, d->transforms + size_t(d->num_transforms) * 48);

  auto *s = new muonx_scene();
  s->ctx = ctx;
  uint64_t total = 0;
  int rc = MUONX_OK;
  // empty slots point at the sentinel primitive np (a one-primitive leaf in
  // either leaf encoding: count field 0, and the sentinel carries the end flag)
  const std::vector<uint4> qnodes = QuantizeWide(wnodes, kLeafBit | np);
  std::vector<uint32_t> lprims;
  for (uint32_t k = 0; k < np; ++k)
    if (plight[k] >= 0) lprims.push_back(k);
  const bool cull_mis = lprims.size() <= kMaxLightPrims;
  if (!cull_mis || lprims.empty()) lprims.assign(1, 0u);  // keep the buffer non-empty
  if ((rc = Upload(s->nodes, nodes, total)) || (rc = Upload(s->wnodes, wnodes, total)) ||
      (rc = Upload(s->qnodes, qnodes, total)) ||
      (rc = Upload(s->prims, prims, total)) ||
      (rc = Upload(s->xforms, xf, total)) || (rc = Upload(s->prim_id, pid, total)) ||
      (rc = Upload(s->prim_mat, pmat, total)) || (rc = Upload(s->prim_light, plight, total)) ||
      (rc = Upload(s->materials, mats, total)) || (rc = Upload(s->lights, lights, total)) ||
      (rc = Upload(s->light_prims, lprims, total))) {
    delete s;
    return rc;
  }
  if (any_vn) {
    if ((rc = Upload(s->prim_vn, pvn, total)) || (rc = Upload(s->prim_use_vn, puvn, total))) {
      delete s;
      return rc;
    }
  }
  DevSceneView &v = s->view;
  v.nodes = s->nodes.as<float4>();
  v.wnodes = s->wnodes.as<float4>();
  v.qnodes = s->qnodes.as<uint4>();
  v.num_wnodes = uint32_t(wnodes.size() / 8);
  v.wstack = wstack;
  v.leaf_counts = leaf_counts ? 1u : 0u;
  v.root_leaf = root_leaf ? 1u : 0u;
  v.linear = (d->acceleration == 0 && root_leaf) ? 1u : 0u;
  v.light_prims = s->light_prims.as<uint32_t>();
  v.cull_mis = cull_mis ? 1u : 0u;
  {
    uint32_t ne = 0;
    for (uint32_t k = 0; k < np; ++k) ne += plight[k] >= 0 ? 1u : 0u;
    v.num_light_prims = cull_mis ? ne : 0u;
  }
  v.prims = s->prims.as<float4>();
  v.xforms = s->xforms.as<float>();
  v.prim_id = s->prim_id.as<int32_t>();
  v.prim_mat = s->prim_mat.as<int32_t>();
  v.prim_light = s->prim_light.as<int32_t>();
  v.prim_vn = any_vn ? s->prim_vn.as<float>() : nullptr;
  v.prim_use_vn = any_vn ? s->prim_use_vn.as<uint8_t>() : nullptr;
  v.materials = s->materials.as<float>();
  v.lights = s->lights.as<float>();
  v.num_prims = np;
  v.num_nodes = n_dev;
  v.num_lights = d->num_lights;
  v.num_quad_lights = nquad;
  v.bvh_depth = depth;
  s->bvh_depth = depth;
  s->device_bytes = total;
  s->reference_tree = !device_tree;
  s->build_ms = device_tree ? built.total_ms : 0.0f;
  std::memcpy(s->bounds, d->node_bounds, sizeof(s->bounds));  // the root's box
  s->host_lights = std::move(lights);
  s->host_materials = std::move(mats);
  s->id = RegisterScene();
  *out = s;
  return MUONX_OK;
}

extern "C" void muonx_scene_destroy(muonx_scene *s) {
  if (!s) return;
  // server jobs still in flight keep using the scene's device memory
  while (s->inflight.load() > 0) std::this_thread::sleep_for(std::chrono::milliseconds(1));
  UnregisterScene(s->id);
  if (s->ctx) (void)hipSetDevice(s->ctx->device);
  delete s;
}

extern "C" int muonx_scene_get_info(const muonx_scene *s, muonx_scene_info *out) {
  if (!s || !out) {
    SetError("muonx_scene_get_info: null argument");
    return MUONX_ERR_INVALID;
  }
  out->device_bytes = s->device_bytes;
  out->num_prims = s->view.num_prims;
  out->num_nodes = s->view.num_nodes;
  out->bvh_depth = s->bvh_depth;
  out->num_lights = s->view.num_lights;
  out->num_quad_lights = s->view.num_quad_lights;
  out->device_built = s->reference_tree ? 0u : 1u;
  out->build_ms = s->build_ms;
  return MUONX_OK;
}
