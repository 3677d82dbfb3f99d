// CPU checks of the host-side wide-tree construction (muon_amd/csrc/device/
// wide_bvh.h): the SAH-optimal collapse of a binary BVH into 4-wide nodes and
// the 8-bit quantization of their child boxes. Built with the host compiler
// under AddressSanitizer / UBSan by tests/test_wide_bvh.py; prints "ok" or
// the first failed check and exits non-zero.
//
// On random binary trees (pre-order, as the scene descriptor holds them,
// including flat, coincident, huge and tiny boxes):
//  * every primitive lies in exactly one leaf of the wide tree, every wide
//    node is reached once and has 2..4 children;
//  * each child box is the binary node's box, and contains its subtree;
//  * the stack bound is 3 per wide node along the deepest path;
//  * the collapse's SAH cost is never above the greedy collapse's (expand the
//    child of largest surface area until four), and equals the least cost
//    over all collapses, enumerated one by one, on small trees;
//  * every quantized child plane, as the exact real origin + q * 2^e, lies
//    outside the float box on its side; empty slots hold the inverted box
//    (255 / 0) and the given reference.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <random>
#include <vector>

#include "wide_bvh.h"

using namespace muonx;

namespace {

#define CHECK(cond, ...)                             \
  do {                                               \
    if (!(cond)) {                                   \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                      \
      std::printf("\n");                             \
      std::exit(1);                                  \
    }                                                \
  } while (0)

struct Tree {
  std::vector<float> bounds;  // 6 per node
  std::vector<int32_t> left, right, axis, start, count;
  std::vector<float> prim_box;  // 6 per primitive, leaf order
  muonx_scene_desc desc{};
  uint32_t np = 0;
  void finish() {
    desc = muonx_scene_desc{};
    desc.num_prims = np;
    desc.num_nodes = uint32_t(left.size());
    desc.node_bounds = bounds.data();
    desc.node_left = left.data();
    desc.node_right = right.data();
    desc.node_axis = axis.data();
    desc.node_start = start.data();
    desc.node_count = count.data();
  }
};

void Union(float *dst, const float *b) {
  for (int a = 0; a < 3; ++a) {
    dst[a] = std::min(dst[a], b[a]);
    dst[a + 3] = std::max(dst[a + 3], b[a + 3]);
  }
}

// A binary tree over n random primitive boxes: median splits on a random
// axis, leaves of 1..max_leaf primitives, in pre-order.
Tree RandomTree(std::mt19937 &rng, uint32_t n, uint32_t max_leaf, float scale, int style) {
  Tree t;
  t.np = n;
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::vector<float> box(size_t(n) * 6);
  for (uint32_t i = 0; i < n; ++i) {
    float c[3], h[3];
    for (int a = 0; a < 3; ++a) {
      c[a] = (style == 1 ? 0.5f : u(rng)) * scale;   // style 1: coincident centres
      h[a] = style == 2 ? 0.0f : u(rng) * 0.05f * scale;  // style 2: flat (points)
    }
    if (style == 3 && (i & 1)) h[i % 3] = 0.0f;  // some flat in one axis
    for (int a = 0; a < 3; ++a) {
      box[size_t(i) * 6 + a] = c[a] - h[a];
      box[size_t(i) * 6 + a + 3] = c[a] + h[a];
    }
  }
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = i;
  std::function<int32_t(uint32_t, uint32_t)> build = [&](uint32_t lo, uint32_t hi) -> int32_t {
    const int32_t me = int32_t(t.left.size());
    t.left.push_back(-1);
    t.right.push_back(-1);
    t.axis.push_back(0);
    t.start.push_back(0);
    t.count.push_back(0);
    float b[6] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                  std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
                  -std::numeric_limits<float>::max(), -std::numeric_limits<float>::max()};
    for (uint32_t i = lo; i < hi; ++i) Union(b, &box[size_t(ids[i]) * 6]);
    t.bounds.insert(t.bounds.end(), b, b + 6);
    const uint32_t cnt = hi - lo;
    const uint32_t leaf_size = 1 + uint32_t(rng() % max_leaf);
    if (cnt <= leaf_size && !(me == 0 && cnt >= 2)) {  // the root is always split
      t.start[size_t(me)] = int32_t(t.prim_box.size() / 6);
      t.count[size_t(me)] = int32_t(cnt);
      for (uint32_t i = lo; i < hi; ++i)
        t.prim_box.insert(t.prim_box.end(), &box[size_t(ids[i]) * 6], &box[size_t(ids[i]) * 6] + 6);
      return me;
    }
    const int ax = int(rng() % 3);
    const uint32_t mid = lo + cnt / 2;
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi, [&](uint32_t x, uint32_t y) {
      return box[size_t(x) * 6 + ax] + box[size_t(x) * 6 + ax + 3] < box[size_t(y) * 6 + ax] + box[size_t(y) * 6 + ax + 3];
    });
    t.axis[size_t(me)] = ax;
    const int32_t l = build(lo, mid);
    const int32_t r = build(mid, hi);
    t.left[size_t(me)] = l;
    t.right[size_t(me)] = r;
    return me;
  };
  build(0, n);
  t.finish();
  return t;
}

uint32_t RefsOf(const std::vector<F4> &w, size_t node, int k) {
  const float *r = &w[node * 8 + 6].x;
  uint32_t v;
  std::memcpy(&v, r + k, 4);
  return v;
}

// SAH cost of a collapse with the collapse's own weights.
struct CostModel {
  const Tree &t;
  const std::vector<Span> &span;
  uint32_t merge;
  bool is_leaf(int32_t c) const {
    return t.count[size_t(c)] != 0 || (merge && span[size_t(c)].count && span[size_t(c)].count <= merge);
  }
  uint32_t prims(int32_t c) const { return t.count[size_t(c)] != 0 ? uint32_t(t.count[size_t(c)]) : span[size_t(c)].count; }
  float sa(int32_t c) const { return SurfaceArea(&t.bounds[size_t(c) * 6]); }
  float leaf(int32_t c) const { return sa(c) * float(prims(c)) * kCollapsePrimCost; }
};

// Cost of the greedy collapse (the largest-area internal child is expanded
// until a node has four children).
double GreedyCost(const CostModel &m, int32_t root) {
  double total = 0.0;
  std::vector<int32_t> todo{root};
  while (!todo.empty()) {
    const int32_t n = todo.back();
    todo.pop_back();
    total += double(m.sa(n)) * kCollapseNodeCost;
    std::vector<int32_t> c{m.t.left[size_t(n)], m.t.right[size_t(n)]};
    while (c.size() < 4) {
      int best = -1;
      float bsa = -1.0f;
      for (size_t k = 0; k < c.size(); ++k) {
        if (m.is_leaf(c[k])) continue;
        if (m.sa(c[k]) > bsa) {
          bsa = m.sa(c[k]);
          best = int(k);
        }
      }
      if (best < 0) break;
      const int32_t b = c[size_t(best)];
      c[size_t(best)] = m.t.left[size_t(b)];
      c.insert(c.begin() + best + 1, m.t.right[size_t(b)]);
    }
    for (int32_t x : c) {
      if (m.is_leaf(x)) total += double(m.leaf(x));
      else todo.push_back(x);
    }
  }
  return total;
}

// Least cost over every collapse, by enumerating the cuts explicitly (small
// trees only): each wide node takes any cut of its binary subtree into 2..4
// pieces, each piece a leaf (if it may be one) or another wide node.
double BruteCost(const CostModel &m, int32_t n);
// all cuts of the subtree at n into at most k pieces, as piece lists
void Cuts(const CostModel &m, int32_t n, int k, std::vector<std::vector<int32_t>> &out) {
  out.push_back({n});
  if (k < 2 || m.is_leaf(n)) return;
  for (int i = 1; i < k; ++i) {
    std::vector<std::vector<int32_t>> a, b;
    Cuts(m, m.t.left[size_t(n)], i, a);
    Cuts(m, m.t.right[size_t(n)], k - i, b);
    for (auto &x : a)
      for (auto &y : b) {
        std::vector<int32_t> z = x;
        z.insert(z.end(), y.begin(), y.end());
        if (int(z.size()) <= k) out.push_back(z);
      }
  }
}
double PieceCost(const CostModel &m, int32_t c) { return m.is_leaf(c) ? double(m.leaf(c)) : BruteCost(m, c); }
double BruteCost(const CostModel &m, int32_t n) {
  std::vector<std::vector<int32_t>> a, b;
  double best = 1e300;
  for (int i = 1; i <= 3; ++i) {
    a.clear();
    b.clear();
    Cuts(m, m.t.left[size_t(n)], i, a);
    Cuts(m, m.t.right[size_t(n)], 4 - i, b);
    for (auto &x : a)
      for (auto &y : b) {
        if (x.size() + y.size() > 4) continue;
        double c = double(m.sa(n)) * kCollapseNodeCost;
        for (int32_t p : x) c += PieceCost(m, p);
        for (int32_t p : y) c += PieceCost(m, p);
        best = std::min(best, c);
      }
  }
  return best;
}

// Walks the wide tree; returns its SAH cost and checks its structure.
double CheckWide(const Tree &t, const std::vector<F4> &w, uint32_t merge, uint32_t max_stack) {
  const size_t nw = w.size() / 8;
  std::vector<int> seen_node(nw, 0), seen_prim(t.np, 0);
  // binary node of each leaf start (merge 0: leaves are the binary leaves)
  std::vector<int32_t> leaf_of_start(t.np, -1);
  for (size_t i = 0; i < t.left.size(); ++i)
    if (t.count[i] != 0) leaf_of_start[size_t(t.start[i])] = int32_t(i);
  double cost = 0.0;
  uint32_t deepest = 0;
  struct It {
    size_t node;
    uint32_t depth;
    const float *box;  // the parent's child box (null at the root)
  };
  std::vector<It> todo{{0, 1, nullptr}};
  while (!todo.empty()) {
    const It it = todo.back();
    todo.pop_back();
    CHECK(it.node < nw, "wide node %zu out of range", it.node);
    CHECK(seen_node[it.node] == 0, "wide node %zu reached twice", it.node);
    seen_node[it.node] = 1;
    const F4 *q = &w[it.node * 8];
    uint32_t n;
    std::memcpy(&n, &q[7].x, 4);
    CHECK(n >= 2 && n <= 4, "wide node %zu has %u children", it.node, n);
    float nb[6] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max(), -std::numeric_limits<float>::max(),
                   -std::numeric_limits<float>::max(), -std::numeric_limits<float>::max()};
    for (uint32_t k = 0; k < n; ++k) {
      float cb[6] = {(&q[0].x)[k], (&q[2].x)[k], (&q[4].x)[k], (&q[1].x)[k], (&q[3].x)[k], (&q[5].x)[k]};
      Union(nb, cb);
      const uint32_t ref = RefsOf(w, it.node, int(k));
      if (ref & kLeafBit) {
        deepest = std::max(deepest, it.depth);
        uint32_t first, cnt;
        if (merge) {
          first = ref & kLeafFirstMask;
          cnt = ((ref >> kLeafCountShift) & (kLeafMaxCount - 1u)) + 1u;
        } else {
          first = ref & ~kLeafBit;
          CHECK(first < t.np && leaf_of_start[first] >= 0, "leaf ref %x has no binary leaf", ref);
          cnt = uint32_t(t.count[size_t(leaf_of_start[first])]);
        }
        CHECK(first + cnt <= t.np, "leaf [%u, %u) beyond %u primitives", first, first + cnt, t.np);
        for (uint32_t p = first; p < first + cnt; ++p) {
          CHECK(seen_prim[p] == 0, "primitive %u in two leaves", p);
          seen_prim[p] = 1;
          for (int a = 0; a < 3; ++a)
            CHECK(cb[a] <= t.prim_box[size_t(p) * 6 + a] && cb[a + 3] >= t.prim_box[size_t(p) * 6 + a + 3],
                  "leaf box misses primitive %u", p);
        }
        cost += double(2.0f * ((cb[3] - cb[0]) * (cb[4] - cb[1]) + (cb[4] - cb[1]) * (cb[5] - cb[2]) +
                               (cb[5] - cb[2]) * (cb[3] - cb[0]))) *
                double(cnt) * kCollapsePrimCost;
      } else {
        todo.push_back({size_t(ref), it.depth + 1, nullptr});
        // the child's own children must sit inside this box
        const F4 *c = &w[size_t(ref) * 8];
        uint32_t cn;
        std::memcpy(&cn, &c[7].x, 4);
        for (uint32_t j = 0; j < cn && j < 4; ++j) {
          const float gb[6] = {(&c[0].x)[j], (&c[2].x)[j], (&c[4].x)[j], (&c[1].x)[j], (&c[3].x)[j], (&c[5].x)[j]};
          for (int a = 0; a < 3; ++a)
            CHECK(cb[a] <= gb[a] && cb[a + 3] >= gb[a + 3], "child box of node %zu misses a grandchild", it.node);
        }
      }
    }
    cost += double(2.0f * ((nb[3] - nb[0]) * (nb[4] - nb[1]) + (nb[4] - nb[1]) * (nb[5] - nb[2]) +
                           (nb[5] - nb[2]) * (nb[3] - nb[0]))) *
            kCollapseNodeCost;
  }
  for (size_t i = 0; i < nw; ++i) CHECK(seen_node[i] == 1, "wide node %zu unreachable", i);
  for (uint32_t p = 0; p < t.np; ++p) CHECK(seen_prim[p] == 1, "primitive %u in no leaf", p);
  CHECK(max_stack == std::max(1u, 3u * deepest), "stack bound %u, deepest leaf at wide depth %u", max_stack,
        deepest);
  return cost;
}

void CheckQuantized(const std::vector<F4> &w, const std::vector<U4> &qn, uint32_t empty_ref) {
  const size_t nw = w.size() / 8;
  CHECK(qn.size() == nw * 4, "quantized node count");
  for (size_t i = 0; i < nw; ++i) {
    const F4 *q = &w[i * 8];
    const U4 *o = &qn[i * 4];
    uint32_t n;
    std::memcpy(&n, &q[7].x, 4);
    CHECK((o[0].w >> 24) == n, "node %zu: count %u vs %u", i, o[0].w >> 24, n);
    const float org[3] = {BitsAsFloat(o[0].x), BitsAsFloat(o[0].y), BitsAsFloat(o[0].z)};
    const int ex[3] = {int(o[0].w & 255u) - 127, int((o[0].w >> 8) & 255u) - 127, int((o[0].w >> 16) & 255u) - 127};
    const uint32_t lo_w[3] = {o[1].x, o[1].y, o[1].z}, hi_w[3] = {o[1].w, o[2].x, o[2].y};
    const uint32_t refs[4] = {o[3].x, o[3].y, o[3].z, o[3].w};
    for (uint32_t k = 0; k < 4; ++k) {
      for (int a = 0; a < 3; ++a) {
        const uint32_t ql = (lo_w[a] >> (8 * k)) & 255u, qh = (hi_w[a] >> (8 * k)) & 255u;
        if (k >= n) {
          CHECK(ql == 255u && qh == 0u, "node %zu slot %u axis %d: empty slot planes %u / %u", i, k, a, ql, qh);
          continue;
        }
        const float flo = (&q[2 * a].x)[k], fhi = (&q[2 * a + 1].x)[k];
        CHECK(PlaneValue(org[a], ql, ex[a]) <= double(flo), "node %zu child %u axis %d: low plane above the box", i,
              k, a);
        CHECK(PlaneValue(org[a], qh, ex[a]) >= double(fhi), "node %zu child %u axis %d: high plane below the box", i,
              k, a);
      }
      CHECK(refs[k] == (k < n ? RefsOf(w, i, int(k)) : empty_ref), "node %zu slot %u: reference", i, k);
    }
  }
}

void RunCase(std::mt19937 &rng, uint32_t n, uint32_t max_leaf, float scale, int style, bool brute) {
  Tree t = RandomTree(rng, n, max_leaf, scale, style);
  CHECK(t.count[0] == 0, "root is a leaf (n=%u)", n);
  const std::vector<Span> span = SubtreeSpans(&t.desc);
  for (uint32_t merge : {0u, 2u}) {
    uint32_t max_stack = 0;
    const std::vector<F4> w = CollapseWide(&t.desc, span, merge, &max_stack);
    const double cost = CheckWide(t, w, merge, max_stack);
    const CostModel m{t, span, merge};
    const double greedy = GreedyCost(m, 0);
    CHECK(cost <= greedy * (1.0 + 1e-5) + 1e-30, "collapse cost %.9g above the greedy collapse's %.9g (n=%u)", cost,
          greedy, n);
    if (brute) {
      const double best = BruteCost(m, 0);
      CHECK(cost <= best * (1.0 + 1e-5) + 1e-30 && cost >= best * (1.0 - 1e-5) - 1e-30,
            "collapse cost %.9g, least over all collapses %.9g (n=%u, merge %u)", cost, best, n, merge);
    }
    const uint32_t empty_ref = kLeafBit | n;
    CheckQuantized(w, QuantizeWide(w, empty_ref), empty_ref);
  }
}

}  // namespace

int main() {
  std::mt19937 rng(12345);
  // small trees against the enumeration of all collapses
  for (int rep = 0; rep < 300; ++rep) RunCase(rng, 2 + uint32_t(rng() % 11), 1 + uint32_t(rng() % 3), 1.0f, 0, true);
  // larger trees, several box styles and scales
  const float scales[] = {1.0f, 1e-20f, 1e20f, 3.0e5f};
  for (int rep = 0; rep < 60; ++rep) {
    const int style = rep % 4;
    RunCase(rng, 50 + uint32_t(rng() % 3000), 1 + uint32_t(rng() % 4), scales[(rep / 4) % 4], style, false);
  }
  std::printf("ok\n");
  return 0;
}
