// Host SAH BVH build with the tree shape of muon/acceleration.cc:204-458:
// 12-bucket SAH on the widest centroid axis, cost 0.125 + (nL*SA_L +
// nR*SA_R)/SA, leaf when n < cost, std::partition by bucket and the
// std::nth_element midpoint fallback. The same libstdc++ algorithms are
// called with the same predicates on the same input order, so the tree (and
// the leaf order of the primitives) is the reference's.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <limits>
#include <thread>

#include "host_scene.h"
#include "xform.h"

namespace muonx {
namespace host {
namespace {

struct Box {
  V3 mn{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
        std::numeric_limits<float>::infinity()};
  V3 mx{-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
        -std::numeric_limits<float>::infinity()};
  static Box Of2(V3 a, V3 b) {
    Box r;
    r.mn = V3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)};
    r.mx = V3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)};
    return r;
  }
  static Box Point(V3 p) {
    Box r;
    r.mn = p;
    r.mx = p;
    return r;
  }
  Box Add(V3 p) const {
    Box r;
    r.mn = V3{gmin(mn.x, p.x), gmin(mn.y, p.y), gmin(mn.z, p.z)};
    r.mx = V3{gmax(mx.x, p.x), gmax(mx.y, p.y), gmax(mx.z, p.z)};
    return r;
  }
  Box Add(const Box &b) const {
    Box r;
    r.mn = V3{gmin(mn.x, b.mn.x), gmin(mn.y, b.mn.y), gmin(mn.z, b.mn.z)};
    r.mx = V3{gmax(mx.x, b.mx.x), gmax(mx.y, b.mx.y), gmax(mx.z, b.mx.z)};
    return r;
  }
  V3 Dim() const { return mx - mn; }
  float Area() const {
    V3 d = Dim();
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
  }
  int MaxAxis() const {
    V3 d = Dim();
    if (d.x > d.y && d.x > d.z) return 0;
    if (d.y > d.z) return 1;
    return 2;
  }
};

struct PrimInfo {
  size_t index;  // parse order
  Box box;
  V3 centroid;
};

// muon/objects.cc:186-193 (tri: transformed vertices), muon/bounds.cc:52-69
// (sphere: 8 transformed corners of the object box).
Box WorldBox(const HostScene &sc, size_t i) {
  const float *g = &sc.prim_geom[i * 12];
  const float *m = sc.transforms[sc.prim_transform[i]].data();
  if (sc.prim_type[i] == MUONX_PRIM_TRI) {
    V3 a = xform_point(m, V3{g[0], g[1], g[2]});
    V3 b = xform_point(m, V3{g[3], g[4], g[5]});
    V3 c = xform_point(m, V3{g[6], g[7], g[8]});
    return Box::Of2(a, b).Add(c);
  }
  V3 c{g[0], g[1], g[2]};
  float r = g[3];
  Box ob = Box::Of2(c - v3s(r), c + v3s(r));
  V3 lo = ob.mn, hi = ob.mx;
  Box b = Box::Point(xform_point(m, lo));
  b = b.Add(xform_point(m, V3{hi.x, lo.y, lo.z}));
  b = b.Add(xform_point(m, V3{lo.x, hi.y, lo.z}));
  b = b.Add(xform_point(m, V3{lo.x, lo.y, hi.z}));
  b = b.Add(xform_point(m, V3{lo.x, hi.y, hi.z}));
  b = b.Add(xform_point(m, V3{hi.x, lo.y, hi.z}));
  b = b.Add(xform_point(m, V3{hi.x, hi.y, lo.z}));
  return b.Add(xform_point(m, hi));
}

// Nodes of a subtree in pre-order; child links are indices into the same
// buffer until the buffer is spliced into its parent's.
struct NodeBuf {
  std::vector<float> bounds;
  std::vector<int32_t> left, right, axis, start, count;
  int depth = 0;
  int New() {
    const int id = int(left.size());
    left.push_back(-1);
    right.push_back(-1);
    axis.push_back(0);
    start.push_back(0);
    count.push_back(0);
    bounds.resize(bounds.size() + 6, 0.0f);
    return id;
  }
  void SetBounds(int id, const Box &b) {
    float *p = &bounds[size_t(id) * 6];
    p[0] = b.mn.x; p[1] = b.mn.y; p[2] = b.mn.z;
    p[3] = b.mx.x; p[4] = b.mx.y; p[5] = b.mx.z;
  }
  // Appends `o` (a complete subtree); returns the index of its root here.
  int Append(const NodeBuf &o) {
    const int off = int(left.size());
    for (size_t i = 0; i < o.left.size(); ++i) {
      left.push_back(o.left[i] >= 0 ? o.left[i] + off : -1);
      right.push_back(o.right[i] >= 0 ? o.right[i] + off : -1);
    }
    axis.insert(axis.end(), o.axis.begin(), o.axis.end());
    start.insert(start.end(), o.start.begin(), o.start.end());
    count.insert(count.end(), o.count.begin(), o.count.end());
    bounds.insert(bounds.end(), o.bounds.begin(), o.bounds.end());
    depth = std::max(depth, o.depth);
    return off;
  }
};

class Builder {
 public:
  Builder(int partition, std::vector<PrimInfo> &info, int threads)
      : partition_(partition), info_(info), spare_(threads - 1) {}

  // Emits the subtree over info_[start, end) in pre-order into `buf`;
  // returns its root's index there. When both halves of a split are large
  // and a thread is spare, the right half is built on its own thread into
  // its own buffer; both halves are appended in pre-order, so the node
  // numbering is the sequential build's.
  int Build(NodeBuf &buf, size_t start, size_t end, int depth, Box *out_box) {
    buf.depth = std::max(buf.depth, depth);
    const size_t n = end - start;
    if (n == 1) return Leaf(buf, start, n, info_[start].box, out_box);

    Box cb;
    for (size_t i = start; i < end; ++i) cb = cb.Add(info_[i].centroid);
    const int axis = cb.MaxAxis();
    auto first = info_.begin() + start, last = info_.begin() + end;
    size_t split = size_t(-1);
    bool uniform = false;

    if (partition_ == 0) {
      uniform = true;
    } else if (partition_ == 1) {
      const float mid = (comp(cb.mn, axis) + comp(cb.mx, axis)) / 2;
      auto it = std::partition(first, last, [axis, mid](const PrimInfo &p) { return comp(p.centroid, axis) < mid; });
      split = size_t(it - info_.begin());
      uniform = (split == start || split == end);
    } else {
      constexpr size_t kBuckets = 12;
      const float extent = comp(cb.Dim(), axis);
      const float lo = comp(cb.mn, axis);
      auto bucket_of = [extent, lo, axis](const PrimInfo &p) -> size_t {
        float off = extent == 0.0f ? 0.0f : (comp(p.centroid, axis) - lo) / extent;
        size_t b = static_cast<size_t>(off * kBuckets);
        return b == kBuckets ? kBuckets - 1 : b;
      };
      size_t count[kBuckets] = {};
      Box bbox[kBuckets];
      for (size_t i = start; i < end; ++i) {
        size_t b = bucket_of(info_[i]);
        ++count[b];
        bbox[b] = bbox[b].Add(info_[i].box);
      }
      Box all;
      for (size_t b = 0; b < kBuckets; ++b) all = all.Add(bbox[b]);
      float cost[kBuckets - 1] = {};
      size_t acc = 0;
      Box accb;
      for (size_t i = 0; i + 1 < kBuckets; ++i) {
        acc += count[i];
        if (acc == 0) continue;
        accb = accb.Add(bbox[i]);
        cost[i] = acc * accb.Area();
      }
      acc = 0;
      accb = Box();
      for (int i = int(kBuckets) - 2; i >= 0; --i) {
        acc += count[i + 1];
        if (acc == 0) continue;
        accb = accb.Add(bbox[i + 1]);
        cost[i] += acc * accb.Area();
      }
      const float total = all.Area();
      float min_cost = std::numeric_limits<float>::infinity();
      size_t best = 0;
      for (size_t i = 0; i + 1 < kBuckets; ++i) {
        float c = 0.125f + cost[i] / total;
        if (c < min_cost) {
          min_cost = c;
          best = i;
        }
      }
      const float leaf_cost = float(n);
      if (leaf_cost < min_cost) return Leaf(buf, start, n, all, out_box);
      auto it = std::partition(first, last, [&](const PrimInfo &p) { return bucket_of(p) <= best; });
      split = size_t(it - info_.begin());
      uniform = (split == start || split == end);
    }
    if (uniform) {
      split = (start + end) / 2;
      std::nth_element(first, info_.begin() + split, last, [axis](const PrimInfo &a, const PrimInfo &b) {
        return comp(a.centroid, axis) < comp(b.centroid, axis);
      });
    }
    const int me = buf.New();
    Box lb, rb;
    int l, r;
    if (std::min(split - start, end - split) >= kSpawnPrims && TakeThread()) {
      // the two halves touch disjoint ranges of info_
      NodeBuf lbuf, rbuf;
      int lroot = 0, rroot = 0;
      std::exception_ptr err;  // an exception (allocation failure) reaches the caller's thread
      std::thread t([&] {
        try {
          rroot = Build(rbuf, split, end, depth + 1, &rb);
        } catch (...) {
          err = std::current_exception();
        }
      });
      try {
        lroot = Build(lbuf, start, split, depth + 1, &lb);
      } catch (...) {
        t.join();
        spare_.fetch_add(1);
        throw;
      }
      t.join();
      spare_.fetch_add(1);
      if (err) std::rethrow_exception(err);
      l = buf.Append(lbuf) + lroot;
      r = buf.Append(rbuf) + rroot;
    } else {
      l = Build(buf, start, split, depth + 1, &lb);
      r = Build(buf, split, end, depth + 1, &rb);
    }
    Box b = lb.Add(rb);
    buf.left[size_t(me)] = l;
    buf.right[size_t(me)] = r;
    buf.axis[size_t(me)] = axis;
    buf.SetBounds(me, b);
    *out_box = b;
    return me;
  }

 private:
  static constexpr size_t kSpawnPrims = 1 << 15;

  bool TakeThread() {
    int s = spare_.load();
    while (s > 0)
      if (spare_.compare_exchange_weak(s, s - 1)) return true;
    return false;
  }

  int Leaf(NodeBuf &buf, size_t start, size_t n, const Box &b, Box *out_box) {
    int id = buf.New();
    buf.start[size_t(id)] = int32_t(start);
    buf.count[size_t(id)] = int32_t(n);
    buf.SetBounds(id, b);
    *out_box = b;
    return id;
  }

  int partition_;
  std::vector<PrimInfo> &info_;
  std::atomic<int> spare_;
};

}  // namespace

void BuildBVH(HostScene &sc, int acceleration, int partition) {
  const size_t n = sc.num_prims();
  sc.acceleration = acceleration == 0 ? 0 : 1;
  sc.node_bounds.clear();
  sc.node_left.clear();
  sc.node_right.clear();
  sc.node_axis.clear();
  sc.node_start.clear();
  sc.node_count.clear();
  sc.leaf_prims.clear();
  sc.bvh_depth = 0;
  if (n == 0) return;
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<PrimInfo> info(n);
  {
    auto boxes = [&](size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; ++i) {
        Box b = WorldBox(sc, i);
        V3 diag = b.mx - b.mn;
        info[i] = PrimInfo{i, b, b.mn + 0.5f * diag};
      }
    };
    const size_t nt = n >= (size_t(1) << 16) ? hw : 1;
    std::vector<std::thread> ts;  // WorldBox only reads the scene and writes its own slots
    for (size_t t = 1; t < nt; ++t) ts.emplace_back(boxes, n * t / nt, n * (t + 1) / nt);
    boxes(0, n / nt);
    for (auto &t : ts) t.join();
  }
  if (acceleration == 0) {
    // Linear structure (muon/acceleration.cc:14-45): every primitive in parse
    // order; modelled as a single leaf over the scene bounds.
    Box all;
    for (const auto &p : info) all = all.Add(p.box);
    sc.node_left.push_back(-1);
    sc.node_right.push_back(-1);
    sc.node_axis.push_back(0);
    sc.node_start.push_back(0);
    sc.node_count.push_back(int32_t(n));
    sc.node_bounds.insert(sc.node_bounds.end(), {all.mn.x, all.mn.y, all.mn.z, all.mx.x, all.mx.y, all.mx.z});
    for (size_t i = 0; i < n; ++i) sc.leaf_prims.push_back(int32_t(i));
    return;
  }
  Builder b(partition, info, int(hw));
  NodeBuf nodes;
  Box root;
  b.Build(nodes, 0, n, 1, &root);
  sc.node_bounds = std::move(nodes.bounds);
  sc.node_left = std::move(nodes.left);
  sc.node_right = std::move(nodes.right);
  sc.node_axis = std::move(nodes.axis);
  sc.node_start = std::move(nodes.start);
  sc.node_count = std::move(nodes.count);
  sc.bvh_depth = nodes.depth;
  sc.leaf_prims.resize(n);
  for (size_t i = 0; i < n; ++i) sc.leaf_prims[i] = int32_t(info[i].index);
}

}  // namespace host
}  // namespace muonx
