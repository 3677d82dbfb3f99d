// BVH traversal for gfx950: one ray per lane, child-boxes-in-parent binary
// nodes, a per-lane short stack in LDS (entries = child ref + entry distance),
// and the reference's primitive tests in object space.
//
// Parity contract with the reference (muon/acceleration.cc:65-202,
// muon/objects.cc:29-270):
//  * primitive tests are op-for-op the reference's (object-space ray through
//    M^-1 with a re-normalised direction, unnormalised-normal 1e-4 parallel
//    test, edge tests, world distance |M p - o|), compiled without FMA
//    contraction and with IEEE division/sqrt, so distances carry its bits;
//  * acceptance is strict `0 < dist < best`, and children are visited in the
//    reference's DFS order (children[dir[axis] < 0] first), so ties resolve
//    the same way;
//  * box tests multiply by a per-ray reciprocal instead of dividing, and
//    widen the [t_lo, t_hi] interval by 2^-20 relative so that every box the
//    reference's exact slab test (muon/bounds.cc:88-125) accepts is accepted
//    here. Visiting a superset of boxes cannot change the closest hit except
//    for distance ties within rounding.
#ifndef MUONX_DEVICE_TRACE_H_
#define MUONX_DEVICE_TRACE_H_

#include "../common/vecmath.h"
#include "dev_scene.h"
#include "gload.h"

namespace muonx {

constexpr float kWidenLo = 1.0f - 1.0f / 1048576.0f;
constexpr float kWidenHi = 1.0f + 1.0f / 1048576.0f;
constexpr float kSlack = 1.0f / 1048576.0f;

struct TraceRay {
  V3 o, d, inv;
  uint32_t dneg;  // bit a set when d[a] < 0
};

// The reciprocal direction of the fast slab test, infinite for a zero
// component (the sign follows the zero's sign). The reference skips such an
// axis (muon/bounds.cc:98-100); here a plane then yields +-inf, or NaN for an
// origin exactly on the plane, and the slab test's min / max ignore NaN, so
// the axis only rejects boxes the ray cannot reach.
__device__ __forceinline__ float safe_inv(float d) { return 1.0f / d; }

__device__ __forceinline__ TraceRay make_ray(V3 o, V3 d) {
  TraceRay r;
  r.o = o;
  r.d = d;
  r.inv = V3{safe_inv(d.x), safe_inv(d.y), safe_inv(d.z)};
  r.dneg = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
  return r;
}

// The largest float below b, for b > 0: l < b exactly when l <= below(b).
__device__ __forceinline__ float below(float b) { return __uint_as_float(__float_as_uint(b) - 1u); }

// Conservative slab test against `best_m` = below(best); returns the widened
// entry distance in `lo`. Accepts when the widened interval [l, h] is not
// empty and starts before best (l >= 0, so h >= 0; a box touching the origin
// from behind, h == 0, is visited and finds nothing).
__device__ __forceinline__ bool box_hit(const TraceRay &r, float mnx, float mxx, float mny, float mxy, float mnz,
                                        float mxz, float best_m, float &lo) {
  // entry (near) and exit (far) plane per axis from the reciprocal's sign;
  // a NaN distance (origin on a plane of a zero-direction axis) drops out of
  // the max / min
  const bool sx = signbit(r.inv.x), sy = signbit(r.inv.y), sz = signbit(r.inv.z);
  const float ax = ((sx ? mxx : mnx) - r.o.x) * r.inv.x, bx = ((sx ? mnx : mxx) - r.o.x) * r.inv.x;
  const float ay = ((sy ? mxy : mny) - r.o.y) * r.inv.y, by = ((sy ? mny : mxy) - r.o.y) * r.inv.y;
  const float az = ((sz ? mxz : mnz) - r.o.z) * r.inv.z, bz = ((sz ? mnz : mxz) - r.o.z) * r.inv.z;
  float l = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f));
  float h = fminf(fminf(bx, by), bz);
  l = l * kWidenLo;
  h = h * kWidenHi;
  lo = l;
  return l <= fminf(h, best_m);
}

// The reference's slab test (muon/bounds.cc:88-125) with its divisions and
// zero-direction skip; returns its t_min in `lo`. With this test and the
// exact t_min kept on the stack, the traversal makes the reference's accept
// / reject decision for every node: the near child is tested with the same
// best distance the reference uses when it descends, the far child's
// distance-independent part (overlap, t_max > 0) is decided at the parent
// and its `t_min < best` part when it is popped, as the reference does.
// Skipping the root's own test is exact: child intervals are contained in
// the root's, since correctly rounded arithmetic is monotonic.
__device__ __forceinline__ bool box_hit_exact(const TraceRay &r, float mnx, float mxx, float mny, float mxy,
                                              float mnz, float mxz, float best, float &lo) {
  float tmin = 0.0f, tmax = __builtin_inff();
  const float mn[3] = {mnx, mny, mnz}, mx[3] = {mxx, mxy, mxz};
  const float o[3] = {r.o.x, r.o.y, r.o.z}, d[3] = {r.d.x, r.d.y, r.d.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (d[a] == 0.0f) continue;
    float ta = (mn[a] - o[a]) / d[a];
    float tb = (mx[a] - o[a]) / d[a];
    if (ta > tb) {
      const float tmp = ta;
      ta = tb;
      tb = tmp;
    }
    tmin = gmax(tmin, ta);
    tmax = gmin(tmax, tb);
  }
  lo = tmin;
  return !(tmin > tmax) && tmax > 0.0f && tmin < best;
}

// muon/objects.cc:58-175 in object space. On a hit returns the object-space
// point; optionally the unnormalised barycentric numerators (u, v, w) and n2.
__device__ __forceinline__ bool tri_test(V3 a, V3 b, V3 c, V3 o, V3 d, V3 &p) {
  V3 n = cross(b - a, c - a);
  float dn = dot(n, d);
  if (dn > -kEps && dn < kEps) return false;
  float t = (dot(a, n) - dot(o, n)) / dn;
  if (t < 0.0f) return false;
  p = o + d * t;
  if (dot(n, cross(b - a, p - a)) < 0.0f) return false;
  if (dot(n, cross(c - b, p - b)) < 0.0f) return false;
  if (dot(n, cross(a - c, p - c)) < 0.0f) return false;
  return true;
}

// muon/objects.cc:195-270 in object space (direction assumed unit length).
__device__ __forceinline__ bool sphere_test(V3 c, float rad, V3 o, V3 d, V3 &p) {
  V3 oc = o - c;
  float bp = dot(d, oc);
  float cc = dot(oc, oc) - rad * rad;
  float disc = bp * bp - cc;
  if (disc < 0.0f) return false;
  float sd = sqrtf(disc);
  float r0 = -bp + sd, r1 = -bp - sd;
  if (r0 < 0.0f && r1 < 0.0f) return false;
  float t = gmin(r0, r1);
  if (t < 0.0f) t = gmax(r0, r1);
  p = o + d * t;
  return true;
}

// Ray.Transform(M^-1) of muon/ray.cc:7-11.
__device__ __forceinline__ void to_object(const float *xf, V3 o, V3 d, V3 &oo, V3 &od) {
  float mi[16];
  load_xform(xf + 16, mi);
  oo = xform_point(mi, o);
  od = normalize(xform_vec(mi, d));
}

struct TraceCounters {
  uint32_t nodes = 0;   // internal nodes fetched (two child boxes each)
  uint32_t prims = 0;   // primitive tests (the reference's object tests)
  uint32_t ohits = 0;   // object hits (as muon/acceleration.cc:108-114,177-179 count them)
  // bounds tests / hits: in the exact mode the reference's (muon/acceleration.cc:92,101,162,171);
  // in the fast mode the child boxes the device tree's nodes test and pass
  uint32_t btests = 0;
  uint32_t bhits = 0;
};

// Tests leaf-order primitive i against the ray; returns the world distance
// on an object-space hit.
struct ObjCache {
  int32_t tid;
  V3 o, d;
};

__device__ __forceinline__ bool prim_hit(const DevSceneView &s, uint32_t i, const TraceRay &r, ObjCache &oc,
                                         float &dist, uint32_t &flags) {
  const float4 q0 = ldg(s.prims + 3 * i);
  const float4 q1 = ldg(s.prims + 3 * i + 1);
  const uint32_t tw = __float_as_uint(q0.w);
  flags = __float_as_uint(q1.w);
  const int32_t tid = int32_t(tw & 0x7fffffffu);
  const float *xf = s.xforms + size_t(tid) * 48;
  if (tid != oc.tid) {
    to_object(xf, r.o, r.d, oc.o, oc.d);
    oc.tid = tid;
  }
  V3 p;
  bool hit;
  if (tw & kSphereBit) {
    hit = sphere_test(V3{q0.x, q0.y, q0.z}, q1.x, oc.o, oc.d, p);
  } else {
    const float4 q2 = ldg(s.prims + 3 * i + 2);
    hit = tri_test(V3{q0.x, q0.y, q0.z}, V3{q1.x, q1.y, q1.z}, V3{q2.x, q2.y, q2.z}, oc.o, oc.d, p);
  }
  if (!hit) return false;
  float m[16];
  load_xform(xf, m);
  V3 pos = xform_point(m, p);
  dist = length(pos - r.o);
  return true;
}

// Per-lane traversal stacks. Entries are (child ref, entry distance).
//
// LdsStack: CAP entries in LDS at base[k * stride] (stride = block size, so
// lane accesses are bank-conflict free). The caller sizes CAP >= BVH depth.
template <int CAP>
struct LdsStack {
  uint2 *base;
  int stride;
  int sp;
  __device__ __forceinline__ LdsStack(uint2 *b, int s) : base(b), stride(s), sp(0) {}
  __device__ __forceinline__ void reset() { sp = 0; }
  __device__ __forceinline__ void push(uint2 v) {
    if (sp < CAP) {
      base[sp * stride] = v;
      ++sp;
    }
  }
  // Pushes c, b, a (the last m of them, m <= 3) in that order: a ends on top.
  // Three unconditional writes, c then b then a: an entry that is not pushed
  // is written to a's slot first and overwritten by a.
  __device__ __forceinline__ void push_last(int m, uint2 a, uint2 b, uint2 c) {
    if (m <= 0) return;
    if (sp + m > CAP) {
      if (m >= 3) push(c);
      if (m >= 2) push(b);
      push(a);
      return;
    }
    const int pa = sp + m - 1, pb = m >= 2 ? sp + m - 2 : pa, pc = m >= 3 ? sp : pa;
    base[pc * stride] = c;
    base[pb * stride] = b;
    base[pa * stride] = a;
    sp += m;
  }
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ uint2 pop() {
    --sp;
    return base[sp * stride];
  }
};

// RingStack: the top S entries (S a power of two) in an LDS ring, older
// entries spilled to a per-thread global region g[k * gstride], k <
// gcap. Keeps LDS use small (high occupancy) while supporting deep trees.
template <int S>
struct RingStack {
  uint2 *base;
  int stride;
  uint2 *g;
  int gstride;
  int gcap;
  int sp;
  __device__ __forceinline__ RingStack(uint2 *b, int s, uint2 *gb, int gs, int gc)
      : base(b), stride(s), g(gb), gstride(gs), gcap(gc), sp(0) {}
  __device__ __forceinline__ void reset() { sp = 0; }
  __device__ __forceinline__ void push(uint2 v) {
    const int slot = sp & (S - 1);
    if (sp >= S) {
      if (sp - S >= gcap) return;  // sized from the BVH depth: never taken
      g[(sp - S) * gstride] = base[slot * stride];
    }
    base[slot * stride] = v;
    ++sp;
  }
  // Pushes c, b, a (the last m of them, m <= 3) in that order: a ends on top.
  // Without a spill (the common case) the entries go straight to their ring
  // slots (all below S) with three unconditional writes, c then b then a: an
  // entry that is not pushed is written to a's slot first and overwritten.
  __device__ __forceinline__ void push_last(int m, uint2 a, uint2 b, uint2 c) {
    if (m <= 0) return;
    if (sp + m > S) {
      if (m >= 3) push(c);
      if (m >= 2) push(b);
      push(a);
      return;
    }
    const int pa = sp + m - 1, pb = m >= 2 ? sp + m - 2 : pa, pc = m >= 3 ? sp : pa;
    base[pc * stride] = c;
    base[pb * stride] = b;
    base[pa * stride] = a;
    sp += m;
  }
  __device__ __forceinline__ bool empty() const { return sp == 0; }
  __device__ __forceinline__ uint2 pop() {
    --sp;
    const int slot = sp & (S - 1);
    const uint2 v = base[slot * stride];
    if (sp >= S) base[slot * stride] = g[(sp - S) * gstride];
    return v;
  }
};

// Traverses the BVH. ANY: stop at the first primitive with 0 < dist < tmax
// (HasIntersection); otherwise closest hit (Intersect). EXACT selects the
// reference's division slab test (identical node visits, used by the
// deterministic integrators and the exact query mode); the default is the
// conservative reciprocal test. WW ("while-while"): each lane walks internal
// nodes until it holds a leaf, then the wavefront tests leaves together,
// instead of interleaving node and leaf steps lane by lane. Returns the
// leaf-order primitive index or -1; `best` holds the distance on return.
// One internal-node step: tests both child boxes, pushes the far child when
// both are hit, and moves `ref` to the next node; returns false when the
// stack is exhausted.
template <bool COUNT, bool EXACT, class Stack>
__device__ __forceinline__ bool node_step(const DevSceneView &s, const TraceRay &r, float best, Stack &stack,
                                          uint32_t &ref, TraceCounters &cnt) {
  const float4 *np = s.nodes + 4 * size_t(ref);
  const float4 n0 = ldg(np), n1 = ldg(np + 1), n2 = ldg(np + 2);
  const int4 n3 = ldg(reinterpret_cast<const int4 *>(np + 3));
  if (COUNT) cnt.nodes++;
  float lo0, lo1;
  bool h0, h1;
  if (EXACT) {
    h0 = box_hit_exact(r, n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, best, lo0);
    h1 = box_hit_exact(r, n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, best, lo1);
  } else {
    const float bm = below(best);
    h0 = box_hit(r, n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, bm, lo0);
    h1 = box_hit(r, n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, bm, lo1);
  }
  if (COUNT && !EXACT) {  // the fast mode's own work: both child boxes tested
    cnt.btests += 2;
    cnt.bhits += uint32_t(h0) + uint32_t(h1);
  }
  const uint32_t axis = (uint32_t(n3.z) >> 29) & 3u;
  const bool swap = (r.dneg >> axis) & 1u;  // near child is children[1]
  const uint32_t near_ref = swap ? uint32_t(n3.y) : uint32_t(n3.x);
  const uint32_t far_ref = swap ? uint32_t(n3.x) : uint32_t(n3.y);
  const bool hn = swap ? h1 : h0;
  const bool hf = swap ? h0 : h1;
  const float lof = swap ? lo0 : lo1;
  if (COUNT && EXACT) {
    // The reference (muon/acceleration.cc:137,197) always pushes the far
    // child and tests its box only when it pops it, which an any-hit query
    // that stops at its first hit (:182-183) never does. So with the near box hit the
    // far child is pushed too, with an infinite key when its box was missed
    // here, and its test is counted at its pop (pop_next_counted).
    if (hn) {
      cnt.btests++;
      cnt.bhits++;
      stack.push(make_uint2(far_ref, hf ? __float_as_uint(lof) : 0x7f800000u));
      ref = near_ref;
      return true;
    }
    cnt.btests += 2;  // the near box, then the far one popped at once
    if (hf) {
      cnt.bhits++;
      ref = far_ref;
      return true;
    }
    return false;
  }
  if (hn) {
    if (hf) stack.push(make_uint2(far_ref, __float_as_uint(lof)));
    ref = near_ref;
    return true;
  }
  if (hf) {
    ref = far_ref;
    return true;
  }
  return false;
}

// Pops the next entry whose box starts before `best`.
template <class Stack>
__device__ __forceinline__ bool pop_next(Stack &stack, float best, uint32_t &ref) {
  while (!stack.empty()) {
    const uint2 e = stack.pop();
    if (__uint_as_float(e.y) < best) {
      ref = e.x;
      return true;
    }
  }
  return false;
}

// pop_next for the counted exact traversal: every popped far child is the
// reference's bounds test at its pop, a hit when its box still starts
// before `best`.
template <class Stack>
__device__ __forceinline__ bool pop_next_counted(Stack &stack, float best, uint32_t &ref, TraceCounters &cnt) {
  while (!stack.empty()) {
    const uint2 e = stack.pop();
    cnt.btests++;
    if (__uint_as_float(e.y) < best) {
      cnt.bhits++;
      ref = e.x;
      return true;
    }
  }
  return false;
}

// Tests the primitives of the leaf at `ref`; returns true if ANY found a hit.
// A wide-tree leaf (WIDE) whose reference carries a count tests that many
// primitives; otherwise the leaf ends at the primitive flagged last.
template <bool ANY, bool COUNT, bool WIDE = false>
__device__ __forceinline__ bool leaf_step(const DevSceneView &s, const TraceRay &r, float &best, uint32_t ref,
                                          ObjCache &oc, int32_t &hit_prim, TraceCounters &cnt) {
  uint32_t i = ref & ~kLeafBit, left = 0;
  if (WIDE && s.leaf_counts) {
    i = ref & kLeafFirstMask;
    left = ((ref >> kLeafCountShift) & (kLeafMaxCount - 1u)) + 1u;
  }
  while (true) {
    float dist;
    uint32_t flags;
    if (COUNT) cnt.prims++;
    const bool ph = prim_hit(s, i, r, oc, dist, flags);
    // an object hit is Primitive::Intersect's (closest hit) or
    // HasIntersection's (any hit: also 0 < dist < tmax), muon/objects.cc:9-17
    if (COUNT && ph && (!ANY || (dist > 0.0f && dist < best))) cnt.ohits++;
    if (ph && dist > 0.0f && dist < best) {
      best = dist;
      hit_prim = int32_t(i);
      if (ANY) return true;
    }
    if (left ? --left == 0u : (flags & 1u) != 0u) return false;
    ++i;
  }
}

// Traverses the BVH. ANY: stop at the first primitive with 0 < dist < tmax
// (HasIntersection); otherwise closest hit (Intersect). EXACT selects the
// reference's division slab test (identical node visits, used by the
// deterministic integrators and the exact query mode); the default is the
// conservative reciprocal test. WW ("while-while"): every lane walks
// internal nodes until it holds a leaf, then the wavefront tests leaves
// together; otherwise each iteration does one node or one leaf step per lane.
// Returns the leaf-order primitive index or -1; `best` holds the distance.
template <bool ANY, bool COUNT, bool EXACT, bool WW, class Stack>
__device__ __forceinline__ int32_t trace_bvh_s(const DevSceneView &s, const TraceRay &r, float &best, Stack &stack,
                                               TraceCounters &cnt) {
  int32_t hit_prim = -1;
  ObjCache oc;
  oc.tid = -1;
  uint32_t ref = 0;  // root: internal node 0
  constexpr bool kRefStats = COUNT && EXACT;
  if (kRefStats && s.linear) {
    // Linear::Intersect tests every primitive and counts object tests and
    // hits only; Linear::HasIntersection counts nothing
    // (muon/acceleration.cc:14-45). The one leaf is node 0's first child.
    if constexpr (ANY) {
      TraceCounters none;
      return trace_bvh_s<ANY, false, EXACT, WW>(s, r, best, stack, none);
    }
    ref = __float_as_uint(ldg(s.nodes + 3).x);
  } else if (kRefStats) {
    // The reference tests the root's box first (muon/acceleration.cc:92-99);
    // here it is the union of node 0's child boxes. A ray that misses it
    // misses every box below, so stopping changes no result. A root that is
    // a leaf sits in node 0's first child (the second is a sentinel).
    const float4 *np = s.nodes;
    const float4 n0 = ldg(np), n1 = ldg(np + 1), n2 = ldg(np + 2);
    float lo;
    cnt.btests++;
    if (!box_hit_exact(r, fminf(n0.x, n1.x), fmaxf(n0.y, n1.y), fminf(n0.z, n1.z), fmaxf(n0.w, n1.w),
                       fminf(n2.x, n2.z), fmaxf(n2.y, n2.w), best, lo))
      return -1;
    cnt.bhits++;
    if (s.root_leaf) ref = __float_as_uint(np[3].x);
  }
  auto pop = [&](uint32_t &rf) {
    if constexpr (kRefStats) return pop_next_counted(stack, best, rf, cnt);
    else return pop_next(stack, best, rf);
  };
  if (WW) {
    while (true) {
      bool alive = true;
      while (!(ref & kLeafBit)) {
        if (!node_step<COUNT, EXACT>(s, r, best, stack, ref, cnt) && !pop(ref)) {
          alive = false;
          break;
        }
      }
      if (!alive) break;
      if (leaf_step<ANY, COUNT>(s, r, best, ref, oc, hit_prim, cnt)) return hit_prim;
      if (!pop(ref)) break;
    }
  } else {
    while (true) {
      if (!(ref & kLeafBit)) {
        if (node_step<COUNT, EXACT>(s, r, best, stack, ref, cnt)) continue;
      } else {
        if (leaf_step<ANY, COUNT>(s, r, best, ref, oc, hit_prim, cnt)) return hit_prim;
      }
      if (!pop(ref)) break;
    }
  }
  return hit_prim;
}

// Makes a loaded value count as used at this point. A node's child
// references are needed only when a child box is hit; without this the
// compiler sinks their load below the box test, which costs a second
// dependent fetch of the node on every step. The empty asm moves nothing and
// reads no memory; the reference load then issues with the box loads.
__device__ __forceinline__ void use_now(uint4 &v) { asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w)); }

// One wide-node step: tests the (up to 4) child boxes with the conservative
// slab test, moves `ref` to the nearest hit child and pushes the other hit
// children farthest first (so they pop nearest first). Returns false when no
// child is hit.
template <bool COUNT, class Stack>
__device__ __forceinline__ bool wnode_step(const DevSceneView &s, const TraceRay &r, float best, Stack &stack,
                                           uint32_t &ref, TraceCounters &cnt) {
  const float4 *np = s.wnodes + 8 * size_t(ref);
  const float4 lx = ldg(np), hx = ldg(np + 1), ly = ldg(np + 2), hy = ldg(np + 3), lz = ldg(np + 4),
               hz = ldg(np + 5);
  uint4 rf = ldg(reinterpret_cast<const uint4 *>(np + 6));
  const uint32_t n = __float_as_uint(ldg(np + 7).x);
  if (COUNT) cnt.nodes++;
  constexpr float kInf = __builtin_inff();
  float t0, t1, t2, t3;
  const float bm = below(best);
  const bool h0 = box_hit(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, bm, t0);
  const bool h1 = box_hit(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, bm, t1) && n > 1;
  const bool h2 = box_hit(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, bm, t2) && n > 2;
  const bool h3 = box_hit(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, bm, t3) && n > 3;
  // (distance, ref) keys; misses sort last
  use_now(rf);
  float k0 = h0 ? t0 : kInf, k1 = h1 ? t1 : kInf, k2 = h2 ? t2 : kInf, k3 = h3 ? t3 : kInf;
  uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
  const int nhit = int(h0) + int(h1) + int(h2) + int(h3);
  if (COUNT) {
    cnt.btests += n;
    cnt.bhits += uint32_t(nhit);
  }
  if (nhit == 0) return false;
  auto cswap = [](float &ka, uint32_t &ra, float &kb, uint32_t &rb) {
    const bool sw = kb < ka;
    const float tk = sw ? kb : ka;
    const uint32_t tr = sw ? rb : ra;
    kb = sw ? ka : kb;
    rb = sw ? ra : rb;
    ka = tk;
    ra = tr;
  };
  // sorting network for 4 keys (ascending)
  cswap(k0, r0, k1, r1);
  cswap(k2, r2, k3, r3);
  cswap(k0, r0, k2, r2);
  cswap(k1, r1, k3, r3);
  cswap(k1, r1, k2, r2);
  stack.push_last(nhit - 1, make_uint2(r1, __float_as_uint(k1)), make_uint2(r2, __float_as_uint(k2)),
                  make_uint2(r3, __float_as_uint(k3)));
  ref = r0;
  return true;
}

// Node formats of the fast traversal mode.
constexpr int kNodesBinary = 2;  // binary child-boxes-in-parent nodes
constexpr int kNodesWide = 4;    // 4-wide nodes, float child boxes
constexpr int kNodesWideQ = 5;   // 4-wide nodes, 8-bit quantized child boxes

// Byte k of a packed plane word as a float (v_cvt_f32_ubyte<k>).
__device__ __forceinline__ float plane_q(uint32_t word, int k) { return float((word >> (8 * k)) & 255u); }

// One quantized wide-node step (as wnode_step). Per axis, the distance to
// the plane origin + q * 2^e is folded to one fused multiply-add,
//   t = q * A + B,  A = 2^e * inv (exact),  B = (origin - o) * inv,
// instead of decoding the plane and then subtracting and multiplying. The
// fold loses the plane's own rounding, so the host quantizes against the
// planes' exact real values; the rounding of B (two operations) and of the
// fma leaves |t' - t| <= ~3 eps |B| + eps |t|. The slack 2^-20 |B| per axis
// (down on entry planes, up on exit planes) covers the first term and the
// 2^-20 relative widening the second, so the test stays a superset of the
// exact one. Reciprocals are clamped to +-2^64: a zero direction component
// then gives finite plane distances whose sign says whether the origin lies
// in the slab (an infinite one would make inf - inf and lose the axis), and
// the per-axis slack keeps a ray that runs in an exit plane inside the box
// (B is then nonzero unless the child is flat at the node's origin, and a
// flat child holds only primitives parallel to that ray, which never hit).
// Empty slots hold an inverted box and a sentinel leaf (the host's
// quantizer), so no slot is skipped by the node's child count.
// ANY (an any-hit walk): the children are not ordered by distance. Until the
// walk returns at its first hit, `best` stays the ray's tmax, so every
// pushed child passes the pop test anyway; the hit children are moved ahead
// of the missed ones in child order by flags alone and pushed with key 0.
template <bool COUNT, bool ANY, class Stack>
__device__ __forceinline__ bool qnode_step(const DevSceneView &s, const TraceRay &r, float best, Stack &stack,
                                           uint32_t &ref, TraceCounters &cnt) {
  const uint4 *np = s.qnodes + 4 * size_t(ref);
  const uint4 n0 = ldg(np), n1 = ldg(np + 1), n2 = ldg(np + 2);
  uint4 rf = ldg(np + 3);
  if (COUNT) cnt.nodes++;
  const float ox = __uint_as_float(n0.x), oy = __uint_as_float(n0.y), oz = __uint_as_float(n0.z);
  const float sx = __uint_as_float((n0.w & 255u) << 23), sy = __uint_as_float(((n0.w >> 8) & 255u) << 23);
  const float sz = __uint_as_float(((n0.w >> 16) & 255u) << 23);
  const uint32_t n = n0.w >> 24;
  constexpr float kInf = __builtin_inff();
  // Per axis, the words holding the planes the ray enters (near) and leaves
  // (far) by: the min planes where the reciprocal direction is positive, the
  // max planes where it is negative, chosen once for the four children. A
  // NaN distance (zero direction component, origin on the plane) drops out
  // of the max / min below.
  const uint32_t mx = uint32_t(__float_as_int(r.inv.x) >> 31), my = uint32_t(__float_as_int(r.inv.y) >> 31),
                 mz = uint32_t(__float_as_int(r.inv.z) >> 31);
  const uint32_t nxw = (n1.w & mx) | (n1.x & ~mx), fxw = (n1.x & mx) | (n1.w & ~mx);
  const uint32_t nyw = (n2.x & my) | (n1.y & ~my), fyw = (n1.y & my) | (n2.x & ~my);
  const uint32_t nzw = (n2.y & mz) | (n1.z & ~mz), fzw = (n1.z & mz) | (n2.y & ~mz);
  const float bm = below(best);
  constexpr float kBig = 18446744073709551616.0f;  // 2^64
  const float ix = __builtin_amdgcn_fmed3f(r.inv.x, -kBig, kBig), iy = __builtin_amdgcn_fmed3f(r.inv.y, -kBig, kBig),
              iz = __builtin_amdgcn_fmed3f(r.inv.z, -kBig, kBig);
  const float Ax = sx * ix, Ay = sy * iy, Az = sz * iz;
  const float Bx = (ox - r.o.x) * ix, By = (oy - r.o.y) * iy, Bz = (oz - r.o.z) * iz;
  const float ex = fabsf(Bx) * kSlack, ey = fabsf(By) * kSlack, ez = fabsf(Bz) * kSlack;
  const float Bxl = Bx - ex, Bxh = Bx + ex, Byl = By - ey, Byh = By + ey, Bzl = Bz - ez, Bzh = Bz + ez;
  float t[4];
  bool h[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float ax = __builtin_fmaf(plane_q(nxw, k), Ax, Bxl), bx = __builtin_fmaf(plane_q(fxw, k), Ax, Bxh);
    const float ay = __builtin_fmaf(plane_q(nyw, k), Ay, Byl), by = __builtin_fmaf(plane_q(fyw, k), Ay, Byh);
    const float az = __builtin_fmaf(plane_q(nzw, k), Az, Bzl), bz = __builtin_fmaf(plane_q(fzw, k), Az, Bzh);
    const float l = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f)) * kWidenLo;
    const float hh = fminf(fminf(bx, by), bz) * kWidenHi;
    t[k] = l;
    h[k] = l <= fminf(hh, bm);
  }
  use_now(rf);
  const int nhit = int(h[0]) + int(h[1]) + int(h[2]) + int(h[3]);
  if (COUNT) {
    cnt.btests += n;
    cnt.bhits += uint32_t(nhit);
  }
  if (nhit == 0) return false;
  uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
  if constexpr (ANY) {
    bool h0 = h[0], h1 = h[1], h2 = h[2], h3 = h[3];
    // moves a hit child in front of a missed one (a 0/1 sorting network)
    auto cpart = [](bool &ha, uint32_t &ra, bool &hb, uint32_t &rb) {
      const bool sw = hb && !ha;
      const uint32_t tr = sw ? rb : ra;
      rb = sw ? ra : rb;
      ra = tr;
      const bool o = ha || hb;
      hb = ha && hb;
      ha = o;
    };
    cpart(h0, r0, h1, r1);
    cpart(h2, r2, h3, r3);
    cpart(h0, r0, h2, r2);
    cpart(h1, r1, h3, r3);
    cpart(h1, r1, h2, r2);
    stack.push_last(nhit - 1, make_uint2(r1, 0u), make_uint2(r2, 0u), make_uint2(r3, 0u));
    ref = r0;
    return true;
  } else {
    float k0 = h[0] ? t[0] : kInf, k1 = h[1] ? t[1] : kInf, k2 = h[2] ? t[2] : kInf, k3 = h[3] ? t[3] : kInf;
    auto cswap = [](float &ka, uint32_t &ra, float &kb, uint32_t &rb) {
      const bool sw = kb < ka;
      const float tk = sw ? kb : ka;
      const uint32_t tr = sw ? rb : ra;
      kb = sw ? ka : kb;
      rb = sw ? ra : rb;
      ka = tk;
      ra = tr;
    };
    cswap(k0, r0, k1, r1);
    cswap(k2, r2, k3, r3);
    cswap(k0, r0, k2, r2);
    cswap(k1, r1, k3, r3);
    cswap(k1, r1, k2, r2);
    stack.push_last(nhit - 1, make_uint2(r1, __float_as_uint(k1)), make_uint2(r2, __float_as_uint(k2)),
                    make_uint2(r3, __float_as_uint(k3)));
    ref = r0;
    return true;
  }
}

// One internal-node step of the fast mode for node format NODES.
template <int NODES, bool COUNT, bool ANY, class Stack>
__device__ __forceinline__ bool fast_step(const DevSceneView &s, const TraceRay &r, float best, Stack &stack,
                                          uint32_t &ref, TraceCounters &cnt) {
  if constexpr (NODES == kNodesWideQ) return qnode_step<COUNT, ANY>(s, r, best, stack, ref, cnt);
  else if constexpr (NODES == kNodesWide) return wnode_step<COUNT>(s, r, best, stack, ref, cnt);
  else return node_step<COUNT, false>(s, r, best, stack, ref, cnt);
}

// Closest / any hit over the wide BVH (fast mode only: children are visited
// nearest first, so ties may resolve differently from the reference's order).
template <int NODES, bool ANY, bool COUNT, bool WW, class Stack>
__device__ __forceinline__ int32_t trace_bvh4_s(const DevSceneView &s, const TraceRay &r, float &best, Stack &stack,
                                                TraceCounters &cnt) {
  int32_t hit_prim = -1;
  ObjCache oc;
  oc.tid = -1;
  uint32_t ref = 0;  // root: wide node 0
  if (WW) {
    while (true) {
      bool alive = true;
      while (!(ref & kLeafBit)) {
        if (!fast_step<NODES, COUNT, ANY>(s, r, best, stack, ref, cnt) && !pop_next(stack, best, ref)) {
          alive = false;
          break;
        }
      }
      if (!alive) break;
      if (leaf_step<ANY, COUNT, true>(s, r, best, ref, oc, hit_prim, cnt)) return hit_prim;
      if (!pop_next(stack, best, ref)) break;
    }
  } else {
    while (true) {
      if (!(ref & kLeafBit)) {
        if (fast_step<NODES, COUNT, ANY>(s, r, best, stack, ref, cnt)) continue;
      } else {
        if (leaf_step<ANY, COUNT, true>(s, r, best, ref, oc, hit_prim, cnt)) return hit_prim;
      }
      if (!pop_next(stack, best, ref)) break;
    }
  }
  return hit_prim;
}

// Dispatch on the node format (kNodes*).
template <int WIDTH, bool ANY, bool COUNT, bool WW, class Stack>
__device__ __forceinline__ int32_t trace_fast(const DevSceneView &s, const TraceRay &r, float &best, Stack &stack,
                                              TraceCounters &cnt) {
  if constexpr (WIDTH == kNodesBinary) return trace_bvh_s<ANY, COUNT, false, WW>(s, r, best, stack, cnt);
  else return trace_bvh4_s<WIDTH, ANY, COUNT, WW>(s, r, best, stack, cnt);
}

// Resumable fast-mode traversal of one lane, for kernels that hand a lane a
// new ray as soon as its current one is finished instead of waiting for the
// slowest ray of the wavefront. round() walks internal nodes to the next leaf
// and tests it (one while-while step); it returns true when the ray is done.
template <int WIDTH, bool ANY, bool COUNT>
struct LaneTrace {
  TraceRay r;
  float best;
  int32_t hit;
  uint32_t ref;
  ObjCache oc;
  template <class Stack>
  __device__ __forceinline__ void start(const TraceRay &ray, float tmax, Stack &st) {
    r = ray;
    best = tmax;
    hit = -1;
    ref = 0;
    oc.tid = -1;
    st.reset();
  }
  template <class Stack>
  __device__ __forceinline__ bool round(const DevSceneView &s, Stack &st, TraceCounters &cnt) {
    while (!(ref & kLeafBit)) {
      if (!fast_step<WIDTH, COUNT, ANY>(s, r, best, st, ref, cnt) && !pop_next(st, best, ref)) return true;
    }
    if (leaf_step<ANY, COUNT, WIDTH != kNodesBinary>(s, r, best, ref, oc, hit, cnt)) return true;
    return !pop_next(st, best, ref);
  }
};

// Fixed LDS stack at stack[k * stride], k < STACK (STACK >= BVH depth).
template <int STACK, bool ANY, bool COUNT, bool EXACT = false>
__device__ __forceinline__ int32_t trace_bvh(const DevSceneView &s, const TraceRay &r, float &best, uint2 *stack,
                                             int stride, TraceCounters &cnt) {
  LdsStack<STACK> st(stack, stride);
  return trace_bvh_s<ANY, COUNT, EXACT, false>(s, r, best, st, cnt);
}

}  // namespace muonx

#endif
