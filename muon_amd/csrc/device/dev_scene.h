// Device-resident scene layout (HBM, struct-of-arrays).
//
// BVH: the host's binary SAH tree with each node's two child boxes stored in
// the parent ("child boxes in parent"), 64 B per internal node = 4 x float4:
//   n0 = (c0.min.x, c0.max.x, c0.min.y, c0.max.y)
//   n1 = (c1.min.x, c1.max.x, c1.min.y, c1.max.y)
//   n2 = (c0.min.z, c0.max.z, c1.min.z, c1.max.z)
//   n3 = (ref0, ref1, meta0, meta1) as ints
// ref_i is the child's node index, or for a leaf the first primitive (leaf
// order). meta_i bit 31 marks a leaf, bits 0..28 its primitive count;
// bits 29..30 of meta0 hold this node's split axis. One node fetch tests both
// children, and the near child (reference order: children[dir[axis] < 0]
// first, muon/acceleration.cc:77-82,134-138) is chosen from the axis.
//
// Wide BVH: the binary tree collapsed to 4-ary nodes of least SAH cost
// (wide_bvh.h), 128 B per node = 8 x float4:
//   w0 = child min.x[0..3]  w1 = child max.x[0..3]
//   w2 = child min.y[0..3]  w3 = child max.y[0..3]
//   w4 = child min.z[0..3]  w5 = child max.z[0..3]
//   w6 = child refs[0..3] (node index, or a leaf reference, below)
//   w7 = (child count, 0, 0, 0)
// Leaves are primitive ranges of the binary tree: its leaves, or (when leaf
// references carry counts) whole binary subtrees of a few primitives.
//
// Quantized wide BVH: the same nodes in 64 B = 4 x uint4, child planes as
// 8-bit offsets from a per-node origin in steps of a per-axis power of two:
//   q0 = (origin.xyz, biased exponents x | y << 8 | z << 16 | count << 24)
//   q1 = (min.x[4], min.y[4], min.z[4], max.x[4]) one byte per child
//   q2 = (max.y[4], max.z[4], 0, 0)
//   q3 = child refs[4]
// A decoded plane is origin + q * 2^e; every decoded box contains the float
// box.
//
// Primitives (leaf order), 48 B = 3 x float4:
//   tri:    (v0.xyz, tid)  (v1.xyz, 0)  (v2.xyz, 0)     object space
//   sphere: (c.xyz, tid | kSphereBit)  (r, 0, 0, 0)  (0)
// tid indexes 48-float transform records (M, M^-1, M^-T column-major).
#ifndef MUONX_DEVICE_DEV_SCENE_H_
#define MUONX_DEVICE_DEV_SCENE_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bvh_layout.h"

namespace muonx {

constexpr uint32_t kCountMask = 0x1fffffffu;
constexpr uint32_t kMaxLightPrims = 16;
constexpr uint32_t kSphereBit = 0x80000000u;

// Material record, 16 floats: ambient3 diffuse3 specular3 emission3
// shininess roughness brdf reflectiveness (precomputed, muon/materials.cc:113-123,270-293).
constexpr int kMatStride = 16;
// Light record, 16 floats: kind color3 a3 b3 c3 area(quad) — see muonx_scene_desc.
constexpr int kLightStride = 16;

struct DevSceneView {
  const float4 *nodes;      // 4 per internal node
  const float4 *wnodes;     // 8 per wide node
  const uint4 *qnodes;      // 4 per wide node, 8-bit quantized child boxes
  const float4 *prims;      // 3 per primitive, leaf order
  const float *xforms;      // 48 per transform
  const int32_t *prim_id;   // leaf order -> parse order
  const int32_t *prim_mat;  // leaf order -> material
  const int32_t *prim_light;
  const float *prim_vn;     // 9 per primitive (leaf order) or nullptr
  const uint8_t *prim_use_vn;
  const float *materials;
  const float *lights;
  uint32_t num_prims;
  uint32_t num_nodes;
  uint32_t num_lights;
  uint32_t num_quad_lights;
  uint32_t bvh_depth;
  uint32_t num_wnodes;
  uint32_t wstack;          // bound on stack entries of a wide traversal
  uint32_t leaf_counts;     // 1: wide leaf references carry their primitive count
  uint32_t root_leaf;       // 1: the BVH root is a leaf (node 0 holds it and a sentinel leaf)
  uint32_t linear;          // 1: the linear structure (one leaf in parse order)
  // Emitting primitives (prim_light >= 0), leaf order, when there are at
  // most kMaxLightPrims of them (cull_mis = 1): a MIS ray that hits none of
  // them needs no traversal.
  const uint32_t *light_prims;
  uint32_t num_light_prims;
  uint32_t cull_mis;
  // World box around those primitives (min xyz, max xyz), padded so that a
  // ray that hits one of them passes its slab test: a MIS ray that misses the
  // box skips their primitive tests.
  float light_box[6];
};

}  // namespace muonx

#endif
