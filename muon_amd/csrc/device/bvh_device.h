// Device-side BVH build (bvh_device.hip) for scenes handed over without a
// host-built tree.
#ifndef MUONX_DEVICE_BVH_DEVICE_H_
#define MUONX_DEVICE_BVH_DEVICE_H_

#include <cstdint>
#include <vector>

#include "runtime.h"

namespace muonx {

// A binary BVH in the scene descriptor's layout: nodes in pre-order, leaves
// of one primitive each, `leaf` the parse-order ids in leaf order.
struct DeviceBvh {
  std::vector<float> bounds;  // 6 per node
  std::vector<int32_t> left, right, axis, start, count, leaf;
  uint32_t depth = 0;
  float device_ms = 0.0f;  // the build's kernels and sort
  float total_ms = 0.0f;   // with the copies and the host's pre-order pass
};

// Builds the BVH of the descriptor's primitives (world boxes of their
// transformed geometry) on the context's device.
int BuildDeviceBvh(muonx_ctx *ctx, const muonx_scene_desc *d, DeviceBvh *out);

}  // namespace muonx

#endif
