// Reference encodings of the BVH layouts shared by the device traversal and
// the host code that builds the device trees (no HIP types, so host-only
// code and CPU tests can include it).
#ifndef MUONX_DEVICE_BVH_LAYOUT_H_
#define MUONX_DEVICE_BVH_LAYOUT_H_

#include <cstdint>

namespace muonx {

constexpr uint32_t kLeafBit = 0x80000000u;
// Leaf references of the wide trees when the scene's leaves carry their
// primitive count (DevSceneView::leaf_counts): leaf bit | (count - 1) << 24 |
// first primitive, for scenes of fewer than 2^24 primitives and leaves of at
// most 128. Otherwise leaf bit | first primitive, and a leaf ends at the
// primitive flagged last (as in the binary tree).
constexpr uint32_t kLeafCountShift = 24;
constexpr uint32_t kLeafFirstMask = (1u << kLeafCountShift) - 1u;
constexpr uint32_t kLeafMaxCount = 128;

}  // namespace muonx

#endif
