#ifndef MUONX_DEVICE_QUERY_H_
#define MUONX_DEVICE_QUERY_H_

#include <hip/hip_runtime.h>

#include "dev_scene.h"
#include "runtime.h"

namespace muonx {

constexpr int kQueryBlock = 128;

int LaunchQuery(muonx_ctx *ctx, const DevSceneView &s, bool any, bool exact, const float *ro, const float *rd,
                const float *tmax, uint32_t n, int32_t *prim, float *t, uint8_t *occ, unsigned long long *counters,
                hipStream_t stream, bool noalloc = false);

// Looks up the device address of the static scratch of
// muonx_intersect_noalloc for `device` (which must be current); once per
// device, at context creation.
int InitStaticScratch(int device);

}  // namespace muonx

#endif
