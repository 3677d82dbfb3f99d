// Batched closest-hit / any-hit queries over device-resident rays
// (muonx_intersect_closest / muonx_intersect_any).
#include "query.h"
#include "runtime.h"
#include "trace.h"

namespace muonx {

template <int STACK, bool ANY, bool COUNT, bool EXACT>
__global__ void __launch_bounds__(kQueryBlock) k_query(DevSceneView s, const float *__restrict__ ro,
                                                       const float *__restrict__ rd, const float *__restrict__ tmax,
                                                       uint32_t n, int32_t *__restrict__ prim_out,
                                                       float *__restrict__ t_out, uint8_t *__restrict__ occ_out,
                                                       unsigned long long *__restrict__ counters) {
  __shared__ uint2 stack[STACK * kQueryBlock];
  const uint32_t i = blockIdx.x * kQueryBlock + threadIdx.x;
  TraceCounters cnt{0, 0};
  if (i < n) {
    const V3 o{ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
    const V3 d{rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
    float best = tmax ? tmax[i] : __builtin_inff();
    const TraceRay r = make_ray(o, d);
    const int32_t p = trace_bvh<STACK, ANY, COUNT, EXACT>(s, r, best, stack + threadIdx.x, kQueryBlock, cnt);
    if (ANY) {
      occ_out[i] = p >= 0 ? 1 : 0;
    } else {
      prim_out[i] = p >= 0 ? s.prim_id[p] : -1;
      if (t_out) t_out[i] = p >= 0 ? best : 0.0f;
    }
  }
  if (COUNT) {
    // wave-level reduction, one atomic per wave
    unsigned long long a = cnt.nodes, b = cnt.prims;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off);
      b += __shfl_down(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&counters[0], a);
      atomicAdd(&counters[1], b);
    }
  }
}

template <bool ANY, bool COUNT, bool EXACT>
int LaunchQueryT(const DevSceneView &s, const float *ro, const float *rd, const float *tmax, uint32_t n,
                 int32_t *prim, float *t, uint8_t *occ, unsigned long long *counters, hipStream_t stream) {
  if (n == 0) return MUONX_OK;
  const dim3 grid((n + kQueryBlock - 1) / kQueryBlock);
  const uint32_t need = s.bvh_depth;  // stack entries <= internal ancestors < depth
  if (need <= 24) {
    hipLaunchKernelGGL((k_query<24, ANY, COUNT, EXACT>), grid, dim3(kQueryBlock), 0, stream, s, ro, rd, tmax, n, prim, t,
                       occ, counters);
  } else if (need <= 48) {
    hipLaunchKernelGGL((k_query<48, ANY, COUNT, EXACT>), grid, dim3(kQueryBlock), 0, stream, s, ro, rd, tmax, n, prim, t,
                       occ, counters);
  } else if (need <= 96) {
    hipLaunchKernelGGL((k_query<96, ANY, COUNT, EXACT>), grid, dim3(kQueryBlock), 0, stream, s, ro, rd, tmax, n, prim, t,
                       occ, counters);
  } else {
    SetError("BVH depth " + std::to_string(need) + " exceeds the traversal stack (96)");
    return MUONX_ERR_CAPACITY;
  }
  MX_HIP(hipGetLastError());
  return MUONX_OK;
}

int LaunchQuery(const DevSceneView &s, bool any, bool exact, const float *ro, const float *rd, const float *tmax,
                uint32_t n, int32_t *prim, float *t, uint8_t *occ, unsigned long long *counters, hipStream_t stream) {
#define MX_Q(A, C, E) return LaunchQueryT<A, C, E>(s, ro, rd, tmax, n, prim, t, occ, counters, stream)
  const bool cnt = counters != nullptr;
  if (any) {
    if (exact) {
      if (cnt) MX_Q(true, true, true);
      MX_Q(true, false, true);
    }
    if (cnt) MX_Q(true, true, false);
    MX_Q(true, false, false);
  }
  if (exact) {
    if (cnt) MX_Q(false, true, true);
    MX_Q(false, false, true);
  }
  if (cnt) MX_Q(false, true, false);
  MX_Q(false, false, false);
#undef MX_Q
}

}  // namespace muonx

using namespace muonx;

extern "C" int muonx_intersect(muonx_scene *scene, const muonx_query *q, void *stream) {
  if (!scene || !q) {
    SetError("muonx_intersect: null argument");
    return MUONX_ERR_INVALID;
  }
  if (q->n && (!q->ray_o || !q->ray_d || (q->any_hit ? !q->occluded : !q->prim_id))) {
    SetError("muonx_intersect: missing ray or output array");
    return MUONX_ERR_INVALID;
  }
  if (q->mode != MUONX_TRAVERSAL_FAST && q->mode != MUONX_TRAVERSAL_EXACT) {
    SetError("muonx_intersect: unknown traversal mode");
    return MUONX_ERR_INVALID;
  }
  MX_HIP(hipSetDevice(scene->ctx->device));
  return LaunchQuery(scene->view, q->any_hit != 0, q->mode == MUONX_TRAVERSAL_EXACT, q->ray_o, q->ray_d, q->tmax,
                     q->n, q->prim_id, q->t, q->occluded, q->counters, static_cast<hipStream_t>(stream));
}

extern "C" int muonx_intersect_closest(muonx_scene *scene, const float *ro, const float *rd, const float *tmax,
                                       uint32_t n, int32_t *prim_id, float *t, void *stream) {
  muonx_query q{ro, rd, tmax, n, 0, MUONX_TRAVERSAL_FAST, prim_id, t, nullptr, nullptr};
  return muonx_intersect(scene, &q, stream);
}

extern "C" int muonx_intersect_any(muonx_scene *scene, const float *ro, const float *rd, const float *tmax,
                                   uint32_t n, uint8_t *occluded, void *stream) {
  muonx_query q{ro, rd, tmax, n, 1, MUONX_TRAVERSAL_FAST, nullptr, nullptr, occluded, nullptr};
  return muonx_intersect(scene, &q, stream);
}
