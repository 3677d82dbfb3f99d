// Batched closest-hit / any-hit queries over device-resident rays
// (muonx_intersect_closest / muonx_intersect_any).
#include "query.h"
#include "runtime.h"
#include "trace.h"
#include "traverse_launch.h"

namespace muonx {

template <int KIND, bool WW, int WIDTH, bool ANY, bool COUNT, bool EXACT>
__global__ void __launch_bounds__(kQueryBlock) k_query(DevSceneView s, StackArgs sa, const float *__restrict__ ro,
                                                       const float *__restrict__ rd, const float *__restrict__ tmax,
                                                       uint32_t n, int32_t *__restrict__ prim_out,
                                                       float *__restrict__ t_out, uint8_t *__restrict__ occ_out,
                                                       unsigned long long *__restrict__ counters) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kQueryBlock];
  TraceCounters cnt{0, 0};
  for (uint32_t i = blockIdx.x * kQueryBlock + threadIdx.x; i < n; i += gridDim.x * kQueryBlock) {
    const V3 o{ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]};
    const V3 d{rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]};
    float best = tmax ? tmax[i] : __builtin_inff();
    const TraceRay r = make_ray(o, d);
    auto st = make_stack<KIND>(lds, kQueryBlock, sa);
    int32_t p;
    if constexpr (EXACT) p = trace_bvh_s<ANY, COUNT, true, false>(s, r, best, st, cnt);
    else p = trace_fast<WIDTH, ANY, COUNT, WW>(s, r, best, st, cnt);
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

namespace {
template <bool ANY, bool COUNT, bool EXACT>
struct QueryLaunch {
  const TraverseCfg &cfg;
  const DevSceneView &s;
  const float *ro, *rd, *tmax;
  uint32_t n;
  int32_t *prim;
  float *t;
  uint8_t *occ;
  unsigned long long *counters;
  hipStream_t stream;
  template <int KIND, bool WW, int WIDTH>
  void run() const {
    const StackArgs sa{cfg.spill, cfg.gcap};
    hipLaunchKernelGGL((k_query<KIND, WW, WIDTH, ANY, COUNT, EXACT>), dim3(cfg.grid), dim3(kQueryBlock), 0, stream,
                       s, sa, ro, rd, tmax, n, prim, t, occ, counters);
  }
};
}  // namespace

int LaunchQuery(muonx_ctx *ctx, const DevSceneView &s, bool any, bool exact, const float *ro, const float *rd,
                const float *tmax, uint32_t n, int32_t *prim, float *t, uint8_t *occ, unsigned long long *counters,
                hipStream_t stream) {
  if (n == 0) return MUONX_OK;
  TraverseCfg cfg;
  int rc = PlanTraverse(ctx, s, n, kQueryBlock, exact, cfg);
  if (rc != MUONX_OK) return rc;
  if ((rc = BeginSpill(ctx, cfg, stream)) != MUONX_OK) return rc;
#define MX_Q(A, C, E) DispatchTraverse(cfg, QueryLaunch<A, C, E>{cfg, s, ro, rd, tmax, n, prim, t, occ, counters, stream})
#define MX_QE(A, C) DispatchExact(cfg, QueryLaunch<A, C, true>{cfg, s, ro, rd, tmax, n, prim, t, occ, counters, stream})
  const bool cnt = counters != nullptr;
  if (exact) {
    // exact mode: full LDS stack, one node or leaf step per iteration
    if (any) MX_QE(true, false);
    else if (cnt) MX_QE(false, true);
    else MX_QE(false, false);
  } else if (any) {
    if (cnt) MX_Q(true, true, false);
    else MX_Q(true, false, false);
  } else {
    if (cnt) MX_Q(false, true, false);
    else MX_Q(false, false, false);
  }
#undef MX_Q
#undef MX_QE
  MX_HIP(hipGetLastError());
  return EndSpill(ctx, cfg, stream);
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
  std::lock_guard<std::mutex> lock(scene->ctx->mu);
  MX_HIP(hipSetDevice(scene->ctx->device));
  return LaunchQuery(scene->ctx, scene->view, q->any_hit != 0, q->mode == MUONX_TRAVERSAL_EXACT, q->ray_o, q->ray_d, q->tmax,
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
