// Batched closest-hit / any-hit queries over device-resident rays
// (muonx_intersect_closest / muonx_intersect_any).
#include "query.h"
#include "runtime.h"
#include "trace.h"
#include "traverse_launch.h"
#include "wave.h"

#include <mutex>
#include <string>

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

// Fast mode with per-lane ray fetch: a lane whose ray is finished takes the
// next unclaimed ray once at least `fetch_min` lanes of its wavefront are
// idle, so long rays no longer hold idle lanes (claims: WaveFetch).

template <int KIND, int WIDTH, bool ANY, bool COUNT>
__global__ void __launch_bounds__(kQueryBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_query_dyn(DevSceneView s, StackArgs sa, const float *__restrict__ ro, const float *__restrict__ rd,
                const float *__restrict__ tmax, uint32_t n, int32_t *__restrict__ prim_out, float *__restrict__ t_out,
                uint8_t *__restrict__ occ_out, unsigned long long *__restrict__ counters, uint32_t *next,
                int fetch_min) {
  __shared__ uint2 lds[StackOf<KIND>::kLds * kQueryBlock];
  TraceCounters cnt{0, 0};
  auto st = make_stack<KIND>(lds, kQueryBlock, sa);
  LaneTrace<WIDTH, ANY, COUNT> lt;
  uint32_t idx = 0;
  // chunks of at most 128 rays: 2^24 closest-hit rays over the level-7
  // icosphere trace in 3.245 ms against 3.285 ms with 256 (and 3.41 ms with
  // 512, 3.45 ms with 64); the engine's kernels keep 256, with which the
  // path-traced render of the same scene is 4 % faster than with 128
  WaveFetch wf(n, 128);
  bool have = false, drained = false;
  while (true) {
    const bool need = !have && !drained;
    const uint64_t needm = __ballot(need);
    if (needm && (__popcll(needm) >= fetch_min || __ballot(have) == 0)) {
      const uint32_t k = wf.claim(needm, next);
      if (need) {
        if (k < n) {
          idx = k;
          const V3 o{ro[3 * k], ro[3 * k + 1], ro[3 * k + 2]};
          const V3 d{rd[3 * k], rd[3 * k + 1], rd[3 * k + 2]};
          lt.start(make_ray(o, d), tmax ? tmax[k] : __builtin_inff(), st);
          have = true;
        } else {
          drained = true;
        }
      }
    }
    if (__ballot(have) == 0) break;
    if (have && lt.round(s, st, cnt)) {
      if (ANY) {
        occ_out[idx] = lt.hit >= 0 ? 1 : 0;
      } else {
        prim_out[idx] = lt.hit >= 0 ? s.prim_id[lt.hit] : -1;
        if (t_out) t_out[idx] = lt.hit >= 0 ? lt.best : 0.0f;
      }
      have = false;
    }
  }
  if (COUNT) {
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
template <bool ANY, bool COUNT>
struct QueryDynLaunch {
  const TraverseCfg &cfg;
  const DevSceneView &s;
  const float *ro, *rd, *tmax;
  uint32_t n;
  int32_t *prim;
  float *t;
  uint8_t *occ;
  unsigned long long *counters;
  uint32_t *next;
  int fetch_min;
  hipStream_t stream;
  template <int KIND, bool WW, int WIDTH>
  void run() const {
    const StackArgs sa{cfg.spill, cfg.gcap};
    hipLaunchKernelGGL((k_query_dyn<KIND, WIDTH, ANY, COUNT>), dim3(cfg.grid), dim3(kQueryBlock), 0, stream, s, sa,
                       ro, rd, tmax, n, prim, t, occ, counters, next, fetch_min);
  }
};

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

// Static scratch of muonx_intersect_noalloc: the claim counter of the
// per-lane fetch kernels (the first 256 bytes) and the traversal stacks'
// spill region (the rest). Every GPU holds its own copy until the process
// ends; its device address is looked up once per device by
// InitStaticScratch (with the device current), never used by name on the
// host.
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
constexpr size_t kStaticScratchHead = 256;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

void *StaticScratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }
}  // namespace

int InitStaticScratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices) {
    SetError("static scratch: device index " + std::to_string(device) + " above the table of " +
             std::to_string(kMaxScratchDevices));
    return MUONX_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return MUONX_OK;
  void *p = nullptr;
  MX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = p;
  return MUONX_OK;
}

// Launches the query kernels. With `noalloc` the spill region and the claim
// counter come from the static scratch (the grid is capped so that the spill
// fits), the counter is cleared by a kernel rather than a memset node, and
// no event or buffer is created.
int LaunchQuery(muonx_ctx *ctx, const DevSceneView &s, bool any, bool exact, const float *ro, const float *rd,
                const float *tmax, uint32_t n, int32_t *prim, float *t, uint8_t *occ, unsigned long long *counters,
                hipStream_t stream, bool noalloc) {
  if (n == 0) return MUONX_OK;
  TraverseCfg cfg;
  unsigned char *sbase = nullptr;
  int rc;
  if (noalloc) {
    sbase = static_cast<unsigned char *>(StaticScratch(ctx->device));
    if (!sbase) {
      SetError("muonx_intersect_noalloc: no static scratch address for device " + std::to_string(ctx->device));
      return MUONX_ERR_STATE;
    }
    rc = PlanTraverseShape(ctx->tuning, ctx->num_cus, s.bvh_depth, s.wstack, n, kQueryBlock, cfg, exact);
    if (rc != MUONX_OK) return rc;
    if (cfg.gcap > 0) {
      const size_t per_block = size_t(kQueryBlock) * size_t(cfg.gcap) * sizeof(uint2);
      const size_t max_blocks = (kStaticScratchBytes - kStaticScratchHead) / per_block;
      if (max_blocks == 0) {
        SetError("muonx_intersect_noalloc: a traversal stack of " + std::to_string(cfg.kind + cfg.gcap) +
                 " entries per lane needs " + std::to_string(per_block) + " bytes of spill per block, more than the " +
                 std::to_string(kStaticScratchBytes - kStaticScratchHead) + "-byte static scratch");
        return MUONX_ERR_CAPACITY;
      }
      cfg.grid = int(std::min<size_t>(size_t(cfg.grid), max_blocks));
      cfg.threads = cfg.grid * kQueryBlock;
      cfg.spill = reinterpret_cast<uint2 *>(sbase + kStaticScratchHead);
    }
  } else {
    rc = PlanTraverse(ctx, s, n, kQueryBlock, exact, cfg);
    if (rc != MUONX_OK) return rc;
    if ((rc = BeginSpill(ctx, cfg, stream)) != MUONX_OK) return rc;
  }
#define MX_Q(A, C, E) DispatchTraverse(cfg, QueryLaunch<A, C, E>{cfg, s, ro, rd, tmax, n, prim, t, occ, counters, stream})
#define MX_QE(A, C) DispatchExact(cfg, QueryLaunch<A, C, true>{cfg, s, ro, rd, tmax, n, prim, t, occ, counters, stream})
  const bool cnt = counters != nullptr;
  const int fetch = ctx->tuning.ray_fetch;
  if (!exact && fetch > 0 && n <= (1u << 31)) {  // the claim counter stays below 2^32 (n + chunk x waves)
    // per-lane ray fetch; the claim counter is ordered like the spill region
    uint32_t *next;
    if (noalloc) {
      next = reinterpret_cast<uint32_t *>(sbase);
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, next);
    } else {
      MX_HIP(ctx->fetch_ctr.Reserve(sizeof(uint32_t)));
      if (!ctx->spill_done) MX_HIP(hipEventCreateWithFlags(&ctx->spill_done, hipEventDisableTiming));
      MX_HIP(hipStreamWaitEvent(stream, ctx->spill_done, 0));
      MX_HIP(hipMemsetAsync(ctx->fetch_ctr.ptr, 0, sizeof(uint32_t), stream));
      next = ctx->fetch_ctr.as<uint32_t>();
    }
#define MX_QD(A, C) DispatchTraverse(cfg, QueryDynLaunch<A, C>{cfg, s, ro, rd, tmax, n, prim, t, occ, counters, next, fetch, stream})
    if (any) {
      if (cnt) MX_QD(true, true);
      else MX_QD(true, false);
    } else {
      if (cnt) MX_QD(false, true);
      else MX_QD(false, false);
    }
#undef MX_QD
    MX_HIP(hipGetLastError());
    if (!noalloc) MX_HIP(hipEventRecord(ctx->spill_done, stream));
    return MUONX_OK;
  }
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
  return noalloc ? MUONX_OK : EndSpill(ctx, cfg, stream);
}

}  // namespace muonx

using namespace muonx;

namespace {
int CheckQuery(const char *fn, muonx_scene *scene, const muonx_query *q) {
  if (!scene || !q) {
    SetError(std::string(fn) + ": null argument");
    return MUONX_ERR_INVALID;
  }
  if (q->n && (!q->ray_o || !q->ray_d || (q->any_hit ? !q->occluded : !q->prim_id))) {
    SetError(std::string(fn) + ": missing ray or output array");
    return MUONX_ERR_INVALID;
  }
  if (q->mode != MUONX_TRAVERSAL_FAST && q->mode != MUONX_TRAVERSAL_EXACT) {
    SetError(std::string(fn) + ": unknown traversal mode");
    return MUONX_ERR_INVALID;
  }
  return MUONX_OK;
}

int RunQuery(muonx_scene *scene, const muonx_query *q, void *stream, bool noalloc) {
  if (q->mode == MUONX_TRAVERSAL_EXACT && !scene->reference_tree) {
    SetError("exact traversal needs the reference's SAH tree; this scene's BVH was built on the device");
    return MUONX_ERR_STATE;
  }
  std::lock_guard<std::mutex> lock(scene->ctx->mu);
  MX_HIP(hipSetDevice(scene->ctx->device));
  return LaunchQuery(scene->ctx, scene->view, q->any_hit != 0, q->mode == MUONX_TRAVERSAL_EXACT, q->ray_o, q->ray_d,
                     q->tmax, q->n, q->prim_id, q->t, q->occluded, q->counters, static_cast<hipStream_t>(stream),
                     noalloc);
}
}  // namespace

extern "C" int muonx_intersect(muonx_scene *scene, const muonx_query *q, void *stream) {
  const int rc = CheckQuery("muonx_intersect", scene, q);
  return rc != MUONX_OK ? rc : RunQuery(scene, q, stream, false);
}

extern "C" int muonx_intersect_noalloc(muonx_scene *scene, const muonx_query *q, void *stream) {
  const int rc = CheckQuery("muonx_intersect_noalloc", scene, q);
  return rc != MUONX_OK ? rc : RunQuery(scene, q, stream, true);
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
