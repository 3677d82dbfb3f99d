// Launch-side selection of the traversal variant (stack policy, loop form)
// from a context's tuning, shared by the query and path-tracer kernels.
#ifndef MUONX_DEVICE_TRAVERSE_LAUNCH_H_
#define MUONX_DEVICE_TRAVERSE_LAUNCH_H_

#include <algorithm>

#include "runtime.h"
#include "trace.h"

namespace muonx {

// Stack variants: kinds 24/48/96 = full LDS stack of that capacity, 8/16 =
// LDS ring with HBM spill.
struct TraverseCfg {
  int kind;     // 24, 48, 96, 8 or 16
  bool ww;
  int width;    // node format: kNodesBinary, kNodesWide, kNodesWideQ
  int grid;     // persistent blocks
  int threads;  // grid * block
  uint2 *spill;
  int gcap;     // spill entries per thread
};

// Everything a kernel needs to build its stack.
struct StackArgs {
  uint2 *spill;
  int gcap;
};

template <int KIND>
struct StackOf;
template <>
struct StackOf<24> { using T = LdsStack<24>; static constexpr int kLds = 24; };
template <>
struct StackOf<48> { using T = LdsStack<48>; static constexpr int kLds = 48; };
template <>
struct StackOf<96> { using T = LdsStack<96>; static constexpr int kLds = 96; };
template <>
struct StackOf<8> { using T = RingStack<8>; static constexpr int kLds = 8; };
template <>
struct StackOf<16> { using T = RingStack<16>; static constexpr int kLds = 16; };

template <int KIND>
__device__ __forceinline__ typename StackOf<KIND>::T make_stack(uint2 *lds, int block, const StackArgs &sa) {
  if constexpr (KIND == 8 || KIND == 16) {
    const int gtid = int(blockIdx.x) * block + int(threadIdx.x);
    const int nthreads = int(gridDim.x) * block;
    return typename StackOf<KIND>::T(lds + threadIdx.x, block, sa.spill + gtid, nthreads, sa.gcap);
  } else {
    return typename StackOf<KIND>::T(lds + threadIdx.x, block);
  }
}

// Resolves the variant for a scene (binary depth, wide-stack bound): node
// format, grid, stack kind and the spill entries per thread (cfg.gcap; the
// spill region then needs cfg.threads * cfg.gcap uint2, not yet assigned).
// `want_exact` selects the exact mode: binary nodes and full LDS stacks (the
// deterministic modes are not performance-critical and keep the simplest
// form). Allocates nothing.
inline int PlanTraverseShape(const muonx_tuning &t, int num_cus, uint32_t bvh_depth, uint32_t wstack,
                             uint64_t work_items, int block, TraverseCfg &cfg, bool want_exact = false) {
  const int cus = num_cus > 0 ? num_cus : 256;
  cfg.width = want_exact || t.bvh_width != 4 ? kNodesBinary : (t.node_bits == 8 ? kNodesWideQ : kNodesWide);
  const int bpc = std::max(1, t.blocks_per_cu);
  cfg.grid = int(std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(cus) * bpc, (work_items + block - 1) / block)));
  cfg.threads = cfg.grid * block;
  cfg.ww = t.while_while != 0 && !want_exact;
  cfg.spill = nullptr;
  cfg.gcap = 0;
  // binary: entries never exceed depth - 1; wide: the collapse's bound
  const int need = int(cfg.width != kNodesBinary ? std::max<uint32_t>(wstack, 1) : bvh_depth);
  const bool ring = !want_exact && (t.stack_entries == 8 || t.stack_entries == 16);
  if (ring && need <= t.stack_entries) {
    cfg.kind = need <= 8 ? 8 : 16;  // a ring that never spills
    return MUONX_OK;
  }
  if (ring || (!want_exact && need > 96)) {
    cfg.kind = ring ? t.stack_entries : 16;
    cfg.gcap = need - cfg.kind;
    return MUONX_OK;
  }
  if (need <= 24) cfg.kind = 24;
  else if (need <= 48) cfg.kind = 48;
  else if (need <= 96) cfg.kind = 96;
  else {
    SetError("BVH depth " + std::to_string(need) + " exceeds the exact-mode traversal stack (96)");
    return MUONX_ERR_CAPACITY;
  }
  return MUONX_OK;
}

// PlanTraverseShape, then reserves the spill region in `spill`.
inline int PlanTraverseInto(const muonx_tuning &t, int num_cus, uint32_t bvh_depth, uint32_t wstack,
                            uint64_t work_items, int block, DevBuf &spill, TraverseCfg &cfg, bool want_exact = false) {
  const int rc = PlanTraverseShape(t, num_cus, bvh_depth, wstack, work_items, block, cfg, want_exact);
  if (rc != MUONX_OK || cfg.gcap == 0) return rc;
  MX_HIP(spill.Reserve(size_t(cfg.threads) * cfg.gcap * sizeof(uint2)));
  cfg.spill = spill.as<uint2>();
  return MUONX_OK;
}

inline int PlanTraverse(muonx_ctx *ctx, const DevSceneView &s, uint64_t work_items, int block, bool want_exact,
                        TraverseCfg &cfg) {
  return PlanTraverseInto(ctx->tuning, ctx->num_cus, s.bvh_depth, s.wstack, work_items, block, ctx->spill, cfg,
                          want_exact);
}

// Kernels that use the spill region wait for its previous user when issued
// on another stream.
inline int BeginSpill(muonx_ctx *ctx, const TraverseCfg &cfg, hipStream_t stream) {
  if (!cfg.spill) return MUONX_OK;
  if (!ctx->spill_done) MX_HIP(hipEventCreateWithFlags(&ctx->spill_done, hipEventDisableTiming));
  MX_HIP(hipStreamWaitEvent(stream, ctx->spill_done, 0));
  return MUONX_OK;
}
inline int EndSpill(muonx_ctx *ctx, const TraverseCfg &cfg, hipStream_t stream) {
  if (!cfg.spill) return MUONX_OK;
  MX_HIP(hipEventRecord(ctx->spill_done, stream));
  return MUONX_OK;
}

// Calls F::template run<KIND, WW, WIDTH>() for the planned variant.
template <int W, class F>
inline void DispatchKind(const TraverseCfg &cfg, F &&f) {
  switch (cfg.kind) {
    case 8: cfg.ww ? f.template run<8, true, W>() : f.template run<8, false, W>(); break;
    case 16: cfg.ww ? f.template run<16, true, W>() : f.template run<16, false, W>(); break;
    case 24: cfg.ww ? f.template run<24, true, W>() : f.template run<24, false, W>(); break;
    case 48: cfg.ww ? f.template run<48, true, W>() : f.template run<48, false, W>(); break;
    default: cfg.ww ? f.template run<96, true, W>() : f.template run<96, false, W>(); break;
  }
}
template <class F>
inline void DispatchTraverse(const TraverseCfg &cfg, F &&f) {
  if (cfg.width == kNodesWideQ) DispatchKind<kNodesWideQ>(cfg, f);
  else if (cfg.width == kNodesWide) DispatchKind<kNodesWide>(cfg, f);
  else DispatchKind<kNodesBinary>(cfg, f);
}
// Exact mode: binary nodes, full LDS stack, if-if loop.
template <class F>
inline void DispatchExact(const TraverseCfg &cfg, F &&f) {
  switch (cfg.kind) {
    case 24: f.template run<24, false, 2>(); break;
    case 48: f.template run<48, false, 2>(); break;
    default: f.template run<96, false, 2>(); break;
  }
}

}  // namespace muonx

#endif
