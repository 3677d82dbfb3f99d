// Library-internal runtime objects behind the opaque C handles.
#ifndef MUONX_DEVICE_RUNTIME_H_
#define MUONX_DEVICE_RUNTIME_H_

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/muonx.h"
#include "dev_scene.h"

namespace muonx {

void SetError(const std::string &msg);

#define MX_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      ::muonx::SetError(std::string(#call) + ": " + hipGetErrorString(e_));                \
      return MUONX_ERR_HIP;                                                               \
    }                                                                                     \
  } while (0)

// Device allocation that frees itself.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { Free(); }
  void Free() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  hipError_t Alloc(size_t n) {
    Free();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&ptr, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  // Grows (without preserving contents) to at least n bytes.
  hipError_t Reserve(size_t n) {
    if (n <= bytes && ptr) return hipSuccess;
    return Alloc(n);
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
};

struct RenderWorkspace;

// Scenes get ids that are never reused, so caches keyed by id cannot mistake
// a new scene for a destroyed one at the same address.
uint64_t RegisterScene();
void UnregisterScene(uint64_t id);
bool SceneLive(uint64_t id);

}  // namespace muonx

struct muonx_ctx {
  int device = 0;
  std::mutex mu;
  muonx::RenderWorkspace *ws = nullptr;
  int num_cus = 0;
  muonx_tuning tuning{8, 1, 16, 8192, 4, 48, 8, 0};  // best of the sweeps on the level-7 icosphere
  muonx::DevBuf spill;      // per-thread traversal stack spill region
  muonx::DevBuf fetch_ctr;  // ray claim counter of the per-lane fetch kernels
  // Orders kernels that share `spill` when they are issued on different streams.
  hipEvent_t spill_done = nullptr;
};

struct muonx_scene {
  muonx_ctx *ctx = nullptr;
  uint64_t id = 0;
  std::atomic<int> inflight{0};  // server jobs that use the scene
  muonx::DevBuf nodes, wnodes, qnodes, prims, xforms, prim_id, prim_mat, prim_light, prim_vn, prim_use_vn, materials, lights;
  muonx::DevBuf light_prims;  // emitting primitives (leaf order) when few
  muonx::DevSceneView view{};
  uint32_t bvh_depth = 0;
  uint64_t device_bytes = 0;
  bool reference_tree = true;  // false: the BVH was built on the device (exact traversal refused)
  float build_ms = 0.0f;       // that device build
  float bounds[6] = {0, 0, 0, 0, 0, 0};  // world bounds: min xyz, max xyz
  std::vector<float> host_lights;     // 16 per light
  std::vector<float> host_materials;  // 16 per material
};

#endif
