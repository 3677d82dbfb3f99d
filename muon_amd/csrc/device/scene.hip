// Scene upload: converts the host descriptor (parse-order primitives + the
// binary SAH BVH in pre-order) into the device layout of dev_scene.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../common/vecmath.h"
#include "bvh_device.h"
#include "runtime.h"
#include "wide_bvh.h"

namespace muonx {
namespace {

// World box around the first `n` primitives of `ids` (leaf order), padded by
// 1e-3 of its diagonal plus 1e-5 of its largest coordinate: the primitive
// tests run in object space, so a ray that hits one of them passes within
// rounding of its transformed vertices (a sphere: of its transformed object
// box). All zeros when n is 0.
void EmitterBox(const std::vector<float4> &prims, const std::vector<float> &xf, const std::vector<uint32_t> &ids,
                uint32_t n, float box[6]) {
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (uint32_t q = 0; q < n && q < ids.size(); ++q) {
    const float4 *p = &prims[size_t(ids[q]) * 3];
    uint32_t tw;
    std::memcpy(&tw, &p[0].w, 4);
    const float *m = &xf[size_t(tw & 0x7fffffffu) * 48];  // column-major M
    auto add = [&](double x, double y, double z) {
      const double w[3] = {m[0] * x + m[4] * y + m[8] * z + m[12], m[1] * x + m[5] * y + m[9] * z + m[13],
                           m[2] * x + m[6] * y + m[10] * z + m[14]};
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], w[a]);
        hi[a] = std::max(hi[a], w[a]);
      }
    };
    if (tw & kSphereBit) {
      const double r = p[1].x;
      for (int c = 0; c < 8; ++c)
        add(p[0].x + ((c & 1) ? r : -r), p[0].y + ((c & 2) ? r : -r), p[0].z + ((c & 4) ? r : -r));
    } else {
      add(p[0].x, p[0].y, p[0].z);
      add(p[1].x, p[1].y, p[1].z);
      add(p[2].x, p[2].y, p[2].z);
    }
  }
  if (!(lo[0] <= hi[0])) {
    for (int a = 0; a < 6; ++a) box[a] = 0.0f;
    return;
  }
  double diag = 0.0, big = 0.0;
  for (int a = 0; a < 3; ++a) {
    diag += (hi[a] - lo[a]) * (hi[a] - lo[a]);
    big = std::max(big, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
  }
  const double pad = 1e-3 * std::sqrt(diag) + 1e-5 * big + 1e-30;
  for (int a = 0; a < 3; ++a) {
    box[a] = std::nextafter(float(lo[a] - pad), -HUGE_VALF);
    box[a + 3] = std::nextafter(float(hi[a] + pad), HUGE_VALF);
  }
}

// Reflectiveness of Phong (muon/materials.cc:113-123) and GGX (:270-293).
float Reflectiveness(const float *m) {
  const float s = ((0.0f + m[6]) + m[7] + m[8]) / 3.0f;
  const float d = ((0.0f + m[3]) + m[4] + m[5]) / 3.0f;
  const float denom = d + s;
  if (int(m[14]) == MUONX_BRDF_GGX) {
    if (denom == 0.0f) return 1.0f;
    return gmax(s / denom, 0.25f);
  }
  return denom > 0.0f ? s / denom : 0.0f;
}

template <typename T>
int Upload(DevBuf &buf, const std::vector<T> &host, uint64_t &total) {
  MX_HIP(buf.Alloc(host.size() * sizeof(T)));
  if (!host.empty()) MX_HIP(hipMemcpy(buf.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  total += buf.bytes;
  return MUONX_OK;
}

}  // namespace

namespace {
std::mutex g_scene_mu;
std::unordered_set<uint64_t> g_live_scenes;
uint64_t g_next_scene_id = 1;
}  // namespace

uint64_t RegisterScene() {
  std::lock_guard<std::mutex> lock(g_scene_mu);
  const uint64_t id = g_next_scene_id++;
  g_live_scenes.insert(id);
  return id;
}

void UnregisterScene(uint64_t id) {
  std::lock_guard<std::mutex> lock(g_scene_mu);
  g_live_scenes.erase(id);
}

bool SceneLive(uint64_t id) {
  std::lock_guard<std::mutex> lock(g_scene_mu);
  return g_live_scenes.count(id) != 0;
}

}  // namespace muonx

using namespace muonx;

extern "C" int muonx_scene_create(muonx_ctx *ctx, const muonx_scene_desc *din, muonx_scene **out) {
  if (!ctx || !din || !out) {
    SetError("muonx_scene_create: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = nullptr;
  const uint32_t np = din->num_prims;
  if (np == 0) {
    SetError("muonx_scene_create: scene has no primitives");  // the reference dereferences a null BVH root
    return MUONX_ERR_INVALID;
  }
  if (!din->prim_type || !din->prim_geom || !din->prim_transform || !din->prim_material || !din->transforms ||
      !din->materials) {
    SetError("muonx_scene_create: descriptor is missing arrays");
    return MUONX_ERR_INVALID;
  }
  for (uint32_t i = 0; i < np; ++i) {
    if (din->prim_transform[i] < 0 || uint32_t(din->prim_transform[i]) >= din->num_transforms) {
      SetError("muonx_scene_create: invalid transform index");
      return MUONX_ERR_INVALID;
    }
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  MX_HIP(hipSetDevice(ctx->device));
  // a descriptor without a tree: build one on the device
  DeviceBvh built;
  muonx_scene_desc dd = *din;
  const bool device_tree = din->num_nodes == 0;
  if (device_tree) {
    const int brc = BuildDeviceBvh(ctx, din, &built);
    if (brc != MUONX_OK) return brc;
    dd.num_nodes = uint32_t(built.left.size());
    dd.node_bounds = built.bounds.data();
    dd.node_left = built.left.data();
    dd.node_right = built.right.data();
    dd.node_axis = built.axis.data();
    dd.node_start = built.start.data();
    dd.node_count = built.count.data();
    dd.leaf_prims = built.leaf.data();
    dd.acceleration = 1;
  }
  const muonx_scene_desc *d = &dd;
  if (!d->node_bounds || !d->node_left || !d->node_right || !d->node_axis || !d->node_start || !d->node_count ||
      !d->leaf_prims) {
    SetError("muonx_scene_create: descriptor is missing arrays");
    return MUONX_ERR_INVALID;
  }

  // ---- primitives in leaf order (+ one sentinel that never hits)
  std::vector<float4> prims(size_t(np + 1) * 3, make_float4(0, 0, 0, 0));
  std::vector<int32_t> pid(np + 1, -1), pmat(np + 1, 0), plight(np + 1, -1);
  std::vector<float> pvn;
  std::vector<uint8_t> puvn;
  bool any_vn = false;
  if (d->prim_use_vnormal && d->prim_vnormal)
    for (uint32_t i = 0; i < np; ++i) any_vn |= d->prim_use_vnormal[i] != 0;
  if (any_vn) {
    pvn.assign(size_t(np + 1) * 9, 0.0f);
    puvn.assign(np + 1, 0);
  }
  for (uint32_t k = 0; k < np; ++k) {
    const int32_t id = d->leaf_prims[k];
    if (id < 0 || uint32_t(id) >= np) {
      SetError("muonx_scene_create: leaf_prims holds an invalid id");
      return MUONX_ERR_INVALID;
    }
    const float *g = d->prim_geom + size_t(id) * 12;
    const int32_t tid = d->prim_transform[id];
    if (tid < 0 || uint32_t(tid) >= d->num_transforms) {
      SetError("muonx_scene_create: invalid transform index");
      return MUONX_ERR_INVALID;
    }
    const int32_t mid = d->prim_material[id];
    if (mid < 0 || uint32_t(mid) >= d->num_materials) {
      SetError("muonx_scene_create: invalid material index");
      return MUONX_ERR_INVALID;
    }
    uint32_t tw = uint32_t(tid);
    float4 *q = &prims[size_t(k) * 3];
    if (d->prim_type[id] == MUONX_PRIM_SPHERE) {
      tw |= kSphereBit;
      q[0] = make_float4(g[0], g[1], g[2], BitsAsFloat(tw));
      q[1] = make_float4(g[3], 0, 0, 0);
    } else {
      q[0] = make_float4(g[0], g[1], g[2], BitsAsFloat(tw));
      q[1] = make_float4(g[3], g[4], g[5], 0);
      q[2] = make_float4(g[6], g[7], g[8], 0);
    }
    pid[k] = id;
    pmat[k] = mid;
    plight[k] = d->prim_light ? d->prim_light[id] : -1;
    if (any_vn) {
      puvn[k] = d->prim_use_vnormal[id];
      std::memcpy(&pvn[size_t(k) * 9], d->prim_vnormal + size_t(id) * 9, 9 * sizeof(float));
    }
  }
  // Sentinel: transform 0, all-zero triangle (zero normal: never hit), end of leaf.
  prims[size_t(np) * 3 + 1].w = BitsAsFloat(1u);

  // ---- BVH: child boxes in parent; internal binary nodes in pre-order.
  const uint32_t nn = d->num_nodes;
  std::vector<int32_t> dev_index(nn, -1);
  uint32_t n_internal = 0;
  uint32_t depth = 1;
  {
    // depth by a pre-order walk
    std::vector<uint32_t> dep(nn, 1);
    for (uint32_t i = 0; i < nn; ++i) {
      if (d->node_count[i] == 0) {
        const int32_t l = d->node_left[i], r = d->node_right[i];
        if (l <= int32_t(i) || r <= int32_t(i) || uint32_t(l) >= nn || uint32_t(r) >= nn) {
          SetError("muonx_scene_create: BVH is not in pre-order");
          return MUONX_ERR_INVALID;
        }
        dev_index[i] = int32_t(n_internal++);
        dep[l] = dep[r] = dep[i] + 1;
      } else {
        const int64_t s = d->node_start[i], c = d->node_count[i];
        if (s < 0 || c <= 0 || s + c > int64_t(np)) {
          SetError("muonx_scene_create: leaf range out of bounds");
          return MUONX_ERR_INVALID;
        }
        prims[size_t(s + c - 1) * 3 + 1].w = BitsAsFloat(1u);  // last primitive of the leaf
      }
      depth = std::max(depth, dep[i]);
    }
  }
  const bool root_leaf = d->node_count[0] != 0;
  const uint32_t n_dev = root_leaf ? 1 : n_internal;
  std::vector<float4> nodes(size_t(n_dev) * 4);
  auto child_ref = [&](int32_t c) -> uint32_t {
    if (d->node_count[c] != 0) return kLeafBit | uint32_t(d->node_start[c]);
    return uint32_t(dev_index[c]);
  };
  auto child_meta = [&](int32_t c) -> uint32_t { return d->node_count[c] != 0 ? kLeafBit : 0u; };
  auto fill = [&](uint32_t di, const float *b0, const float *b1, uint32_t r0, uint32_t r1, uint32_t m0, uint32_t m1,
                  uint32_t axis) {
    float4 *q = &nodes[size_t(di) * 4];
    q[0] = make_float4(b0[0], b0[3], b0[1], b0[4]);
    q[1] = make_float4(b1[0], b1[3], b1[1], b1[4]);
    q[2] = make_float4(b0[2], b0[5], b1[2], b1[5]);
    q[3] = make_float4(BitsAsFloat(r0), BitsAsFloat(r1), BitsAsFloat(m0 | (axis << 29)),
                       BitsAsFloat(m1));
  };
  if (root_leaf) {
    const float *b = d->node_bounds;
    fill(0, b, b, kLeafBit | uint32_t(d->node_start[0]), kLeafBit | np, kLeafBit, kLeafBit, 0);
    depth = 2;
  } else {
    for (uint32_t i = 0; i < nn; ++i) {
      if (d->node_count[i] != 0) continue;
      const int32_t l = d->node_left[i], r = d->node_right[i];
      fill(uint32_t(dev_index[i]), d->node_bounds + size_t(l) * 6, d->node_bounds + size_t(r) * 6, child_ref(l),
           child_ref(r), child_meta(l), child_meta(r), uint32_t(d->node_axis[i]) & 3u);
    }
  }

  // ---- wide BVH (fast traversal mode)
  // Leaf references carry counts (and small subtrees merge into one leaf)
  // when every count and first primitive fits the reference.
  uint32_t max_leaf = 0;
  for (uint32_t i = 0; i < nn; ++i) max_leaf = std::max<uint32_t>(max_leaf, uint32_t(std::max(0, d->node_count[i])));
  const bool leaf_counts = np < (1u << kLeafCountShift) && max_leaf <= kLeafMaxCount;
  uint32_t merge = 0;
  if (leaf_counts) {
    merge = std::min(kLeafMerge, kLeafMaxCount);
  }
  std::vector<F4> wnodes;
  uint32_t wstack = 1;
  if (root_leaf) {
    wnodes.assign(8, F4{0, 0, 0, 0});
    const float *b = d->node_bounds;
    wnodes[0] = F4{b[0], 0, 0, 0};
    wnodes[1] = F4{b[3], 0, 0, 0};
    wnodes[2] = F4{b[1], 0, 0, 0};
    wnodes[3] = F4{b[4], 0, 0, 0};
    wnodes[4] = F4{b[2], 0, 0, 0};
    wnodes[5] = F4{b[5], 0, 0, 0};
    const uint32_t ref = leaf_counts ? kLeafBit | ((uint32_t(d->node_count[0]) - 1u) << kLeafCountShift) |
                                           uint32_t(d->node_start[0])
                                     : kLeafBit | uint32_t(d->node_start[0]);
    wnodes[6] = F4{BitsAsFloat(ref), 0, 0, 0};
    wnodes[7] = F4{BitsAsFloat(1u), 0, 0, 0};
  } else {
    wnodes = CollapseWide(d, SubtreeSpans(d), merge, &wstack);
  }

  std::vector<float> mats(size_t(d->num_materials) * kMatStride);
  for (uint32_t i = 0; i < d->num_materials; ++i) {
    std::memcpy(&mats[size_t(i) * 16], d->materials + size_t(i) * 16, 16 * sizeof(float));
    mats[size_t(i) * 16 + 15] = Reflectiveness(&mats[size_t(i) * 16]);
  }
  std::vector<float> lights(size_t(std::max(1u, d->num_lights)) * kLightStride, 0.0f);
  uint32_t nquad = 0;
  for (uint32_t i = 0; i < d->num_lights; ++i) {
    std::memcpy(&lights[size_t(i) * 16], d->lights + size_t(i) * 16, 16 * sizeof(float));
    if (int(lights[size_t(i) * 16]) == MUONX_LIGHT_QUAD) {
      ++nquad;
      // slot 15: |e0 x e1| (muon/integration.cc:383-384)
      const float *L = &lights[size_t(i) * 16];
      V3 e0{L[7], L[8], L[9]}, e1{L[10], L[11], L[12]};
      lights[size_t(i) * 16 + 15] = length(cross(e0, e1));
    }
  }
  std::vector<float> xf(d->transforms, d->transforms + size_t(d->num_transforms) * 48);

  auto *s = new muonx_scene();
  s->ctx = ctx;
  uint64_t total = 0;
  int rc = MUONX_OK;
  // empty slots point at the sentinel primitive np (a one-primitive leaf in
  // either leaf encoding: count field 0, and the sentinel carries the end flag)
  const std::vector<U4> qnodes = QuantizeWide(wnodes, kLeafBit | np);
  std::vector<uint32_t> lprims;
  for (uint32_t k = 0; k < np; ++k)
    if (plight[k] >= 0) lprims.push_back(k);
  const bool cull_mis = lprims.size() <= kMaxLightPrims;
  if (!cull_mis || lprims.empty()) lprims.assign(1, 0u);  // keep the buffer non-empty
  if ((rc = Upload(s->nodes, nodes, total)) || (rc = Upload(s->wnodes, wnodes, total)) ||
      (rc = Upload(s->qnodes, qnodes, total)) ||
      (rc = Upload(s->prims, prims, total)) ||
      (rc = Upload(s->xforms, xf, total)) || (rc = Upload(s->prim_id, pid, total)) ||
      (rc = Upload(s->prim_mat, pmat, total)) || (rc = Upload(s->prim_light, plight, total)) ||
      (rc = Upload(s->materials, mats, total)) || (rc = Upload(s->lights, lights, total)) ||
      (rc = Upload(s->light_prims, lprims, total))) {
    delete s;
    return rc;
  }
  if (any_vn) {
    if ((rc = Upload(s->prim_vn, pvn, total)) || (rc = Upload(s->prim_use_vn, puvn, total))) {
      delete s;
      return rc;
    }
  }
  DevSceneView &v = s->view;
  v.nodes = s->nodes.as<float4>();
  v.wnodes = s->wnodes.as<float4>();
  v.qnodes = s->qnodes.as<uint4>();
  v.num_wnodes = uint32_t(wnodes.size() / 8);
  v.wstack = wstack;
  v.leaf_counts = leaf_counts ? 1u : 0u;
  v.root_leaf = root_leaf ? 1u : 0u;
  v.linear = (d->acceleration == 0 && root_leaf) ? 1u : 0u;
  v.light_prims = s->light_prims.as<uint32_t>();
  v.cull_mis = cull_mis ? 1u : 0u;
  {
    uint32_t ne = 0;
    for (uint32_t k = 0; k < np; ++k) ne += plight[k] >= 0 ? 1u : 0u;
    v.num_light_prims = cull_mis ? ne : 0u;
    EmitterBox(prims, xf, cull_mis ? lprims : std::vector<uint32_t>{}, v.num_light_prims, v.light_box);
  }
  v.prims = s->prims.as<float4>();
  v.xforms = s->xforms.as<float>();
  v.prim_id = s->prim_id.as<int32_t>();
  v.prim_mat = s->prim_mat.as<int32_t>();
  v.prim_light = s->prim_light.as<int32_t>();
  v.prim_vn = any_vn ? s->prim_vn.as<float>() : nullptr;
  v.prim_use_vn = any_vn ? s->prim_use_vn.as<uint8_t>() : nullptr;
  v.materials = s->materials.as<float>();
  v.lights = s->lights.as<float>();
  v.num_prims = np;
  v.num_nodes = n_dev;
  v.num_lights = d->num_lights;
  v.num_quad_lights = nquad;
  v.bvh_depth = depth;
  s->bvh_depth = depth;
  s->device_bytes = total;
  s->reference_tree = !device_tree;
  s->build_ms = device_tree ? built.total_ms : 0.0f;
  std::memcpy(s->bounds, d->node_bounds, sizeof(s->bounds));  // the root's box
  s->host_lights = std::move(lights);
  s->host_materials = std::move(mats);
  s->id = RegisterScene();
  *out = s;
  return MUONX_OK;
}

extern "C" void muonx_scene_destroy(muonx_scene *s) {
  if (!s) return;
  // server jobs still in flight keep using the scene's device memory
  while (s->inflight.load() > 0) std::this_thread::sleep_for(std::chrono::milliseconds(1));
  UnregisterScene(s->id);
  if (s->ctx) (void)hipSetDevice(s->ctx->device);
  delete s;
}

extern "C" int muonx_scene_get_info(const muonx_scene *s, muonx_scene_info *out) {
  if (!s || !out) {
    SetError("muonx_scene_get_info: null argument");
    return MUONX_ERR_INVALID;
  }
  out->device_bytes = s->device_bytes;
  out->num_prims = s->view.num_prims;
  out->num_nodes = s->view.num_nodes;
  out->bvh_depth = s->bvh_depth;
  out->num_lights = s->view.num_lights;
  out->num_quad_lights = s->view.num_quad_lights;
  out->device_built = s->reference_tree ? 0u : 1u;
  out->build_ms = s->build_ms;
  return MUONX_OK;
}
