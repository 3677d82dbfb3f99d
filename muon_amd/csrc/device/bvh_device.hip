// Device-side BVH build for scenes handed over without a tree (the host SAH
// build of muon/acceleration.cc:204-458 skipped for fast admission): a
// linear BVH. Primitive centroids get 30-bit Morton codes in the scene's
// centroid box, a radix sort orders them, the binary radix tree over the
// sorted codes is built one internal node per thread (T. Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees", HPG
// 2012), and node boxes are unions computed bottom-up, the second child to
// arrive at a node computing it. The tree comes back to the host in the
// descriptor's pre-order layout, where the scene upload collapses and
// quantizes it like a host-built tree.
#include "bvh_device.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <string>

#include "../common/vecmath.h"

namespace muonx {
namespace {

constexpr int kBlk = 256;
constexpr uint32_t kLeafRef = 0x80000000u;

// World box of primitive i, as the host builder computes it
// (muon/objects.cc:186-193: transformed triangle vertices; muon/bounds.cc:52-69:
// the 8 transformed corners of a sphere's object box).
__device__ __forceinline__ void world_box(const uint8_t *type, const float *geom, const int32_t *xf_id,
                                          const float *xforms, uint32_t i, V3 &lo, V3 &hi) {
  const float *g = geom + size_t(i) * 12;
  const float *m = xforms + size_t(xf_id[i]) * 48;
  auto add = [&](V3 p) {
    lo = V3{gmin(lo.x, p.x), gmin(lo.y, p.y), gmin(lo.z, p.z)};
    hi = V3{gmax(hi.x, p.x), gmax(hi.y, p.y), gmax(hi.z, p.z)};
  };
  if (type[i] == MUONX_PRIM_TRI) {
    const V3 a = xform_point(m, V3{g[0], g[1], g[2]});
    lo = hi = a;
    add(xform_point(m, V3{g[3], g[4], g[5]}));
    add(xform_point(m, V3{g[6], g[7], g[8]}));
    return;
  }
  const V3 c{g[0], g[1], g[2]};
  const float r = g[3];
  const V3 l = c - v3s(r), h = c + v3s(r);
  lo = hi = xform_point(m, l);
  add(xform_point(m, V3{h.x, l.y, l.z}));
  add(xform_point(m, V3{l.x, h.y, l.z}));
  add(xform_point(m, V3{l.x, l.y, h.z}));
  add(xform_point(m, V3{l.x, h.y, h.z}));
  add(xform_point(m, V3{h.x, l.y, h.z}));
  add(xform_point(m, V3{h.x, h.y, l.z}));
  add(xform_point(m, h));
}

// Floats as unsigned integers of the same order (for atomic min / max).
__device__ __forceinline__ uint32_t ordered(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_init_bounds(uint32_t *cb) {
  if (threadIdx.x < 3) cb[threadIdx.x] = 0xffffffffu;
  else if (threadIdx.x < 6) cb[threadIdx.x] = 0u;
}

// Primitive boxes (6 floats each, parse order) and the centroid box.
__global__ void __launch_bounds__(kBlk) k_prim_boxes(const uint8_t *type, const float *geom, const int32_t *xf_id,
                                                     const float *xforms, uint32_t n, float *box, uint32_t *cb) {
  __shared__ float smin[3][kBlk], smax[3][kBlk];
  V3 cmin{__builtin_inff(), __builtin_inff(), __builtin_inff()}, cmax{-__builtin_inff(), -__builtin_inff(),
                                                                     -__builtin_inff()};
  for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
    V3 lo, hi;
    world_box(type, geom, xf_id, xforms, i, lo, hi);
    float *b = box + size_t(i) * 6;
    b[0] = lo.x, b[1] = lo.y, b[2] = lo.z, b[3] = hi.x, b[4] = hi.y, b[5] = hi.z;
    const V3 c = (lo + hi) * 0.5f;
    cmin = V3{fminf(cmin.x, c.x), fminf(cmin.y, c.y), fminf(cmin.z, c.z)};
    cmax = V3{fmaxf(cmax.x, c.x), fmaxf(cmax.y, c.y), fmaxf(cmax.z, c.z)};
  }
  smin[0][threadIdx.x] = cmin.x, smin[1][threadIdx.x] = cmin.y, smin[2][threadIdx.x] = cmin.z;
  smax[0][threadIdx.x] = cmax.x, smax[1][threadIdx.x] = cmax.y, smax[2][threadIdx.x] = cmax.z;
  __syncthreads();
  for (int s = kBlk / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) {
      for (int a = 0; a < 3; ++a) {
        smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
        smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    atomicMin(&cb[threadIdx.x], ordered(smin[threadIdx.x][0]));
    atomicMax(&cb[3 + threadIdx.x], ordered(smax[threadIdx.x][0]));
  }
}

// 10 bits to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Morton code of each centroid in the centroid box: x in bits 3k+2, y in
// 3k+1, z in 3k.
__global__ void __launch_bounds__(kBlk) k_morton(const float *box, uint32_t n, const uint32_t *cb, uint32_t *code,
                                                 uint32_t *idx) {
  const float lo[3] = {unordered(cb[0]), unordered(cb[1]), unordered(cb[2])};
  const float hi[3] = {unordered(cb[3]), unordered(cb[4]), unordered(cb[5])};
  for (uint32_t i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
    const float *b = box + size_t(i) * 6;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const float c = (b[a] + b[a + 3]) * 0.5f;
      const float ext = hi[a] - lo[a];
      const float t = ext > 0.0f ? (c - lo[a]) / ext : 0.0f;
      q[a] = uint32_t(fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f));
    }
    code[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    idx[i] = i;
  }
}

// Length of the common prefix of sorted keys i and j (the code, then the
// index to break ties), -1 outside [0, n).
__device__ __forceinline__ int prefix(const uint32_t *code, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = code[i], b = code[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz(uint32_t(i) ^ uint32_t(j));
}

// Internal node i of the binary radix tree over n sorted keys: its children
// (a leaf flagged kLeafRef, else an internal node), its parent links and the
// axis of the Morton bit that splits it.
__global__ void __launch_bounds__(kBlk) k_radix_tree(const uint32_t *code, int n, uint32_t *child, int32_t *axis,
                                                     int32_t *iparent, int32_t *lparent) {
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n - 1; i += gridDim.x * kBlk) {
    const int d = prefix(code, n, i, i + 1) - prefix(code, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = prefix(code, n, i, i - d);
    int lmax = 2;
    while (prefix(code, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (prefix(code, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefix(code, n, i, j);
    int s = 0, t = l;
    do {
      t = (t + 1) / 2;
      if (prefix(code, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const uint32_t left = lo == g ? (uint32_t(g) | kLeafRef) : uint32_t(g);
    const uint32_t right = hi == g + 1 ? (uint32_t(g + 1) | kLeafRef) : uint32_t(g + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    if (left & kLeafRef) lparent[g] = i;
    else iparent[g] = i;
    if (right & kLeafRef) lparent[g + 1] = i;
    else iparent[g + 1] = i;
    const uint32_t x = code[lo] ^ code[hi];
    axis[i] = x ? 2 - (31 - int(__clz(x))) % 3 : 0;
  }
}

// Node boxes bottom-up: one thread per leaf walks towards the root; at each
// internal node the first child to arrive stops, the second computes the
// union (both children are then written).
__global__ void __launch_bounds__(kBlk) k_boxes_up(const float *pbox, const uint32_t *idx, int n,
                                                   const uint32_t *child, const int32_t *iparent,
                                                   const int32_t *lparent, uint32_t *arrivals, float *lbox,
                                                   float *ibox) {
  for (int k = blockIdx.x * kBlk + threadIdx.x; k < n; k += gridDim.x * kBlk) {
    const float *src = pbox + size_t(idx[k]) * 6;
    float *dst = lbox + size_t(k) * 6;
    for (int c = 0; c < 6; ++c) dst[c] = src[c];
    int cur = lparent[k];
    // the tree is at most 64 levels deep (prefixes of 30 code and 32 index
    // bits); the bound only guards the walk
    for (int level = 0; cur >= 0 && level < 96; ++level) {
      __threadfence();
      if (atomicAdd(&arrivals[cur], 1u) == 0u) break;
      __threadfence();
      float b[2][6];
      for (int s = 0; s < 2; ++s) {
        const uint32_t r = child[2 * cur + s];
        const volatile float *cb = (r & kLeafRef) ? lbox + size_t(r & ~kLeafRef) * 6 : ibox + size_t(r) * 6;
        for (int c = 0; c < 6; ++c) b[s][c] = cb[c];
      }
      float *o = ibox + size_t(cur) * 6;
      for (int c = 0; c < 3; ++c) {
        o[c] = gmin(b[0][c], b[1][c]);
        o[c + 3] = gmax(b[0][c + 3], b[1][c + 3]);
      }
      cur = iparent[cur];
    }
  }
}

inline int Grid(size_t n) { return int(std::max<size_t>(1, std::min<size_t>((n + kBlk - 1) / kBlk, 8192))); }

}  // namespace

int BuildDeviceBvh(muonx_ctx *ctx, const muonx_scene_desc *d, DeviceBvh *out) {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  const uint32_t n = d->num_prims;
  if (n == 0 || n >= (1u << 30)) {
    SetError("device BVH build: primitive count must be in [1, 2^30)");
    return MUONX_ERR_INVALID;
  }
  hipStream_t st = nullptr;
  MX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBuf type, geom, xfid, xf, pbox, cb, code, code2, idx, idx2, child, axis, ipar, lpar, arr, lbox, ibox, temp;
  const size_t ni = n > 1 ? n - 1 : 1;
  MX_HIP(type.Alloc(n));
  MX_HIP(geom.Alloc(size_t(n) * 12 * sizeof(float)));
  MX_HIP(xfid.Alloc(size_t(n) * sizeof(int32_t)));
  MX_HIP(xf.Alloc(size_t(d->num_transforms) * 48 * sizeof(float)));
  MX_HIP(pbox.Alloc(size_t(n) * 6 * sizeof(float)));
  MX_HIP(cb.Alloc(6 * sizeof(uint32_t)));
  MX_HIP(code.Alloc(size_t(n) * 4));
  MX_HIP(code2.Alloc(size_t(n) * 4));
  MX_HIP(idx.Alloc(size_t(n) * 4));
  MX_HIP(idx2.Alloc(size_t(n) * 4));
  MX_HIP(child.Alloc(ni * 2 * 4));
  MX_HIP(axis.Alloc(ni * 4));
  MX_HIP(ipar.Alloc(ni * 4));
  MX_HIP(lpar.Alloc(size_t(n) * 4));
  MX_HIP(arr.Alloc(ni * 4));
  MX_HIP(lbox.Alloc(size_t(n) * 6 * sizeof(float)));
  MX_HIP(ibox.Alloc(ni * 6 * sizeof(float)));
  MX_HIP(hipMemcpyAsync(type.ptr, d->prim_type, n, hipMemcpyHostToDevice, st));
  MX_HIP(hipMemcpyAsync(geom.ptr, d->prim_geom, size_t(n) * 12 * sizeof(float), hipMemcpyHostToDevice, st));
  MX_HIP(hipMemcpyAsync(xfid.ptr, d->prim_transform, size_t(n) * 4, hipMemcpyHostToDevice, st));
  MX_HIP(hipMemcpyAsync(xf.ptr, d->transforms, size_t(d->num_transforms) * 48 * sizeof(float), hipMemcpyHostToDevice,
                        st));
  hipEvent_t e0, e1;
  MX_HIP(hipEventCreate(&e0));
  MX_HIP(hipEventCreate(&e1));
  MX_HIP(hipEventRecord(e0, st));
  hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, st, cb.as<uint32_t>());
  hipLaunchKernelGGL(k_prim_boxes, dim3(Grid(n)), dim3(kBlk), 0, st, type.as<uint8_t>(), geom.as<float>(),
                     xfid.as<int32_t>(), xf.as<float>(), n, pbox.as<float>(), cb.as<uint32_t>());
  hipLaunchKernelGGL(k_morton, dim3(Grid(n)), dim3(kBlk), 0, st, pbox.as<float>(), n, cb.as<uint32_t>(),
                     code.as<uint32_t>(), idx.as<uint32_t>());
  size_t temp_bytes = 0;
  MX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, code.as<uint32_t>(), code2.as<uint32_t>(),
                                            idx.as<uint32_t>(), idx2.as<uint32_t>(), int(n), 0, 30, st));
  MX_HIP(temp.Alloc(temp_bytes));
  MX_HIP(hipcub::DeviceRadixSort::SortPairs(temp.ptr, temp_bytes, code.as<uint32_t>(), code2.as<uint32_t>(),
                                            idx.as<uint32_t>(), idx2.as<uint32_t>(), int(n), 0, 30, st));
  MX_HIP(hipMemsetAsync(lpar.ptr, 0xff, size_t(n) * 4, st));  // -1: the root leaf of a one-primitive scene
  MX_HIP(hipMemsetAsync(ipar.ptr, 0xff, ni * 4, st));
  MX_HIP(hipMemsetAsync(arr.ptr, 0, ni * 4, st));
  if (n > 1)
    hipLaunchKernelGGL(k_radix_tree, dim3(Grid(n - 1)), dim3(kBlk), 0, st, code2.as<uint32_t>(), int(n),
                       child.as<uint32_t>(), axis.as<int32_t>(), ipar.as<int32_t>(), lpar.as<int32_t>());
  hipLaunchKernelGGL(k_boxes_up, dim3(Grid(n)), dim3(kBlk), 0, st, pbox.as<float>(), idx2.as<uint32_t>(), int(n),
                     child.as<uint32_t>(), ipar.as<int32_t>(), lpar.as<int32_t>(), arr.as<uint32_t>(),
                     lbox.as<float>(), ibox.as<float>());
  MX_HIP(hipGetLastError());
  MX_HIP(hipEventRecord(e1, st));
  std::vector<uint32_t> h_child(n > 1 ? size_t(n - 1) * 2 : 0), h_idx(n);
  std::vector<int32_t> h_axis(n > 1 ? n - 1 : 0);
  std::vector<float> h_lbox(size_t(n) * 6), h_ibox(n > 1 ? size_t(n - 1) * 6 : 0);
  MX_HIP(hipMemcpyAsync(h_idx.data(), idx2.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, st));
  MX_HIP(hipMemcpyAsync(h_lbox.data(), lbox.ptr, h_lbox.size() * 4, hipMemcpyDeviceToHost, st));
  if (n > 1) {
    MX_HIP(hipMemcpyAsync(h_child.data(), child.ptr, h_child.size() * 4, hipMemcpyDeviceToHost, st));
    MX_HIP(hipMemcpyAsync(h_axis.data(), axis.ptr, h_axis.size() * 4, hipMemcpyDeviceToHost, st));
    MX_HIP(hipMemcpyAsync(h_ibox.data(), ibox.ptr, h_ibox.size() * 4, hipMemcpyDeviceToHost, st));
  }
  MX_HIP(hipStreamSynchronize(st));
  float dev_ms = 0.0f;
  MX_HIP(hipEventElapsedTime(&dev_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);

  // pre-order arrays of the descriptor: a node, its left subtree, its right
  // subtree; leaves appear in sorted order, so leaf k holds sorted primitive k
  DeviceBvh &b = *out;
  const size_t nn = 2 * size_t(n) - 1;
  b.bounds.assign(nn * 6, 0.0f);
  b.left.assign(nn, -1);
  b.right.assign(nn, -1);
  b.axis.assign(nn, 0);
  b.start.assign(nn, 0);
  b.count.assign(nn, 0);
  b.leaf.assign(h_idx.begin(), h_idx.end());
  b.depth = 1;
  struct Item {
    uint32_t ref;
    int32_t parent;
    bool right;
    uint32_t depth;
  };
  std::vector<Item> todo;
  todo.push_back({n > 1 ? 0u : kLeafRef, -1, false, 1});
  int32_t next = 0;
  while (!todo.empty() && size_t(next) < nn) {
    const Item it = todo.back();
    todo.pop_back();
    const int32_t at = next++;
    if (it.parent >= 0) (it.right ? b.right : b.left)[size_t(it.parent)] = at;
    b.depth = std::max(b.depth, it.depth);
    const bool leaf = (it.ref & kLeafRef) != 0;
    const uint32_t k = it.ref & ~kLeafRef;
    const float *src = leaf ? &h_lbox[size_t(k) * 6] : &h_ibox[size_t(k) * 6];
    std::copy(src, src + 6, &b.bounds[size_t(at) * 6]);
    if (leaf) {
      b.start[size_t(at)] = int32_t(k);
      b.count[size_t(at)] = 1;
      continue;
    }
    b.axis[size_t(at)] = h_axis[k];
    todo.push_back({h_child[2 * size_t(k) + 1], at, true, it.depth + 1});
    todo.push_back({h_child[2 * size_t(k)], at, false, it.depth + 1});
  }
  if (size_t(next) != nn || !todo.empty()) {
    SetError("device BVH build: the tree does not hold every primitive");
    return MUONX_ERR_STATE;
  }
  b.device_ms = dev_ms;
  b.total_ms = float(std::chrono::duration<double, std::milli>(Clock::now() - t0).count());
  return MUONX_OK;
}

}  // namespace muonx
