// ORACLE — test infrastructure only (see omath.h). extern "C" entry points
// used by tests/ (ctypes), __graft_entry__.smoke() and bench.py's CPU
// baseline. The product never loads this library.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle_core.h"

using namespace oracle;

namespace {

struct Handle {
  std::unique_ptr<Scene> scene;
};

void PutStats(const TraceStats &s, uint64_t *out) {
  if (!out) return;
  out[0] += s.primary_rays;
  out[1] += s.secondary_rays;
  out[2] += s.object_tests;
  out[3] += s.object_hits;
  out[4] += s.bounds_tests;
  out[5] += s.bounds_hits;
  out[6] += s.shadow_rays;
}

template <typename F>
void ParallelFor(int64_t n, int threads, F &&body) {
  if (threads < 1) threads = 1;
  std::atomic<int64_t> next(0);
  const int64_t chunk = 4096;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t]() {
      while (true) {
        int64_t s = next.fetch_add(chunk);
        if (s >= n) break;
        body(t, s, std::min(n, s + chunk));
      }
    });
  }
  for (auto &th : pool) th.join();
}

ParseOptions MakeOpts(int accel, int partition, long long seed) {
  ParseOptions o;
  o.accel = accel == 0 ? AccelType::kLinear : AccelType::kBVH;
  o.partition = partition == 0 ? Partition::kUniform : (partition == 1 ? Partition::kMidpoint : Partition::kSAH);
  o.default_seed = seed;
  return o;
}

void FlattenNodes(const BVHNode *n, std::vector<const BVHNode *> &out) {
  out.push_back(n);
  if (n->num_primitives == 0) {
    FlattenNodes(n->children[0].get(), out);
    FlattenNodes(n->children[1].get(), out);
  }
}

}  // namespace

extern "C" {

void *oracle_load(const char *path, int accel, int partition, long long default_seed) {
  auto h = new Handle();
  h->scene = ParseSceneFile(path, MakeOpts(accel, partition, default_seed));
  return h;
}

void *oracle_load_text(const char *text, int accel, int partition, long long default_seed) {
  auto h = new Handle();
  h->scene = ParseSceneText(text, MakeOpts(accel, partition, default_seed));
  return h;
}

// oracle_load_text with `brdf lambertian` building the Lambertian BRDF
// rather than the reference parser's Phong (a test hook for the Lambertian
// restatement, muon/materials.cc:17-47).
void *oracle_load_text_lambertian(const char *text, long long default_seed) {
  auto h = new Handle();
  ParseOptions o = MakeOpts(1, 2, default_seed);
  o.true_lambertian = true;
  h->scene = ParseSceneText(text, o);
  return h;
}

void oracle_free(void *p) { delete static_cast<Handle *>(p); }

// ints: W, H, spp, num_prims, integrator, nee, is, min_depth, max_depth, rr,
//       light_samples, light_stratify, num_lights, has_camera, warnings, cvn
// floats: gamma, attenuation xyz
int oracle_info(void *p, int32_t *ints, float *floats) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  ints[0] = s.width; ints[1] = s.height; ints[2] = s.pixel_samples;
  ints[3] = int32_t(s.root->primitives_.size());
  ints[4] = int32_t(s.integrator); ints[5] = int32_t(s.nee); ints[6] = int32_t(s.importance_sampling);
  ints[7] = s.min_depth; ints[8] = s.max_depth; ints[9] = s.russian_roulette;
  ints[10] = s.light_samples; ints[11] = s.light_stratify; ints[12] = int32_t(s.lights.size());
  ints[13] = s.camera ? 1 : 0; ints[14] = s.num_warnings; ints[15] = s.compute_vertex_normals;
  floats[0] = s.gamma;
  floats[1] = s.attenuation.x; floats[2] = s.attenuation.y; floats[3] = s.attenuation.z;
  return 0;
}

// eye, u, v, w (3 each), tan_fov_x, tan_fov_y
int oracle_camera(void *p, float *out) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  if (!s.camera) return -1;
  const Camera &c = *s.camera;
  V3 vs[4] = {c.eye, c.u, c.v, c.w};
  for (int i = 0; i < 4; ++i) { out[i * 3] = vs[i].x; out[i * 3 + 1] = vs[i].y; out[i * 3 + 2] = vs[i].z; }
  out[12] = c.tan_fov_x;
  out[13] = c.tan_fov_y;
  return 0;
}

uint32_t oracle_pathtracer_seed(void *p) { return static_cast<Handle *>(p)->scene->pathtracer_seed; }

// Renders the scene as muon's Renderer would with --parallelism=parallelism.
// accum: H*W*3 float sum over samples (row-major), rgb8: H*W*3 film output.
double oracle_render(void *p, int parallelism, float *accum, uint8_t *rgb8, uint64_t *stats) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  Film film(s.width, s.height, s.pixel_samples, s.gamma);
  RenderResult r = Render(s, parallelism, film);
  if (accum) std::memcpy(accum, film.accum.data(), film.accum.size() * sizeof(float));
  if (rgb8) film.ToRGB8(rgb8);
  PutStats(r.stats, stats);
  return r.render_seconds;
}

// Batched closest-hit (Structure::Intersect). prim = parse-order id or -1.
int oracle_intersect(void *p, int64_t n, const float *o, const float *d, int32_t *prim, float *t,
                     float *pos, float *nrm, uint64_t *stats, int threads) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  std::vector<TraceStats> st(std::max(threads, 1));
  ParallelFor(n, threads, [&](int tid, int64_t b, int64_t e) {
    std::vector<BVHNode *> stack;
    for (int64_t i = b; i < e; ++i) {
      Ray r(V3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
      Intersection hit;
      bool ok = s.root->Intersect(stack, st[tid], r, hit);
      prim[i] = ok ? hit.obj->id : -1;
      if (t) t[i] = ok ? hit.distance : 0.0f;
      if (pos) { pos[3 * i] = hit.pos.x; pos[3 * i + 1] = hit.pos.y; pos[3 * i + 2] = hit.pos.z; }
      if (nrm) { nrm[3 * i] = hit.normal.x; nrm[3 * i + 1] = hit.normal.y; nrm[3 * i + 2] = hit.normal.z; }
    }
  });
  for (auto &x : st) PutStats(x, stats);
  return 0;
}

// Batched any-hit (Structure::HasIntersection); tmax may be null (= +inf).
int oracle_occluded(void *p, int64_t n, const float *o, const float *d, const float *tmax, uint8_t *occ,
                    uint64_t *stats, int threads) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  std::vector<TraceStats> st(std::max(threads, 1));
  ParallelFor(n, threads, [&](int tid, int64_t b, int64_t e) {
    std::vector<BVHNode *> stack;
    for (int64_t i = b; i < e; ++i) {
      Ray r(V3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
      float tm = tmax ? tmax[i] : std::numeric_limits<float>::infinity();
      occ[i] = s.root->HasIntersection(stack, st[tid], r, tm) ? 1 : 0;
    }
  });
  for (auto &x : st) PutStats(x, stats);
  return 0;
}

// Per-ray intersection of one primitive (parse-order id) with the ray: the
// object-space test of Primitive::Intersect (muon/objects.cc:29-46). Returns
// 1 on a hit and writes distance.
int oracle_prim_intersect(void *p, int32_t id, const float *o, const float *d, float *dist) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  for (auto &pr : s.root->primitives_) {
    if (pr->id != id) continue;
    Intersection hit;
    if (!pr->Intersect(Ray(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2])), hit)) return 0;
    *dist = hit.distance;
    return 1;
  }
  return -1;
}

// Smallest normalised barycentric coordinate min(u,v,w)/|n|^2 of the ray's
// object-space hit on triangle `id` (muon/objects.cc:113-161, before any
// rejection), or 1 for a sphere. Values near 0 mean the ray passes within
// rounding of the triangle's edge. Returns -1 if the ray is parallel.
int oracle_prim_edge(void *p, int32_t id, const float *o, const float *d, float *out) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  for (auto &pr : s.root->primitives_) {
    if (pr->id != id) continue;
    if (pr->is_sphere) {
      *out = 1.0f;
      return 0;
    }
    auto *t = static_cast<Tri *>(pr.get());
    Ray r = Ray(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2])).Transform(*pr->inv_transform);
    float dn = dot(t->normal_, r.d);
    if (dn == 0.0f) return -1;
    const V3 &a = t->v0_.pos, &b = t->v1_.pos, &c = t->v2_.pos;
    float tt = (dot(a, t->normal_) - dot(r.o, t->normal_)) / dn;
    V3 q = r.At(tt);
    float v = dot(t->normal_, cross(b - a, q - a));
    float w = dot(t->normal_, cross(c - b, q - b));
    float u = dot(t->normal_, cross(a - c, q - c));
    float m = std::min(u, std::min(v, w));
    *out = m / t->normal_length2_;
    return 0;
  }
  return -1;
}

int64_t oracle_bvh_num_nodes(void *p) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  auto *bvh = dynamic_cast<BVH *>(s.root.get());
  if (!bvh || !bvh->root()) return 0;
  std::vector<const BVHNode *> nodes;
  FlattenNodes(bvh->root(), nodes);
  return int64_t(nodes.size());
}

// Depth-first (pre-order, left before right) dump of the binary BVH.
// bounds: 6 floats per node (min xyz, max xyz); count: prims in a leaf (0 for
// internal); start: leaf start in leaf order; axis; left/right: pre-order
// indices of children (-1 for leaves).
int oracle_bvh_dump(void *p, float *bounds, int64_t *count, int64_t *start, int32_t *axis, int64_t *left,
                    int64_t *right) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  auto *bvh = dynamic_cast<BVH *>(s.root.get());
  if (!bvh || !bvh->root()) return -1;
  std::vector<const BVHNode *> nodes;
  FlattenNodes(bvh->root(), nodes);
  std::vector<int64_t> index_of_dummy;
  for (size_t i = 0; i < nodes.size(); ++i) {
    const BVHNode *n = nodes[i];
    bounds[6 * i + 0] = n->bounds.mn.x; bounds[6 * i + 1] = n->bounds.mn.y; bounds[6 * i + 2] = n->bounds.mn.z;
    bounds[6 * i + 3] = n->bounds.mx.x; bounds[6 * i + 4] = n->bounds.mx.y; bounds[6 * i + 5] = n->bounds.mx.z;
    count[i] = int64_t(n->num_primitives);
    start[i] = n->num_primitives ? int64_t(n->start) : 0;
    axis[i] = n->num_primitives ? 0 : n->axis;
    left[i] = right[i] = -1;
  }
  // Children indices: in pre-order the left child follows its parent; the
  // right child follows the whole left subtree.
  std::vector<int64_t> size(nodes.size(), 1);
  for (int64_t i = int64_t(nodes.size()) - 1; i >= 0; --i) {
    if (nodes[i]->num_primitives == 0) {
      int64_t l = i + 1;
      int64_t r = l + size[l];
      size[i] = 1 + size[l] + size[r];
      left[i] = l;
      right[i] = r;
    }
  }
  return 0;
}

// Parse-order ids of the primitives in leaf (build) order.
int oracle_prim_order(void *p, int32_t *ids) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  for (size_t i = 0; i < s.root->primitives_.size(); ++i) ids[i] = s.root->primitives_[i]->id;
  return 0;
}

// Per-primitive data indexed by parse-order id.
//   type: 0 tri, 1 sphere; light: index of the quad light or -1
//   xform: 48 floats (M, M^-1, M^-T column-major)
//   geom: tri v0,v1,v2 (9) or sphere centre,radius (4), zero padded to 12
//   mat: ambient, diffuse, specular, emission (12), shininess, roughness, brdf, reflectiveness
int oracle_prim_dump(void *p, int32_t *type, float *xform, float *geom, float *mat, int32_t *light) {
  Scene &s = *static_cast<Handle *>(p)->scene;
  for (const auto &pr : s.root->primitives_) {
    int i = pr->id;
    type[i] = pr->is_sphere ? 1 : 0;
    const M4 *ms[3] = {pr->transform.get(), pr->inv_transform.get(), pr->inv_transpose_transform.get()};
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) xform[i * 48 + k * 16 + c * 4 + r] = (*ms[k])[c][r];
    float *g = geom + i * 12;
    std::fill(g, g + 12, 0.0f);
    if (pr->is_sphere) {
      auto *sp = static_cast<Sphere *>(pr.get());
      g[0] = sp->pos_.x; g[1] = sp->pos_.y; g[2] = sp->pos_.z; g[3] = sp->radius_;
    } else {
      auto *t = static_cast<Tri *>(pr.get());
      V3 vs[3] = {t->v0_.pos, t->v1_.pos, t->v2_.pos};
      for (int k = 0; k < 3; ++k) { g[3 * k] = vs[k].x; g[3 * k + 1] = vs[k].y; g[3 * k + 2] = vs[k].z; }
    }
    const Material &m = *pr->material;
    V3 cs[4] = {m.ambient, m.diffuse, m.specular, m.emission};
    float *mm = mat + i * 16;
    for (int k = 0; k < 4; ++k) { mm[3 * k] = cs[k].x; mm[3 * k + 1] = cs[k].y; mm[3 * k + 2] = cs[k].z; }
    mm[12] = m.shininess; mm[13] = m.roughness; mm[14] = float(int(m.brdf)); mm[15] = m.Reflectiveness();
    light[i] = -1;
    if (pr->light) {
      for (size_t l = 0; l < s.lights.size(); ++l)
        if (s.lights[l].get() == pr->light) light[i] = int32_t(l);
    }
  }
  return 0;
}

int oracle_num_tiles(int w, int h, int spp, int parallelism) { return NumTiles(w, h, spp, parallelism); }

// The first n outputs of std::mt19937 seeded with `seed`, for known-answer tests.
void oracle_mt19937(uint32_t seed, int64_t n, uint32_t *out) {
  std::mt19937 g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = g();
}

// Uniform floats as muon's UniformRandom produces them.
void oracle_uniform(uint32_t seed, int64_t n, float *out) {
  UniformRandom r(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = r.Next();
}

}  // extern "C"
