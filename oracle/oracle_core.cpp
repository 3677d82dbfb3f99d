// ORACLE — test infrastructure only (see omath.h). Restatement of the
// reference render path; each function cites the reference file:line it
// follows.
#include "oracle_core.h"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>

namespace oracle {

// ---------------------------------------------------------------- Bounds
// muon/bounds.cc:21-31
Bounds::Bounds(V3 a, V3 b) {
  mn = V3(fmin_glm(a.x, b.x), fmin_glm(a.y, b.y), fmin_glm(a.z, b.z));
  mx = V3(fmax_glm(a.x, b.x), fmax_glm(a.y, b.y), fmax_glm(a.z, b.z));
}

// muon/bounds.cc:35-45
int Bounds::MaxAxis() const {
  V3 d = Dimensions();
  if (d.x > d.y && d.x > d.z) return 0;
  if (d.y > d.z) return 1;
  return 2;
}

// muon/bounds.cc:47-50
float Bounds::SurfaceArea() const {
  V3 d = Dimensions();
  return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}

// muon/bounds.cc:52-69
Bounds Bounds::Transform(const M4 &m) const {
  Bounds b(transform_position(m, mn));
  b = Union(b, transform_position(m, V3(mx.x, mn.y, mn.z)));
  b = Union(b, transform_position(m, V3(mn.x, mx.y, mn.z)));
  b = Union(b, transform_position(m, V3(mn.x, mn.y, mx.z)));
  b = Union(b, transform_position(m, V3(mn.x, mx.y, mx.z)));
  b = Union(b, transform_position(m, V3(mx.x, mn.y, mx.z)));
  b = Union(b, transform_position(m, V3(mx.x, mx.y, mn.z)));
  return Union(b, transform_position(m, mx));
}

// muon/bounds.cc:75-86
bool Bounds::HasIntersection(const Ray &ray, float max_distance) const {
  float t_min, t_max;
  if (!Intersect(ray, t_min, t_max)) return false;
  return t_max > 0.0f && t_min < max_distance;
}

// muon/bounds.cc:88-125 (axes with a zero direction component are skipped)
bool Bounds::Intersect(const Ray &ray, float &t_min, float &t_max) const {
  t_min = 0;
  t_max = std::numeric_limits<float>::infinity();
  for (int i = 0; i < 3; ++i) {
    if (ray.d[i] == 0.0f) continue;
    float a = (mn[i] - ray.o[i]) / ray.d[i];
    float b = (mx[i] - ray.o[i]) / ray.d[i];
    if (a > b) std::swap(a, b);
    t_min = fmax_glm(t_min, a);
    t_max = fmin_glm(t_max, b);
    if (t_min > t_max) return false;
  }
  return true;
}

// muon/bounds.cc:127-151
Bounds Bounds::Union(const Bounds &a, const Bounds &b) {
  Bounds r;
  r.mn = V3(fmin_glm(a.mn.x, b.mn.x), fmin_glm(a.mn.y, b.mn.y), fmin_glm(a.mn.z, b.mn.z));
  r.mx = V3(fmax_glm(a.mx.x, b.mx.x), fmax_glm(a.mx.y, b.mx.y), fmax_glm(a.mx.z, b.mx.z));
  return r;
}
Bounds Bounds::Union(const Bounds &a, V3 p) {
  Bounds r;
  r.mn = V3(fmin_glm(a.mn.x, p.x), fmin_glm(a.mn.y, p.y), fmin_glm(a.mn.z, p.z));
  r.mx = V3(fmax_glm(a.mx.x, p.x), fmax_glm(a.mx.y, p.y), fmax_glm(a.mx.z, p.z));
  return r;
}

// ---------------------------------------------------------------- sampling helpers
// muon/transform.cc:5-19
static V3 RotateToOrthonormalFrame(V3 s, V3 w) {
  V3 a = std::fabs(w.y) <= 0.9f ? V3(0.0f, 1.0f, 0.0f) : V3(1.0f, 0.0f, 0.0f);
  V3 u = normalize(cross(a, w));
  V3 v = normalize(cross(w, u));
  return s.x * u + s.y * v + s.z * w;
}

static V3 SphericalDir(float theta, float phi) {
  return V3(std::cos(phi) * std::sin(theta), std::sin(phi) * std::sin(theta), std::cos(theta));
}

// muon/hemisphere_sampling.cc:8-22
static V3 SampleHemisphere(V3 n, UniformRandom &rand) {
  float r1 = rand.Next();
  float r2 = rand.Next();
  float theta = std::acos(r1);
  float phi = 2.0f * kPi * r2;
  return RotateToOrthonormalFrame(SphericalDir(theta, phi), n);
}

// muon/hemisphere_sampling.cc:24-38
static V3 SampleCosine(V3 n, UniformRandom &rand) {
  float r1 = rand.Next();
  float r2 = rand.Next();
  float theta = std::acos(std::sqrt(r1));
  float phi = 2.0f * kPi * r2;
  return RotateToOrthonormalFrame(SphericalDir(theta, phi), n);
}

// ---------------------------------------------------------------- materials
// Phong: muon/materials.cc:113-123; GGX: muon/materials.cc:270-293.
float Material::Reflectiveness() const {
  float s = comp_add(specular) / 3.0f;
  float d = comp_add(diffuse) / 3.0f;
  float denom = d + s;
  if (brdf == BrdfType::kGGX) {
    if (denom == 0.0f) return 1.0f;
    return fmax_glm(s / denom, 0.25f);
  }
  return denom > 0.0f ? s / denom : 0.0f;
}

// muon/materials.cc:131-141
static float GGXD(V3 h, V3 n, float roughness) {
  float half_angle = std::acos(clamp_glm(dot(h, n), 0.0f, 1.0f));
  float a2 = roughness * roughness;
  float denom = kPi * std::pow(std::cos(half_angle), 4.0f) *
                std::pow(a2 + std::pow(std::tan(half_angle), 2.0f), 2.0f);
  return denom > 0.0f ? a2 / denom : 0.0f;
}

// muon/materials.cc:148-158
static float SmithG1(V3 dir, V3 n, float roughness) {
  float dn = dot(dir, n);
  if (dn <= 0.0f) return 0.0f;
  float angle = std::acos(clamp_glm(dn, 0.0f, 1.0f));
  float a2 = roughness * roughness;
  return 2.0f / (1.0f + std::sqrt(1.0f + a2 * std::pow(std::tan(angle), 2.0f)));
}

// Sample: Lambertian muon/materials.cc:17-24, Phong :49-71, GGX :180-205.
V3 Material::Sample(V3 ray_dir, V3 normal, UniformRandom &rand) const {
  if (brdf == BrdfType::kLambertian) return SampleCosine(normal, rand);
  float r0 = rand.Next();
  if (r0 > Reflectiveness()) return SampleCosine(normal, rand);
  float r1 = rand.Next();
  float r2 = rand.Next();
  if (brdf == BrdfType::kPhong) {
    float theta = std::acos(std::pow(r1, 1.0f / (shininess + 1.0f)));
    float phi = 2.0f * kPi * r2;
    return RotateToOrthonormalFrame(SphericalDir(theta, phi), reflect(ray_dir, normal));
  }
  float theta = std::atan(roughness * std::sqrt(r1) / std::sqrt(1.0f - r1));
  float phi = 2.0f * kPi * r2;
  V3 h = RotateToOrthonormalFrame(SphericalDir(theta, phi), normal);
  return reflect(ray_dir, h);
}

// PDF: Lambertian muon/materials.cc:26-34, Phong :73-91, GGX :207-225.
float Material::PDF(V3 in_dir, V3 ray_dir, V3 normal) const {
  if (brdf == BrdfType::kLambertian) return fmax_glm(dot(normal, in_dir), 0.0f) * kOneOverPi;
  float t = Reflectiveness();
  if (brdf == BrdfType::kPhong) {
    return (1.0f - t) * fmax_glm(dot(normal, in_dir), 0.0f) * kOneOverPi +
           t * (shininess + 1) * 0.5f * kOneOverPi *
               std::pow(fmax_glm(dot(reflect(ray_dir, normal), in_dir), 0.0f), shininess);
  }
  V3 h = normalize(in_dir - ray_dir);
  return (1.0f - t) * fmax_glm(dot(normal, in_dir), 0.0f) * kOneOverPi +
         t * GGXD(h, normal, roughness) * fmax_glm(dot(normal, h), 0.0f) / (4.0f * dot(h, in_dir));
}

// Eval: Lambertian muon/materials.cc:36-43, Phong :93-111, GGX :227-266.
V3 Material::Eval(V3 in_dir, V3 ray_dir, V3 normal) const {
  V3 kd = diffuse * kOneOverPi;
  if (brdf == BrdfType::kLambertian) return kd;
  if (brdf == BrdfType::kPhong) {
    V3 spec = specular * (shininess + 2.0f) * (0.5f * kOneOverPi) *
              std::pow(fmax_glm(dot(reflect(ray_dir, normal), in_dir), 0.0f), shininess);
    return kd + spec;
  }
  V3 h = normalize(in_dir - ray_dir);
  V3 fresnel = specular + (1.0f - specular) * std::pow(1.0f - dot(in_dir, h), 5.0f);
  float g = SmithG1(in_dir, normal, roughness) * SmithG1(-ray_dir, normal, roughness);
  float d = GGXD(h, normal, roughness);
  V3 spec = (fresnel * g * d) / (4.0f * dot(in_dir, normal) * dot(-ray_dir, normal));
  return kd + spec;
}

// ---------------------------------------------------------------- lights
// muon/lighting.cc:7-36
bool Quad::Intersect(const Ray &ray, Intersection &out) const {
  V3 normal = normalize(cross(edge1, edge0));
  float rdn = dot(ray.d, normal);
  if (rdn == 0.0f) return false;
  float t = dot(corner - ray.o, normal) / rdn;
  if (t < 0) return false;
  V3 p = ray.At(t);
  V3 pe = p - corner;
  float l0 = length(edge0), l1 = length(edge1);
  float u = dot(pe, edge0) / l0;
  float v = dot(pe, edge1) / l1;
  if (u > 0 && u < l0 && v > 0 && v < l1) {
    out.distance = t;
    out.pos = p;
    out.normal = normal;
    out.obj = nullptr;
    return true;
  }
  return false;
}

// muon/lighting.cc:38-72
ShadingInfo Light::ShadingInfoAt(V3 p) const {
  ShadingInfo info;
  switch (kind) {
    case LightKind::kDirectional:
      info.color = color;
      info.direction = direction;
      info.distance = std::numeric_limits<float>::infinity();
      break;
    case LightKind::kPoint: {
      V3 to_light = pos - p;
      V3 dir = normalize(to_light);
      float r = length(to_light);
      float att = attenuation.x + attenuation.y * r + attenuation.z * r * r;
      info.color = color / att;
      info.direction = dir;
      info.distance = r;
      break;
    }
    case LightKind::kQuad:
      info.color = color;
      info.area = &area;
      break;
  }
  return info;
}

// ---------------------------------------------------------------- primitives
// muon/objects.cc:29-46
bool Primitive::Intersect(const Ray &ray, Intersection &out) {
  Ray t_ray = ray.Transform(*inv_transform);
  if (!IntersectObjectSpace(t_ray, out)) return false;
  out.pos = transform_position(*transform, out.pos);
  out.normal = transform_direction(*inv_transpose_transform, out.normal);
  out.distance = length(out.pos - ray.o);
  return true;
}

// muon/objects.cc:9-17
bool Primitive::HasIntersection(const Ray &ray, float max_distance) {
  Intersection hit;
  if (!Intersect(ray, hit)) return false;
  return hit.distance > 0.0f && hit.distance < max_distance;
}

// muon/objects.cc:48-56
Tri::Tri(Vertex &v0, Vertex &v1, Vertex &v2, bool use_vertex_normals)
    : v0_(v0), v1_(v1), v2_(v2), use_vertex_normals_(use_vertex_normals) {
  V3 e_ba = v1_.pos - v0_.pos;
  V3 e_ca = v2_.pos - v0_.pos;
  normal_ = cross(e_ba, e_ca);
  normal_length2_ = dot(normal_, normal_);
}

// muon/objects.cc:58-175
bool Tri::IntersectObjectSpace(const Ray &ray, Intersection &out) {
  float dn = dot(normal_, ray.d);
  if (dn > -kEpsilon && dn < kEpsilon) return false;
  const V3 &a = v0_.pos, &b = v1_.pos, &c = v2_.pos;
  float t = (dot(a, normal_) - dot(ray.o, normal_)) / dn;
  if (t < 0) return false;
  V3 p = ray.At(t);
  float v = dot(normal_, cross(b - a, p - a));
  if (v < 0) return false;
  float w = dot(normal_, cross(c - b, p - b));
  if (w < 0) return false;
  float u = dot(normal_, cross(a - c, p - c));
  if (u < 0) return false;
  u /= normal_length2_;
  v /= normal_length2_;
  w /= normal_length2_;
  V3 n = use_vertex_normals_ ? w * v0_.normal + u * v1_.normal + v * v2_.normal : normalize(normal_);
  out.distance = t;
  out.pos = p;
  out.normal = n;
  out.obj = this;
  return true;
}

// muon/objects.cc:177-193
Bounds Tri::ObjectBounds() const {
  Bounds b(v0_.pos, v1_.pos);
  return Bounds::Union(b, v2_.pos);
}
Bounds Tri::WorldBounds() const {
  V3 a = transform_position(*transform, v0_.pos);
  V3 b = transform_position(*transform, v1_.pos);
  V3 c = transform_position(*transform, v2_.pos);
  Bounds bb(a, b);
  return Bounds::Union(bb, c);
}

// muon/objects.cc:195-270
bool Sphere::IntersectObjectSpace(const Ray &ray, Intersection &out) {
  V3 oc = ray.o - pos_;
  float bp = dot(ray.d, oc);
  float c = dot(oc, oc) - radius_ * radius_;
  float disc = bp * bp - c;
  if (disc < 0) return false;
  float sd = std::sqrt(disc);
  float r0 = -bp + sd;
  float r1 = -bp - sd;
  if (r0 < 0 && r1 < 0) return false;
  float t = fmin_glm(r0, r1);
  if (t < 0) t = fmax_glm(r0, r1);
  V3 p = ray.At(t);
  out.distance = t;
  out.pos = p;
  out.normal = normalize(p - pos_);
  out.obj = this;
  return true;
}

// ---------------------------------------------------------------- acceleration
// muon/acceleration.cc:14-45
bool Linear::Intersect(std::vector<BVHNode *> &, TraceStats &st, const Ray &ray,
                       Intersection &out) const {
  float min_dist = std::numeric_limits<float>::infinity();
  bool found = false;
  for (const auto &obj : primitives_) {
    ++st.object_tests;
    Intersection hit;
    if (!obj->Intersect(ray, hit)) continue;
    ++st.object_hits;
    if (hit.distance > 0.0f && hit.distance < min_dist) {
      min_dist = hit.distance;
      out = hit;
      found = true;
    }
  }
  return found;
}

bool Linear::HasIntersection(std::vector<BVHNode *> &, TraceStats &, const Ray &ray,
                             float max_distance) const {
  for (const auto &obj : primitives_)
    if (obj->HasIntersection(ray, max_distance)) return true;
  return false;
}

// muon/acceleration.cc:47-53
PrimitiveInfo::PrimitiveInfo(size_t idx, const Bounds &b) : original_index(idx), bounds(b) {
  V3 diag = b.mx - b.mn;
  centroid = b.mn + 0.5f * diag;
}

// muon/acceleration.cc:65-142
bool BVH::Intersect(std::vector<BVHNode *> &frontier, TraceStats &st, const Ray &ray,
                    Intersection &out) const {
  if (!root_) return false;  // the reference dereferences a null root here
  size_t first[3] = {ray.d.x < 0, ray.d.y < 0, ray.d.z < 0};
  float min_dist = std::numeric_limits<float>::infinity();
  bool found = false;
  BVHNode *node = root_.get();
  while (true) {
    ++st.bounds_tests;
    if (!node->bounds.HasIntersection(ray, min_dist)) {
      if (frontier.empty()) break;
      node = frontier.back();
      frontier.pop_back();
      continue;
    }
    ++st.bounds_hits;
    if (node->num_primitives > 0) {
      for (size_t i = node->start; i < node->start + node->num_primitives; ++i) {
        ++st.object_tests;
        Intersection hit;
        if (!primitives_[i]->Intersect(ray, hit)) continue;
        ++st.object_hits;
        if (hit.distance > 0.0f && hit.distance < min_dist) {
          min_dist = hit.distance;
          out = hit;
          found = true;
        }
      }
      if (frontier.empty()) break;
      node = frontier.back();
      frontier.pop_back();
      continue;
    }
    size_t c = first[node->axis];
    frontier.push_back(node->children[1 - c].get());
    node = node->children[c].get();
  }
  return found;
}

// muon/acceleration.cc:144-202
bool BVH::HasIntersection(std::vector<BVHNode *> &frontier, TraceStats &st, const Ray &ray,
                          float max_distance) const {
  if (!root_) return false;
  size_t first[3] = {ray.d.x < 0, ray.d.y < 0, ray.d.z < 0};
  BVHNode *node = root_.get();
  while (true) {
    ++st.bounds_tests;
    if (!node->bounds.HasIntersection(ray, max_distance)) {
      if (frontier.empty()) break;
      node = frontier.back();
      frontier.pop_back();
      continue;
    }
    ++st.bounds_hits;
    if (node->num_primitives > 0) {
      for (size_t i = node->start; i < node->start + node->num_primitives; ++i) {
        ++st.object_tests;
        if (primitives_[i]->HasIntersection(ray, max_distance)) {
          ++st.object_hits;
          frontier.clear();
          return true;
        }
      }
      if (frontier.empty()) break;
      node = frontier.back();
      frontier.pop_back();
      continue;
    }
    size_t c = first[node->axis];
    frontier.push_back(node->children[1 - c].get());
    node = node->children[c].get();
  }
  return false;
}

// muon/acceleration.cc:204-229
void BVH::Init() {
  if (primitives_.empty()) return;
  std::vector<PrimitiveInfo> info;
  info.reserve(primitives_.size());
  for (size_t i = 0; i < primitives_.size(); ++i) info.emplace_back(i, primitives_[i]->WorldBounds());
  root_ = Build(0, primitives_.size(), info);
  std::vector<std::unique_ptr<Primitive>> sorted;
  sorted.reserve(primitives_.size());
  for (const auto &pi : info) sorted.push_back(std::move(primitives_[pi.original_index]));
  primitives_.swap(sorted);
}

namespace {
constexpr size_t kNumBuckets = 12;  // muon/acceleration.cc:239
struct Bucket {
  size_t size = 0;
  Bounds bounds;
};
std::unique_ptr<BVHNode> MakeLeaf(size_t n, size_t start, const Bounds &b) {
  auto node = std::make_unique<BVHNode>();
  node->num_primitives = n;
  node->start = start;
  node->bounds = b;
  return node;
}
}  // namespace

// muon/acceleration.cc:241-458
std::unique_ptr<BVHNode> BVH::Build(size_t start, size_t end, std::vector<PrimitiveInfo> &info) const {
  size_t n = end - start;
  if (n == 1) return MakeLeaf(n, start, info[start].bounds);

  Bounds cb;
  for (size_t i = start; i < end; ++i) cb = Bounds::Union(cb, info[i].centroid);
  int axis = cb.MaxAxis();

  size_t split = size_t(-1);
  auto sit = info.begin() + start;
  auto eit = info.begin() + end;
  bool uniform = false;

  switch (partition_) {
    case Partition::kUniform:
      uniform = true;
      break;
    case Partition::kMidpoint: {
      float mid = (cb.mn[axis] + cb.mx[axis]) / 2;
      auto it = std::partition(sit, eit, [axis, mid](const PrimitiveInfo &p) { return p.centroid[axis] < mid; });
      split = size_t(it - info.begin());
      if (split == start || split == end) uniform = true;
      break;
    }
    case Partition::kSAH: {
      Bucket buckets[kNumBuckets];
      V3 space = cb.Dimensions();
      auto bucket_of = [&](const PrimitiveInfo &p) -> size_t {
        float off = space[axis] == 0.0f ? 0.0f : (p.centroid[axis] - cb.mn[axis]) / space[axis];
        size_t b = static_cast<size_t>(off * kNumBuckets);
        if (b == kNumBuckets) b = kNumBuckets - 1;
        return b;
      };
      for (size_t i = start; i < end; ++i) {
        Bucket &bk = buckets[bucket_of(info[i])];
        bk.size++;
        bk.bounds = Bounds::Union(bk.bounds, info[i].bounds);
      }
      Bounds all;
      for (size_t i = 0; i < kNumBuckets; ++i) all = Bounds::Union(all, buckets[i].bounds);

      constexpr int kSplits = kNumBuckets - 1;
      // The reference leaves entries whose left side is empty uninitialised;
      // bucket 0 always holds the minimum centroid, so that never happens.
      float costs[kSplits] = {};
      {
        size_t cnt = 0;
        Bounds acc;
        for (size_t i = 0; i < kSplits; ++i) {
          cnt += buckets[i].size;
          if (cnt == 0) continue;
          acc = Bounds::Union(acc, buckets[i].bounds);
          costs[i] = cnt * acc.SurfaceArea();
        }
      }
      {
        size_t cnt = 0;
        Bounds acc;
        for (int i = kSplits - 1; i >= 0; --i) {
          cnt += buckets[i + 1].size;
          if (cnt == 0) continue;
          acc = Bounds::Union(acc, buckets[i + 1].bounds);
          costs[i] += cnt * acc.SurfaceArea();
        }
      }
      float total = all.SurfaceArea();
      float min_cost = std::numeric_limits<float>::infinity();
      size_t best = 0;  // uninitialised in the reference when no cost is finite
      for (size_t i = 0; i < size_t(kSplits); ++i) {
        float cost = 0.125f + costs[i] / total;
        if (cost < min_cost) {
          min_cost = cost;
          best = i;
        }
      }
      float leaf_cost = n;
      if (leaf_cost < min_cost) return MakeLeaf(n, start, all);
      auto it = std::partition(sit, eit, [&](const PrimitiveInfo &p) { return bucket_of(p) <= best; });
      split = size_t(it - info.begin());
      if (split == start || split == end) uniform = true;
      break;
    }
  }
  if (uniform) {
    split = (start + end) / 2;
    std::nth_element(sit, info.begin() + split, eit,
                     [axis](const PrimitiveInfo &a, const PrimitiveInfo &b) { return a.centroid[axis] < b.centroid[axis]; });
  }
  auto node = std::make_unique<BVHNode>();
  node->children[0] = Build(start, split, info);
  node->children[1] = Build(split, end, info);
  node->axis = axis;
  node->num_primitives = 0;
  node->bounds = Bounds::Union(node->children[0]->bounds, node->children[1]->bounds);
  return node;
}

// ---------------------------------------------------------------- camera
// muon/camera.cc:15-35
Camera::Camera(V3 e, V3 l, V3 up_, float f, int w_, int h_)
    : eye(e), look_at(l), up(up_), fov(f), width(w_), height(h_) {
  tan_fov_y = std::tan(radians(fov) / 2.0f);
  tan_fov_x = tan_fov_y * (static_cast<float>(width) / height);
  V3 a = eye - look_at;
  w = normalize(a);
  u = normalize(cross(up, w));
  v = cross(w, u);
}

// muon/camera.cc:37-62
Ray Camera::CastRay(float x, float y) const {
  float hw = width / 2.0f, hh = height / 2.0f;
  float ox = (x - hw) / hw;
  float oy = (hh - y) / hh;
  float alpha = tan_fov_x * ox;
  float beta = tan_fov_y * oy;
  V3 dir = normalize(alpha * u + beta * v - w);
  return Ray(eye, dir);
}

// ---------------------------------------------------------------- parser
namespace {

void Trim(std::string &s) {
  s.erase(s.begin(), std::find_if(s.begin(), s.end(), [](unsigned char c) { return !std::isspace(c); }));
  s.erase(std::find_if(s.rbegin(), s.rend(), [](unsigned char c) { return !std::isspace(c); }).base(), s.end());
}

// muon/parser.h:18-57 and muon/parser.cc:121-159
struct Workspace {
  std::unique_ptr<std::mt19937> seedgen;
  std::shared_ptr<Material> material;
  std::unique_ptr<Scene> scene;
  std::vector<std::shared_ptr<M4>> transforms{std::make_shared<M4>(1.0f)};
  std::shared_ptr<M4> inv = std::make_shared<M4>(1.0f);
  std::shared_ptr<M4> inv_t = std::make_shared<M4>(1.0f);

  void UpdateCached() {
    inv = std::make_shared<M4>(inverse(*transforms.back()));
    inv_t = std::make_shared<M4>(transpose(*inv));
  }
  void Multiply(const M4 &m) {
    transforms.back() = std::make_shared<M4>(mul(*transforms.back(), m));
    UpdateCached();
  }
  void Push() {
    transforms.push_back(transforms.back());
    UpdateCached();
  }
  void Pop() {
    if (transforms.size() <= 1) {  // undefined in the reference; ignore
      ++scene->num_warnings;
      return;
    }
    transforms.pop_back();
    UpdateCached();
  }
  void GenMaterial() { material = std::make_shared<Material>(*material); }
  void Update(Primitive &p) {
    p.material = material;
    p.transform = transforms.back();
    p.inv_transform = inv;
    p.inv_transpose_transform = inv_t;
  }
};

enum class Cmd {
  kRandomSeed, kFilmSize, kMinDepth, kMaxDepth, kOutput, kGamma, kIntegrator, kPixelSamples,
  kLightSamples, kLightStratify, kNEE, kRR, kIS, kCamera, kLoad, kCVN, kSphere, kStartMesh,
  kEndMesh, kVertex, kVertexNormal, kTri, kTriNormal, kTranslate, kRotate, kScale, kPush, kPop,
  kDirectional, kPoint, kAttenuation, kQuadLight, kBRDF, kAmbient, kDiffuse, kSpecular,
  kShininess, kRoughness, kEmission
};

// muon/parser.cc:79-119
const std::map<std::string, Cmd> &CommandMap() {
  static const std::map<std::string, Cmd> m = {
      {"random_seed", Cmd::kRandomSeed}, {"film_size", Cmd::kFilmSize}, {"min_depth", Cmd::kMinDepth},
      {"max_depth", Cmd::kMaxDepth}, {"output", Cmd::kOutput}, {"gamma", Cmd::kGamma},
      {"integrator", Cmd::kIntegrator}, {"pixel_samples", Cmd::kPixelSamples},
      {"light_samples", Cmd::kLightSamples}, {"light_stratify", Cmd::kLightStratify},
      {"next_event_estimation", Cmd::kNEE}, {"russian_roulette", Cmd::kRR},
      {"importance_sampling", Cmd::kIS}, {"camera", Cmd::kCamera}, {"load", Cmd::kLoad},
      {"compute_vertex_normals", Cmd::kCVN}, {"sphere", Cmd::kSphere}, {"start_mesh", Cmd::kStartMesh},
      {"end_mesh", Cmd::kEndMesh}, {"vertex", Cmd::kVertex}, {"vertex_normal", Cmd::kVertexNormal},
      {"tri", Cmd::kTri}, {"tri_normal", Cmd::kTriNormal}, {"translate", Cmd::kTranslate},
      {"rotate", Cmd::kRotate}, {"scale", Cmd::kScale}, {"push_transform", Cmd::kPush},
      {"pop_transform", Cmd::kPop}, {"directional_light", Cmd::kDirectional},
      {"point_light", Cmd::kPoint}, {"attenuation", Cmd::kAttenuation}, {"quad_light", Cmd::kQuadLight},
      {"brdf", Cmd::kBRDF}, {"ambient", Cmd::kAmbient}, {"diffuse", Cmd::kDiffuse},
      {"specular", Cmd::kSpecular}, {"shininess", Cmd::kShininess}, {"roughness", Cmd::kRoughness},
      {"emission", Cmd::kEmission},
  };
  return m;
}

bool OnOff(const std::string &s, bool &out) {
  if (s == "on") { out = true; return true; }
  if (s == "off") { out = false; return true; }
  return false;
}

}  // namespace

// muon/parser.cc:180-211 (defaults) and :230-880 (command loop).
std::unique_ptr<Scene> ParseSceneText(const std::string &text, const ParseOptions &opts) {
  Workspace ws;
  ws.seedgen = std::make_unique<std::mt19937>(
      opts.default_seed >= 0 ? static_cast<unsigned int>(opts.default_seed) : std::random_device()());
  ws.material = std::make_shared<Material>();
  ws.material->ambient = V3(0.2f);
  ws.material->shininess = 1.0f;
  ws.material->roughness = 0.0f;
  ws.material->brdf = BrdfType::kPhong;
  ws.scene = std::make_unique<Scene>();
  Scene &sc = *ws.scene;
  if (opts.accel == AccelType::kLinear)
    sc.root = std::make_unique<Linear>();
  else
    sc.root = std::make_unique<BVH>(opts.partition);
  int next_id = 0;
  auto add_prim = [&](std::unique_ptr<Primitive> p) {
    p->id = next_id++;
    sc.root->AddPrimitive(std::move(p));
  };
  auto bad = [&]() { ++sc.num_warnings; };

  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    Trim(line);
    if (line.empty() || line[0] == '#') continue;
    std::istringstream iss(line);
    std::string cmd;
    iss >> cmd;
    auto it = CommandMap().find(cmd);
    if (it == CommandMap().end()) {
      bad();
      continue;
    }
    switch (it->second) {
      case Cmd::kRandomSeed: {
        unsigned int seed;
        iss >> seed;
        if (iss.fail()) { bad(); break; }
        ws.seedgen = std::make_unique<std::mt19937>(seed);
      }
        // Falls through into film_size exactly as muon/parser.cc:267-276 does.
        [[fallthrough]];
      case Cmd::kFilmSize: {
        int w, h;
        iss >> w >> h;
        if (iss.fail()) { bad(); break; }
        sc.width = w;
        sc.height = h;
        break;
      }
      case Cmd::kMinDepth: {
        int v; iss >> v;
        if (iss.fail()) { bad(); break; }
        sc.min_depth = v;
        break;
      }
      case Cmd::kMaxDepth: {
        int v; iss >> v;
        if (iss.fail()) { bad(); break; }
        sc.max_depth = v;
        break;
      }
      case Cmd::kOutput: {
        std::string v; iss >> v;
        if (iss.fail()) { bad(); break; }
        sc.output = v;
        break;
      }
      case Cmd::kGamma: {
        float v; iss >> v;
        if (iss.fail()) { bad(); break; }
        sc.gamma = v;
        break;
      }
      case Cmd::kIntegrator: {
        std::string t; iss >> t;
        if (iss.fail()) { bad(); break; }
        if (t == "normals") sc.integrator = IntegratorType::kNormals;
        else if (t == "albedo") sc.integrator = IntegratorType::kAlbedo;
        else if (t == "depth") sc.integrator = IntegratorType::kDepth;
        else if (t == "raytracer") sc.integrator = IntegratorType::kRaytracer;
        else if (t == "analyticdirect") sc.integrator = IntegratorType::kAnalyticDirect;
        else if (t == "pathtracer") {
          sc.integrator = IntegratorType::kPathTracer;
          sc.pathtracer_seed = (*ws.seedgen)();
        } else bad();
        break;
      }
      case Cmd::kPixelSamples: {
        int v; iss >> v;
        if (iss.fail()) { bad(); break; }
        sc.pixel_samples = v;
        break;
      }
      case Cmd::kLightSamples: {
        int v; iss >> v;
        if (iss.fail()) { bad(); break; }
        sc.light_samples = v;
        break;
      }
      case Cmd::kLightStratify: {
        std::string v; iss >> v;
        if (iss.fail() || !OnOff(v, sc.light_stratify)) bad();
        break;
      }
      case Cmd::kNEE: {
        std::string v; iss >> v;
        if (iss.fail()) { bad(); break; }
        if (v == "off") sc.nee = NEE::kOff;
        else if (v == "on") sc.nee = NEE::kOn;
        else if (v == "mis") sc.nee = NEE::kMIS;
        else bad();
        break;
      }
      case Cmd::kRR: {
        std::string v; iss >> v;
        if (iss.fail() || !OnOff(v, sc.russian_roulette)) bad();
        break;
      }
      case Cmd::kIS: {
        std::string v; iss >> v;
        if (iss.fail()) { bad(); break; }
        if (v == "hemisphere") sc.importance_sampling = Sampling::kHemisphere;
        else if (v == "cosine") sc.importance_sampling = Sampling::kCosine;
        else if (v == "brdf") sc.importance_sampling = Sampling::kBRDF;
        else bad();
        break;
      }
      case Cmd::kCamera: {
        float ex, ey, ez, lx, ly, lz, ux, uy, uz, fov;
        iss >> ex >> ey >> ez >> lx >> ly >> lz >> ux >> uy >> uz >> fov;
        if (iss.fail()) { bad(); break; }
        sc.camera = std::make_unique<Camera>(V3(ex, ey, ez), V3(lx, ly, lz), V3(ux, uy, uz), fov, sc.width, sc.height);
        ws.Push();
        break;
      }
      case Cmd::kLoad: {
        bad();  // assimp mesh import is outside the restated path
        break;
      }
      case Cmd::kCVN: {
        std::string v; iss >> v;
        if (iss.fail() || !OnOff(v, sc.compute_vertex_normals)) bad();
        break;
      }
      case Cmd::kSphere: {
        float x, y, z, r;
        iss >> x >> y >> z >> r;
        if (iss.fail()) { bad(); break; }
        auto s = std::make_unique<Sphere>(V3(x, y, z), r);
        ws.Update(*s);
        add_prim(std::move(s));
        break;
      }
      case Cmd::kStartMesh:
        sc.meshes.emplace_back();
        sc.cur_mesh = int(sc.meshes.size()) - 1;
        break;
      case Cmd::kEndMesh:
        sc.cur_mesh = 0;
        break;
      case Cmd::kVertex: {
        float x, y, z;
        iss >> x >> y >> z;
        if (iss.fail()) { bad(); break; }
        Vertex v;
        v.pos = V3(x, y, z);
        v.normal = V3(0.0f);
        sc.meshes[sc.cur_mesh].push_back(v);
        break;
      }
      case Cmd::kVertexNormal:
      case Cmd::kTriNormal:
        break;
      case Cmd::kTri: {
        int a, b, c;
        iss >> a >> b >> c;
        if (iss.fail()) { bad(); break; }
        auto &verts = sc.meshes[sc.cur_mesh];
        int nv = int(verts.size());
        if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) {  // out of range: UB in the reference
          bad();
          break;
        }
        auto t = std::make_unique<Tri>(verts[a], verts[b], verts[c], sc.compute_vertex_normals);
        ws.Update(*t);
        if (sc.compute_vertex_normals) {
          V3 n = t->Normal();
          verts[a].normal += n;
          verts[b].normal += n;
          verts[c].normal += n;
        }
        add_prim(std::move(t));
        break;
      }
      case Cmd::kTranslate: {
        float x, y, z;
        iss >> x >> y >> z;
        if (iss.fail()) { bad(); break; }
        ws.Multiply(translate(V3(x, y, z)));
        break;
      }
      case Cmd::kRotate: {
        float x, y, z, ang;
        iss >> x >> y >> z >> ang;
        if (iss.fail()) { bad(); break; }
        ws.Multiply(rotate(radians(ang), V3(x, y, z)));
        break;
      }
      case Cmd::kScale: {
        float x, y, z;
        iss >> x >> y >> z;
        if (iss.fail()) { bad(); break; }
        ws.Multiply(scale(V3(x, y, z)));
        break;
      }
      case Cmd::kPush: ws.Push(); break;
      case Cmd::kPop: ws.Pop(); break;
      case Cmd::kDirectional: {
        float x, y, z, r, g, b;
        iss >> x >> y >> z >> r >> g >> b;
        if (iss.fail()) { bad(); break; }
        auto l = std::make_unique<Light>();
        l->kind = LightKind::kDirectional;
        l->color = V3(r, g, b);
        l->direction = V3(x, y, z);
        sc.lights.push_back(std::move(l));
        break;
      }
      case Cmd::kPoint: {
        float x, y, z, r, g, b;
        iss >> x >> y >> z >> r >> g >> b;
        if (iss.fail()) { bad(); break; }
        auto l = std::make_unique<Light>();
        l->kind = LightKind::kPoint;
        l->color = V3(r, g, b);
        l->pos = V3(x, y, z);
        l->attenuation = sc.attenuation;
        sc.lights.push_back(std::move(l));
        break;
      }
      case Cmd::kAttenuation: {
        float c0, c1, c2;
        iss >> c0 >> c1 >> c2;
        if (iss.fail()) { bad(); break; }
        sc.attenuation = V3(c0, c1, c2);
        break;
      }
      case Cmd::kQuadLight: {
        float cx, cy, cz, e0x, e0y, e0z, e1x, e1y, e1z, r, g, b;
        iss >> cx >> cy >> cz >> e0x >> e0y >> e0z >> e1x >> e1y >> e1z >> r >> g >> b;
        if (iss.fail()) { bad(); break; }
        V3 color(r, g, b), corner(cx, cy, cz), e0(e0x, e0y, e0z), e1(e1x, e1y, e1z);
        auto l = std::make_unique<Light>();
        l->kind = LightKind::kQuad;
        l->color = color;
        l->area = Quad{corner, e0, e1};
        // muon/parser.cc:736-763: two emissive identity-transform tris.
        sc.gen_vertices.emplace_back();
        Vertex &va = sc.gen_vertices.back();
        sc.gen_vertices.emplace_back();
        Vertex &vb = sc.gen_vertices.back();
        sc.gen_vertices.emplace_back();
        Vertex &vc = sc.gen_vertices.back();
        sc.gen_vertices.emplace_back();
        Vertex &vd = sc.gen_vertices.back();
        va.pos = corner;
        vb.pos = corner + e0;
        vc.pos = corner + e1;
        vd.pos = corner + e0 + e1;
        auto t0 = std::make_unique<Tri>(va, vc, vb, false);
        auto t1 = std::make_unique<Tri>(vb, vc, vd, false);
        auto mat = std::make_shared<Material>();
        mat->brdf = BrdfType::kPhong;
        mat->emission = color;
        auto ident = std::make_shared<M4>(1.0f);
        for (Tri *t : {t0.get(), t1.get()}) {
          t->material = mat;
          t->light = l.get();
          t->transform = ident;
          t->inv_transform = ident;
          t->inv_transpose_transform = ident;
        }
        add_prim(std::move(t0));
        add_prim(std::move(t1));
        sc.lights.push_back(std::move(l));
        break;
      }
      case Cmd::kBRDF: {
        std::string v; iss >> v;
        if (iss.fail()) { bad(); break; }
        BrdfType t;
        // "lambertian" builds a Phong BRDF: the switch in muon/parser.cc:165-178
        // falls through from kLambertian into kPhong.
        if (v == "lambertian" && opts.true_lambertian) t = BrdfType::kLambertian;
        else if (v == "lambertian" || v == "phong") t = BrdfType::kPhong;
        else if (v == "ggx") t = BrdfType::kGGX;
        else { bad(); break; }
        ws.GenMaterial();
        ws.material->brdf = t;
        break;
      }
      case Cmd::kAmbient:
      case Cmd::kDiffuse:
      case Cmd::kSpecular:
      case Cmd::kEmission: {
        float r, g, b;
        iss >> r >> g >> b;
        if (iss.fail()) { bad(); break; }
        ws.GenMaterial();
        V3 c(r, g, b);
        if (it->second == Cmd::kAmbient) ws.material->ambient = c;
        else if (it->second == Cmd::kDiffuse) ws.material->diffuse = c;
        else if (it->second == Cmd::kSpecular) ws.material->specular = c;
        else ws.material->emission = c;
        break;
      }
      case Cmd::kShininess:
      case Cmd::kRoughness: {
        float v; iss >> v;
        if (iss.fail()) { bad(); break; }
        ws.GenMaterial();
        if (it->second == Cmd::kShininess) ws.material->shininess = v;
        else ws.material->roughness = v;
        break;
      }
    }
  }
  // muon/parser.cc:862-871
  if (sc.compute_vertex_normals) {
    for (auto &mesh : sc.meshes)
      for (auto &v : mesh)
        if (length2(v.normal) > 0.0f) v.normal = normalize(v.normal);
  }
  sc.root->Init();
  sc.seedgen = std::move(ws.seedgen);
  return std::move(ws.scene);
}

std::unique_ptr<Scene> ParseSceneFile(const std::string &path, const ParseOptions &opts) {
  std::ifstream f(path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ParseSceneText(ss.str(), opts);
}

// ---------------------------------------------------------------- integrators
// muon/integration.cc:19-45
V3 Integrator::Trace(const Ray &ray, V3 throughput, int depth) {
  if (scene_.max_depth != -1 &&
      depth > (scene_.nee != NEE::kOff ? scene_.max_depth - 1 : scene_.max_depth))
    return V3(0.0f);
  if (depth == 0) ++stats.primary_rays;
  else ++stats.secondary_rays;
  Intersection hit;
  if (scene_.root->Intersect(stack_, stats, ray, hit)) return Shade(hit, ray, throughput, depth);
  return V3(0.0f);
}

namespace {

// muon/integration.cc:47-51
class NormalsTracer : public Integrator {
 public:
  using Integrator::Integrator;
  std::unique_ptr<Integrator> Clone() const override { return std::make_unique<NormalsTracer>(scene_); }
 protected:
  V3 Shade(const Intersection &hit, const Ray &, V3, int) override { return hit.normal * 0.5f + 0.5f; }
};

// muon/integration.cc:57-60
class AlbedoTracer : public Integrator {
 public:
  using Integrator::Integrator;
  std::unique_ptr<Integrator> Clone() const override { return std::make_unique<AlbedoTracer>(scene_); }
 protected:
  V3 Shade(const Intersection &hit, const Ray &, V3, int) override {
    return hit.obj->material->diffuse + hit.obj->material->specular;
  }
};

// muon/integration.cc:66-72
class DepthTracer : public Integrator {
 public:
  using Integrator::Integrator;
  std::unique_ptr<Integrator> Clone() const override { return std::make_unique<DepthTracer>(scene_); }
 protected:
  V3 Shade(const Intersection &hit, const Ray &, V3, int) override { return V3(1.0f / (1.0f + hit.distance)); }
};

// muon/integration.cc:78-120
class Raytracer : public Integrator {
 public:
  using Integrator::Integrator;
  std::unique_ptr<Integrator> Clone() const override { return std::make_unique<Raytracer>(scene_); }
 protected:
  V3 Shade(const Intersection &hit, const Ray &ray, V3 throughput, int depth) override {
    const Material &m = *hit.obj->material;
    V3 color = m.ambient + m.emission;
    V3 shift = hit.pos + kEpsilon * hit.normal;
    for (const auto &light : scene_.lights) {
      ShadingInfo info = light->ShadingInfoAt(hit.pos);
      Ray shadow(shift, info.direction);
      if (Occluded(shadow, info.distance)) continue;
      V3 diffuse = m.diffuse * fmax_glm(dot(hit.normal, info.direction), 0.0f);
      V3 half = normalize(info.direction - ray.d);
      V3 spec = m.specular * std::pow(fmax_glm(dot(hit.normal, half), 0.0f), m.shininess);
      color += info.color * (diffuse + spec);
    }
    if (m.specular.x > 0.0f || m.specular.y > 0.0f || m.specular.z > 0.0f) {
      Ray refl(shift, reflect(ray.d, hit.normal));
      V3 rc = Trace(refl, throughput, depth + 1);
      color += m.specular * rc;
    }
    return color;
  }
};

// muon/integration.cc:126-199
class AnalyticDirect : public Integrator {
 public:
  using Integrator::Integrator;
  std::unique_ptr<Integrator> Clone() const override { return std::make_unique<AnalyticDirect>(scene_); }
 protected:
  V3 Shade(const Intersection &hit, const Ray &, V3, int) override {
    const Material &m = *hit.obj->material;
    V3 color = m.emission;
    for (const auto &light : scene_.lights) {
      ShadingInfo info = light->ShadingInfoAt(hit.pos);
      if (info.area == nullptr) continue;
      V3 verts[4];
      verts[0] = info.area->corner;
      verts[1] = info.area->corner + info.area->edge0;
      verts[2] = info.area->corner + info.area->edge0 + info.area->edge1;
      verts[3] = info.area->corner + info.area->edge1;
      V3 brdf = m.diffuse * kOneOverPi;
      V3 radiance = info.color;
      V3 irr(0.0f);
      for (int i = 0; i < 4; ++i) {
        int ni = (i + 1) % 4;
        V3 ui = normalize(verts[i] - hit.pos);
        V3 un = normalize(verts[ni] - hit.pos);
        float theta = std::acos(dot(ui, un));
        V3 gamma = normalize(cross(ui, un));
        irr += theta * gamma;
      }
      irr *= 0.5f;
      color += brdf * radiance * dot(irr, hit.normal);
    }
    return color;
  }
};

// muon/integration.cc:332-338
float PowerHeuristic(float p0, float p1) {
  float denom = p0 * p0 + p1 * p1;
  if (denom == 0.0f) return 0.0f;
  return (p0 * p0) / denom;
}

// muon/integration.cc:205-630
class PathTracer : public Integrator {
 public:
  PathTracer(Scene &scene, unsigned int seed) : Integrator(scene), rand_(seed) {}
  PathTracer(const PathTracer &o) : Integrator(o.scene_), rand_(o.rand_) {}
  std::unique_ptr<Integrator> Clone() const override { return std::make_unique<PathTracer>(*this); }

 protected:
  // muon/integration.cc:205-271
  V3 Shade(const Intersection &hit, const Ray &ray, V3 T, int depth) override {
    V3 shift = hit.pos + kEpsilon * hit.normal;
    if (depth < scene_.min_depth) return ShadeIndirect(hit, shift, ray, T, depth);
    V3 color;
    if ((scene_.nee != NEE::kOff && depth > 0) || dot(hit.normal, -ray.d) < 0.0f)
      color = V3(0.0f);
    else
      color = T * hit.obj->material->emission;
    color += ShadeDirect(hit, shift, ray, T);
    color += ShadeIndirect(hit, shift, ray, T, depth);
    return color;
  }

 private:
  // muon/integration.cc:273-292
  V3 ShadeDirect(const Intersection &hit, V3 shift, const Ray &ray, V3 T) {
    switch (scene_.nee) {
      case NEE::kOff: return V3(0.0f);
      case NEE::kOn: return ShadeDirectNEE(hit, shift, ray, T, false);
      case NEE::kMIS: {
        V3 a = ShadeDirectNEE(hit, shift, ray, T, true);
        V3 b = ShadeDirectIS(hit, shift, ray, T);
        return a + b;
      }
    }
    return V3(0.0f);
  }

  // muon/integration.cc:294-328
  V3 ShadeIndirect(const Intersection &hit, V3 shift, const Ray &ray, V3 T, int depth) {
    V3 dir, next_T;
    float pdf;
    if (!SampleReflection(hit, ray, T, dir, pdf, next_T)) return V3(0.0f);
    if (scene_.russian_roulette) {
      float q = fmin_glm(comp_max(next_T), 1.0f);
      if (q < rand_.Next()) return V3(0.0f);
      next_T /= q;
    }
    return Trace(Ray(shift, dir), next_T, depth + 1);
  }

  // muon/integration.cc:340-466
  V3 ShadeDirectNEE(const Intersection &hit, V3 shift, const Ray &ray, V3 T, bool mis) {
    V3 color(0.0f);
    const Material &mat = *hit.obj->material;
    for (const auto &light : scene_.lights) {
      ShadingInfo info = light->ShadingInfoAt(hit.pos);
      if (info.area == nullptr) {
        Ray shadow(shift, info.direction);
        if (Occluded(shadow, info.distance)) continue;
        V3 irr = info.color;
        if (info.distance < std::numeric_limits<float>::infinity()) irr /= info.distance * info.distance;
        float cosi = fmax_glm(dot(hit.normal, info.direction), 0.0f);
        V3 f = mat.Eval(info.direction, ray.d, hit.normal);
        color += irr * cosi * f;
        continue;
      }
      float area = length(cross(info.area->edge0, info.area->edge1));
      V3 rn = normalize(cross(info.area->edge0, info.area->edge1));
      // glm::sqrt(int) resolves to std::sqrt(int), i.e. a double square root.
      int strata = scene_.light_stratify ? int(std::sqrt(scene_.light_samples)) : 1;
      int per = scene_.light_stratify ? 1 : scene_.light_samples;
      V3 sum(0.0f);
      for (int i = 0; i < strata; i++) {
        for (int j = 0; j < strata; j++) {
          for (int k = 0; k < per; k++) {
            float u = rand_.Next();
            float v = rand_.Next();
            V3 lp = info.area->corner + (i + u) / strata * info.area->edge0 +
                    (j + v) / strata * info.area->edge1;
            lp -= kEpsilon * rn;
            V3 to_light = lp - shift;
            V3 ldir = normalize(to_light);
            float ldist = length(to_light);
            Ray shadow(shift, ldir);
            if (Occluded(shadow, ldist)) continue;
            float cosi = fmax_glm(dot(hit.normal, ldir), 0.0f);
            float cosl = fmax_glm(dot(rn, ldir), 0.0f);
            float d2 = ldist * ldist;
            float g = cosi * cosl / d2;
            V3 f = mat.Eval(ldir, ray.d, hit.normal);
            V3 contrib = g * f;
            if (mis) {
              float lpdf = NEEPDF(shadow, hit.pos);
              float bpdf = ImportanceSamplingPDF(ldir, hit, ray);
              contrib *= PowerHeuristic(lpdf, bpdf);
            }
            sum += contrib;
          }
        }
      }
      sum *= area / scene_.light_samples;
      color += info.color * sum;
    }
    return T * color;
  }

  // muon/integration.cc:468-501
  V3 ShadeDirectIS(const Intersection &hit, V3 shift, const Ray &ray, V3 T) {
    V3 dir, next_T;
    float bpdf;
    if (!SampleReflection(hit, ray, T, dir, bpdf, next_T)) return V3(0.0f);
    Ray sr(shift, dir);
    ++stats.secondary_rays;
    Intersection nh;
    if (scene_.root->Intersect(stack_, stats, sr, nh) && nh.obj->light != nullptr) {
      ShadingInfo info = nh.obj->light->ShadingInfoAt(nh.pos);
      if (info.area == nullptr) return V3(0.0f);
      float lpdf = NEEPDF(sr, hit.pos);
      float w = PowerHeuristic(bpdf, lpdf);
      return next_T * info.color * w;
    }
    return V3(0.0f);
  }

  // muon/integration.cc:503-562
  bool SampleReflection(const Intersection &hit, const Ray &ray, V3 T, V3 &dir, float &pdf, V3 &next_T) {
    const Material &mat = *hit.obj->material;
    switch (scene_.importance_sampling) {
      case Sampling::kHemisphere: dir = SampleHemisphere(hit.normal, rand_); break;
      case Sampling::kCosine: dir = SampleCosine(hit.normal, rand_); break;
      case Sampling::kBRDF: dir = mat.Sample(ray.d, hit.normal, rand_); break;
    }
    if (dot(dir, hit.normal) <= 0.0f) return false;
    V3 f = mat.Eval(dir, ray.d, hit.normal);
    pdf = ImportanceSamplingPDF(dir, hit, ray);
    switch (scene_.importance_sampling) {
      case Sampling::kHemisphere:
        next_T = T * 2.0f * kPi * f * fmax_glm(dot(hit.normal, dir), 0.0f);
        break;
      case Sampling::kCosine:
        next_T = T * kPi * f;
        break;
      case Sampling::kBRDF:
        next_T = T * f / pdf * fmax_glm(dot(hit.normal, dir), 0.0f);
        break;
    }
    return true;
  }

  // muon/integration.cc:564-608
  float NEEPDF(const Ray &sr, V3 hit_pos) {
    float lpdf = 0.0f;
    int affecting = 0;
    for (const auto &light : scene_.lights) {
      ShadingInfo info = light->ShadingInfoAt(hit_pos);
      if (info.area == nullptr) continue;
      ++affecting;
      Intersection lh;
      if (!info.area->Intersect(sr, lh)) continue;
      float area = length(cross(info.area->edge0, info.area->edge1));
      V3 rn = normalize(cross(info.area->edge0, info.area->edge1));
      float cosl = std::fabs(dot(rn, sr.d));
      float denom = area * cosl;
      if (denom == 0.0f) continue;
      V3 pl = lh.pos - sr.o;
      float d = length(pl);
      float d2 = d * d;
      lpdf += d2 / denom;
    }
    if (affecting == 0) return 0.0f;
    lpdf /= affecting;
    return lpdf;
  }

  // muon/integration.cc:610-630
  float ImportanceSamplingPDF(V3 dir, const Intersection &hit, const Ray &ray) {
    switch (scene_.importance_sampling) {
      case Sampling::kHemisphere: return 1.0f / (2.0f * kPi);
      case Sampling::kCosine: return fmax_glm(dot(hit.normal, dir), 0.0f) / kPi;
      case Sampling::kBRDF: return hit.obj->material->PDF(dir, ray.d, hit.normal);
    }
    return 0.0f;
  }

  UniformRandom rand_;
};

}  // namespace

std::unique_ptr<Integrator> MakeIntegrator(Scene &scene) {
  switch (scene.integrator) {
    case IntegratorType::kNormals: return std::make_unique<NormalsTracer>(scene);
    case IntegratorType::kAlbedo: return std::make_unique<AlbedoTracer>(scene);
    case IntegratorType::kDepth: return std::make_unique<DepthTracer>(scene);
    case IntegratorType::kRaytracer: return std::make_unique<Raytracer>(scene);
    case IntegratorType::kAnalyticDirect: return std::make_unique<AnalyticDirect>(scene);
    case IntegratorType::kPathTracer: return std::make_unique<PathTracer>(scene, scene.pathtracer_seed);
  }
  return nullptr;
}

// ---------------------------------------------------------------- film
// muon/film.cc:10-19 (the bounds check throws in the reference)
void Film::SetPixel(int x, int y, V3 c) {
  if (x < 0 || y < 0 || x >= width || y >= height) return;
  float *p = &accum[(size_t(y) * width + x) * 3];
  p[0] = p[0] + c.x;
  p[1] = p[1] + c.y;
  p[2] = p[2] + c.z;
}

// muon/film.cc:21-43: /spp, pow(v, 1/gamma), clamp [0,1], *255 truncated to u8.
void Film::ToRGB8(uint8_t *out) const {
  for (size_t i = 0; i < size_t(width) * height; ++i) {
    for (int k = 0; k < 3; ++k) {
      float v = accum[i * 3 + k] / static_cast<float>(spp);
      v = std::pow(v, 1.0f / gamma);
      v = clamp_glm(v, 0.0f, 1.0f);
      float s = v * 255;
      out[i * 3 + k] = static_cast<uint8_t>(s);
    }
  }
}

// ---------------------------------------------------------------- renderer
// muon/sampling.cc:9-26
int NumTiles(int width, int height, int pixel_samples, int parallelism) {
  uint64_t total = width * height * static_cast<uint64_t>(pixel_samples);
  constexpr uint64_t div = 500 * 50 * 2;
  int by_samples = int(total / div);
  int by_threads = parallelism * 3;
  return std::min(std::max(by_samples, by_threads), height);
}

// muon/sampling.cc:28-63 (tile 0 takes the leftover rows and seed 0)
std::vector<Tile> TileImage(int width, int height, int num_tiles, std::mt19937 &seedgen) {
  std::vector<Tile> tiles;
  int th = height / num_tiles;
  int excess = height - th * num_tiles;
  tiles.push_back(Tile{0, 0, 0, width, th + excess, 0u});
  for (int i = 1; i < num_tiles; ++i) tiles.push_back(Tile{i, 0, i * th + excess, width, th, static_cast<unsigned int>(seedgen())});
  return tiles;
}

namespace {
// muon/sampling.cc:80-121
class Sampler {
 public:
  Sampler(Tile t, int spp) : tile_(t), spp_(spp), rand_(t.random_seed) {}
  bool Next(float &x, float &y) {
    if (cy_ == tile_.height) return false;
    if (cs_ == 0) {
      x = tile_.x + cx_ + 0.5f;
      y = tile_.y + cy_ + 0.5f;
    } else {
      float u = rand_.Next();
      float v = rand_.Next();
      x = tile_.x + cx_ + u;
      y = tile_.y + cy_ + v;
      if (x == (tile_.x + cx_ + 1.0f)) x = std::nextafter(x, -std::numeric_limits<float>::infinity());
      if (y == (tile_.y + cy_ + 1.0f)) y = std::nextafter(y, -std::numeric_limits<float>::infinity());
    }
    if (++cs_ == spp_) {
      ++cx_;
      cs_ = 0;
    }
    if (cx_ == tile_.width) {
      ++cy_;
      cx_ = 0;
    }
    return true;
  }
 private:
  Tile tile_;
  int spp_;
  int cx_ = 0, cy_ = 0, cs_ = 0;
  UniformRandom rand_;
};
}  // namespace

// muon/renderer.cc:18-85 (Render time = Stop - BuildComplete)
RenderResult Render(Scene &scene, int parallelism, Film &film) {
  RenderResult res;
  if (!scene.camera) return res;  // the reference dereferences a null camera
  auto t0 = std::chrono::steady_clock::now();
  int num_tiles = NumTiles(scene.width, scene.height, scene.pixel_samples, parallelism);
  std::vector<Tile> tiles = TileImage(scene.width, scene.height, num_tiles, *scene.seedgen);
  std::mutex qmu, smu;
  auto proto = MakeIntegrator(scene);
  std::vector<std::thread> threads;
  for (int ti = 0; ti < parallelism; ++ti) {
    threads.emplace_back([&]() {
      auto integ = proto->Clone();
      while (true) {
        Tile tile;
        {
          std::lock_guard<std::mutex> g(qmu);
          if (tiles.empty()) break;
          tile = tiles.back();
          tiles.pop_back();
        }
        Sampler sampler(tile, scene.pixel_samples);
        float x, y;
        while (sampler.Next(x, y)) {
          Ray r = scene.camera->CastRay(x, y);
          V3 c = integ->Trace(r);
          film.SetPixel(int(x), int(y), c);
        }
      }
      std::lock_guard<std::mutex> g(smu);
      res.stats += integ->stats;
    });
  }
  for (auto &t : threads) t.join();
  auto t1 = std::chrono::steady_clock::now();
  res.render_seconds = std::chrono::duration<double>(t1 - t0).count();
  return res;
}

}  // namespace oracle
