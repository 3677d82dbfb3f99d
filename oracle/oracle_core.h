// ORACLE — test infrastructure only (see omath.h). A CPU restatement of the
// reference renderer (voithos/muon) kept op-for-op with its C++ sources so
// that it can serve as the parity oracle and the timed CPU baseline. It
// keeps the reference's algorithmic structure: heap-allocated binary BVH
// nodes, per-primitive object-space transforms, optional-style results and
// recursive integrators.
#ifndef MUONX_ORACLE_CORE_H_
#define MUONX_ORACLE_CORE_H_

#include <array>
#include <cstdint>
#include <deque>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "omath.h"

namespace oracle {

// muon/ray.h:9-26, muon/ray.cc:7-13
struct Ray {
  V3 o, d;
  Ray() = default;
  Ray(V3 origin, V3 dir) : o(origin), d(dir) {}
  Ray Transform(const M4 &m) const { return Ray(transform_position(m, o), transform_direction(m, d)); }
  V3 At(float t) const { return o + d * t; }
};

// muon/stats.h:11-44, plus a shadow-ray counter the reference does not keep.
struct TraceStats {
  uint64_t primary_rays = 0, secondary_rays = 0, object_tests = 0, object_hits = 0,
           bounds_tests = 0, bounds_hits = 0, shadow_rays = 0;
  TraceStats &operator+=(const TraceStats &r) {
    primary_rays += r.primary_rays; secondary_rays += r.secondary_rays;
    object_tests += r.object_tests; object_hits += r.object_hits;
    bounds_tests += r.bounds_tests; bounds_hits += r.bounds_hits;
    shadow_rays += r.shadow_rays;
    return *this;
  }
};

// muon/bounds.{h,cc}
struct Bounds {
  V3 mn = V3(std::numeric_limits<float>::infinity());
  V3 mx = V3(-std::numeric_limits<float>::infinity());
  Bounds() = default;
  explicit Bounds(V3 p) : mn(p), mx(p) {}
  Bounds(V3 a, V3 b);
  V3 Dimensions() const { return mx - mn; }
  int MaxAxis() const;
  float SurfaceArea() const;
  Bounds Transform(const M4 &m) const;
  bool HasIntersection(const Ray &ray, float max_distance) const;
  bool Intersect(const Ray &ray, float &t_min, float &t_max) const;
  static Bounds Union(const Bounds &a, const Bounds &b);
  static Bounds Union(const Bounds &a, V3 p);
};

class Primitive;

// muon/types.h:17-28
struct Intersection {
  float distance = 0.0f;
  V3 pos, normal;
  Primitive *obj = nullptr;
};

struct Vertex {
  V3 pos, normal;
};

// Deterministic uniform [0,1) stream with the reference's exact semantics:
// std::mt19937 + std::uniform_real_distribution<float>(0,1) (muon/random.h:27-43).
class UniformRandom {
 public:
  explicit UniformRandom(unsigned int seed) : gen_(seed), dist_(0.0f, 1.0f) {}
  float Next() { return dist_(gen_); }
 private:
  std::mt19937 gen_;
  std::uniform_real_distribution<float> dist_;
};

enum class BrdfType { kLambertian = 0, kPhong = 1, kGGX = 2 };

// muon/materials.{h,cc}. The BRDF is folded into the material as a type tag;
// the lazily cached reflectiveness of the reference is a pure function of the
// material fields, so it is computed on demand.
struct Material {
  V3 ambient, diffuse, specular, emission;
  float shininess = 0.0f, roughness = 0.0f;
  BrdfType brdf = BrdfType::kPhong;

  float Reflectiveness() const;
  V3 Sample(V3 ray_dir, V3 normal, UniformRandom &rand) const;
  float PDF(V3 in_dir, V3 ray_dir, V3 normal) const;
  V3 Eval(V3 in_dir, V3 ray_dir, V3 normal) const;
};

// muon/lighting.{h,cc}
struct Quad {
  V3 corner, edge0, edge1;
  bool Intersect(const Ray &ray, Intersection &out) const;
};
enum class LightKind { kDirectional, kPoint, kQuad };
struct ShadingInfo {
  V3 color, direction;
  float distance = 0.0f;
  const Quad *area = nullptr;
};
struct Light {
  LightKind kind;
  V3 color;
  V3 direction;    // directional
  V3 pos;          // point
  V3 attenuation;  // point
  Quad area;       // quad
  ShadingInfo ShadingInfoAt(V3 p) const;
};

// muon/objects.{h,cc}
class Primitive {
 public:
  virtual ~Primitive() = default;
  virtual Bounds ObjectBounds() const = 0;
  virtual Bounds WorldBounds() const { return ObjectBounds().Transform(*transform); }
  virtual bool IntersectObjectSpace(const Ray &ray, Intersection &out) = 0;
  bool Intersect(const Ray &ray, Intersection &out);
  bool HasIntersection(const Ray &ray, float max_distance);

  std::shared_ptr<M4> transform, inv_transform, inv_transpose_transform;
  std::shared_ptr<Material> material;
  const Light *light = nullptr;
  int id = -1;        // parse (AddPrimitive) order
  bool is_sphere = false;
};

class Tri : public Primitive {
 public:
  Tri(Vertex &v0, Vertex &v1, Vertex &v2, bool use_vertex_normals);
  bool IntersectObjectSpace(const Ray &ray, Intersection &out) override;
  Bounds ObjectBounds() const override;
  Bounds WorldBounds() const override;
  V3 Normal() const { return normal_; }
  Vertex &v0_, &v1_, &v2_;
  V3 normal_;
  float normal_length2_;
  bool use_vertex_normals_;
};

class Sphere : public Primitive {
 public:
  Sphere(V3 pos, float radius) : pos_(pos), radius_(radius) { is_sphere = true; }
  bool IntersectObjectSpace(const Ray &ray, Intersection &out) override;
  Bounds ObjectBounds() const override { return Bounds(pos_ - V3(radius_), pos_ + V3(radius_)); }
  V3 pos_;
  float radius_;
};

// muon/acceleration.{h,cc}
enum class AccelType { kLinear = 0, kBVH = 1 };
enum class Partition { kUniform = 0, kMidpoint = 1, kSAH = 2 };

struct BVHNode {
  size_t num_primitives = 0;
  size_t start = 0;
  int axis = 0;
  std::array<std::unique_ptr<BVHNode>, 2> children;
  Bounds bounds;
};

struct PrimitiveInfo {
  size_t original_index;
  Bounds bounds;
  V3 centroid;
  PrimitiveInfo(size_t idx, const Bounds &b);
};

class Structure {
 public:
  virtual ~Structure() = default;
  void AddPrimitive(std::unique_ptr<Primitive> p) { primitives_.push_back(std::move(p)); }
  virtual void Init() = 0;
  virtual bool Intersect(std::vector<BVHNode *> &stack, TraceStats &st, const Ray &ray,
                         Intersection &out) const = 0;
  virtual bool HasIntersection(std::vector<BVHNode *> &stack, TraceStats &st, const Ray &ray,
                               float max_distance) const = 0;
  std::vector<std::unique_ptr<Primitive>> primitives_;
};

class Linear : public Structure {
 public:
  void Init() override {}
  bool Intersect(std::vector<BVHNode *> &, TraceStats &st, const Ray &ray,
                 Intersection &out) const override;
  bool HasIntersection(std::vector<BVHNode *> &, TraceStats &st, const Ray &ray,
                       float max_distance) const override;
};

class BVH : public Structure {
 public:
  explicit BVH(Partition p) : partition_(p) {}
  void Init() override;
  bool Intersect(std::vector<BVHNode *> &stack, TraceStats &st, const Ray &ray,
                 Intersection &out) const override;
  bool HasIntersection(std::vector<BVHNode *> &stack, TraceStats &st, const Ray &ray,
                       float max_distance) const override;
  const BVHNode *root() const { return root_.get(); }

 private:
  std::unique_ptr<BVHNode> Build(size_t start, size_t end, std::vector<PrimitiveInfo> &info) const;
  Partition partition_;
  std::unique_ptr<BVHNode> root_;
};

// muon/camera.{h,cc}
struct Camera {
  V3 eye, look_at, up;
  float fov = 0.0f, tan_fov_x = 0.0f, tan_fov_y = 0.0f;
  int width = 0, height = 0;
  V3 u, v, w;
  Camera(V3 eye, V3 look_at, V3 up, float fov, int width, int height);
  Ray CastRay(float x, float y) const;
};

enum class NEE { kOff = 0, kOn = 1, kMIS = 2 };
enum class Sampling { kHemisphere = 0, kCosine = 1, kBRDF = 2 };
enum class IntegratorType { kNormals, kAlbedo, kDepth, kRaytracer, kAnalyticDirect, kPathTracer };

// muon/scene.{h,cc} together with the parser's result (muon/parser.h:61-67).
struct Scene {
  int width = 150, height = 150, min_depth = 0, max_depth = 5;
  std::string output = "raytrace.png";
  float gamma = 1.0f;
  bool compute_vertex_normals = false;
  int pixel_samples = 1, light_samples = 1;
  bool light_stratify = false;
  NEE nee = NEE::kOff;
  bool russian_roulette = false;
  Sampling importance_sampling = Sampling::kHemisphere;
  V3 attenuation = V3(1.0f, 0.0f, 0.0f);
  std::unique_ptr<Camera> camera;
  std::unique_ptr<Structure> root;
  std::vector<std::deque<Vertex>> meshes = std::vector<std::deque<Vertex>>(1);
  int cur_mesh = 0;
  std::deque<Vertex> gen_vertices;
  std::vector<std::unique_ptr<Light>> lights;
  // Seed generator state after parsing (muon/parser.cc:874).
  std::unique_ptr<std::mt19937> seedgen;
  IntegratorType integrator = IntegratorType::kRaytracer;
  unsigned int pathtracer_seed = 0;
  int num_warnings = 0;
};

struct ParseOptions {
  AccelType accel = AccelType::kBVH;
  Partition partition = Partition::kSAH;
  // When >= 0, replaces the std::random_device seed of the initial seed
  // generator (muon/parser.cc:181-182) so that runs are reproducible.
  int64_t default_seed = -1;
  // Test hook, not a reference behaviour: `brdf lambertian` builds the
  // Lambertian BRDF (muon/materials.cc:17-47) instead of the Phong BRDF the
  // reference's parser falls through to, so that the Lambertian restatement
  // can be checked.
  bool true_lambertian = false;
};

std::unique_ptr<Scene> ParseSceneFile(const std::string &path, const ParseOptions &opts);
std::unique_ptr<Scene> ParseSceneText(const std::string &text, const ParseOptions &opts);

// Integrators (muon/integration.{h,cc}).
class Integrator {
 public:
  explicit Integrator(Scene &scene) : scene_(scene) {}
  virtual ~Integrator() = default;
  V3 Trace(const Ray &ray) { return Trace(ray, V3(1.0f), 0); }
  V3 Trace(const Ray &ray, V3 throughput, int depth);
  virtual std::unique_ptr<Integrator> Clone() const = 0;
  TraceStats stats;

 protected:
  virtual V3 Shade(const Intersection &hit, const Ray &ray, V3 throughput, int depth) = 0;
  bool Occluded(const Ray &ray, float max_distance) {
    ++stats.shadow_rays;
    return scene_.root->HasIntersection(stack_, stats, ray, max_distance);
  }
  Scene &scene_;
  std::vector<BVHNode *> stack_;
};

std::unique_ptr<Integrator> MakeIntegrator(Scene &scene);

// muon/film.{h,cc}; the accumulator is exposed row-major [y][x][c].
struct Film {
  int width, height, spp;
  float gamma;
  std::vector<float> accum;  // H*W*3
  Film(int w, int h, int s, float g) : width(w), height(h), spp(s), gamma(g), accum(size_t(w) * h * 3, 0.0f) {}
  void SetPixel(int x, int y, V3 c);
  void ToRGB8(uint8_t *out) const;
};

// muon/renderer.cc:18-85 with muon/sampling.cc tiles / sampler.
struct RenderResult {
  TraceStats stats;
  double render_seconds = 0.0;
};
RenderResult Render(Scene &scene, int parallelism, Film &film);

// Tile helpers, muon/sampling.cc:9-63.
struct Tile {
  int idx, x, y, width, height;
  unsigned int random_seed;
};
int NumTiles(int width, int height, int pixel_samples, int parallelism);
std::vector<Tile> TileImage(int width, int height, int num_tiles, std::mt19937 &seedgen);

}  // namespace oracle

#endif
