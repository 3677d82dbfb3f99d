// Host-side scene produced by the .muon parser and the SAH builder: flat
// arrays in the layout of muonx_scene_desc (include/muonx.h), plus the
// render settings and the precomputed camera.
#ifndef MUONX_HOST_HOST_SCENE_H_
#define MUONX_HOST_HOST_SCENE_H_

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/muonx.h"

namespace muonx {
namespace host {

struct HostScene {
  // Render settings; defaults from muon/defaults.h:15-70.
  int width = 150, height = 150, min_depth = 0, max_depth = 5;
  std::string output = "raytrace.png";
  float gamma = 1.0f;
  bool compute_vertex_normals = false;
  int pixel_samples = 1, light_samples = 1;
  bool light_stratify = false;
  int nee = MUONX_NEE_OFF;
  bool russian_roulette = false;
  int importance_sampling = MUONX_IS_HEMISPHERE;
  int integrator = MUONX_INTEGRATOR_RAYTRACER;  // muon/parser.cc:195
  uint32_t path_seed = 0;
  // Camera, precomputed as muon/camera.cc:15-35 does.
  bool has_camera = false;
  float eye[3] = {0, 0, 0}, cam_u[3] = {0, 0, 0}, cam_v[3] = {0, 0, 0}, cam_w[3] = {0, 0, 0};
  float tan_fov_x = 0.0f, tan_fov_y = 0.0f;
  int warnings = 0;

  // Primitives in parse order.
  std::vector<uint8_t> prim_type;
  std::vector<float> prim_geom;      // 12 per prim
  std::vector<float> prim_vnormal;   // 9 per prim
  std::vector<uint8_t> prim_use_vnormal;
  std::vector<int32_t> prim_transform, prim_material, prim_light;

  std::vector<std::array<float, 48>> transforms;
  std::vector<std::array<float, 16>> materials;
  std::vector<std::array<float, 16>> lights;

  // Binary SAH BVH, pre-order.
  std::vector<float> node_bounds;
  std::vector<int32_t> node_left, node_right, node_axis, node_start, node_count;
  std::vector<int32_t> leaf_prims;
  int bvh_depth = 0;
  int acceleration = 1;  // 0: linear structure (one leaf in parse order)

  size_t num_prims() const { return prim_type.size(); }
  muonx_scene_desc Desc() const;
  muonx_render_params Params() const;
};

// Parses .muon text (muon/parser.cc:230-880) and builds the BVH.
// partition: 0 uniform, 1 midpoint, 2 SAH; acceleration 0 = linear (one leaf).
// `load` paths are taken relative to base_dir (the scene file's directory;
// empty: the working directory).
// With build_bvh false (and the BVH structure) the tree is left empty for
// the device to build.
void ParseScene(const std::string &text, int acceleration, int partition, int64_t default_seed,
                HostScene &out, const std::string &base_dir = std::string(), bool build_bvh = true);

// Builds the binary BVH of muon/acceleration.cc:204-458 into `sc`.
void BuildBVH(HostScene &sc, int acceleration, int partition);

}  // namespace host
}  // namespace muonx

#endif
