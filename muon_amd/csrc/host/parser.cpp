// .muon scene parser: the command semantics of muon/parser.cc:230-880,
// writing straight into flat primitive / transform / material / light arrays
// instead of a graph of heap objects. Quirks kept on purpose (they change
// results): random_seed falls through into film_size (:267-276), camera
// pushes the transform stack (:461), "lambertian" builds a Phong BRDF
// (:165-178), quad_light adds two identity-transform emissive tris (:736-763).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <deque>
#include <map>
#include <random>
#include <sstream>
#include <unordered_map>

#include "host_scene.h"
#include "mesh_io.h"
#include "numbers.h"
#include "xform.h"

namespace muonx {
namespace host {
namespace {

struct Vertex {
  V3 pos, normal;
};

struct TriRef {
  int prim;
  int mesh;
  int idx[3];
};

// Bit-identical deduplication of small float records.
template <size_t N>
struct RecordTable {
  std::vector<std::array<float, N>> *items;
  std::map<std::array<uint32_t, N>, int32_t> index;
  int32_t Get(const std::array<float, N> &rec) {
    std::array<uint32_t, N> key;
    std::memcpy(key.data(), rec.data(), sizeof(float) * N);
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    int32_t id = int32_t(items->size());
    items->push_back(rec);
    index.emplace(key, id);
    return id;
  }
};

struct MaterialState {
  float ambient[3] = {0.2f, 0.2f, 0.2f}, diffuse[3] = {0, 0, 0}, specular[3] = {0, 0, 0},
        emission[3] = {0, 0, 0};
  float shininess = 1.0f, roughness = 0.0f;  // muon/defaults.h:57-67
  int brdf = MUONX_BRDF_PHONG;
  std::array<float, 16> Record() const {
    std::array<float, 16> r{};
    for (int i = 0; i < 3; ++i) {
      r[i] = ambient[i];
      r[3 + i] = diffuse[i];
      r[6 + i] = specular[i];
      r[9 + i] = emission[i];
    }
    r[12] = shininess;
    r[13] = roughness;
    r[14] = float(brdf);
    return r;
  }
};

bool ParseOnOff(std::istringstream &iss, bool &out) {
  std::string v;
  iss >> v;
  if (iss.fail()) return false;
  if (v == "on") out = true;
  else if (v == "off") out = false;
  else return false;
  return true;
}

}  // namespace

void ParseScene(const std::string &text, int acceleration, int partition, int64_t default_seed,
                HostScene &sc, const std::string &base_dir, bool build_bvh) {
  std::mt19937 seedgen(default_seed >= 0 ? static_cast<uint32_t>(default_seed) : std::random_device()());
  MaterialState mat;
  std::vector<M4> stack{Identity()};
  M4 inv = Identity(), inv_t = Identity();
  // record ids of the current transform and material, looked up again only
  // after the state changes
  int32_t xf_id = -1, mat_id = -1;
  auto update_cached = [&]() {
    inv = Inverse(stack.back());
    inv_t = Transpose(inv);
    xf_id = -1;
  };
  RecordTable<48> xforms{&sc.transforms, {}};
  RecordTable<16> mats{&sc.materials, {}};
  auto current_xform = [&]() {
    if (xf_id < 0) {
      std::array<float, 48> r;
      std::copy(stack.back().begin(), stack.back().end(), r.begin());
      std::copy(inv.begin(), inv.end(), r.begin() + 16);
      std::copy(inv_t.begin(), inv_t.end(), r.begin() + 32);
      xf_id = xforms.Get(r);
    }
    return xf_id;
  };
  auto current_mat = [&]() {
    if (mat_id < 0) mat_id = mats.Get(mat.Record());
    return mat_id;
  };
  std::vector<std::deque<Vertex>> meshes(1);
  int cur_mesh = 0;
  // vertices of loaded meshes: outside `meshes`, so the final normalisation
  // pass does not reach them (muon/scene.cc:9-14, muon/parser.cc:862-871)
  std::deque<Vertex> gen;
  std::vector<TriRef> tri_refs;
  float attenuation[3] = {1.0f, 0.0f, 0.0f};

  {
    // capacity for the primitives of tri / sphere lines (a hint: quad
    // lights and loaded meshes add more)
    size_t est = 0;
    for (size_t pos = 0; pos < text.size();) {
      size_t q = pos;
      while (q < text.size() && (text[q] == ' ' || text[q] == '\t')) ++q;
      if (text.compare(q, 4, "tri ") == 0 || text.compare(q, 7, "sphere ") == 0) ++est;
      const size_t nl = text.find('\n', pos);
      pos = nl == std::string::npos ? text.size() : nl + 1;
    }
    sc.prim_type.reserve(est);
    sc.prim_geom.reserve(12 * est);
    sc.prim_vnormal.reserve(9 * est);
    sc.prim_use_vnormal.reserve(est);
    sc.prim_transform.reserve(est);
    sc.prim_material.reserve(est);
    sc.prim_light.reserve(est);
  }
  auto add_prim = [&](uint8_t type, const float *geom, int ngeom, int32_t xf, int32_t m, int32_t light) {
    int id = int(sc.prim_type.size());
    sc.prim_type.push_back(type);
    const size_t g0 = sc.prim_geom.size();
    sc.prim_geom.resize(g0 + 12, 0.0f);
    std::memcpy(sc.prim_geom.data() + g0, geom, sizeof(float) * size_t(std::min(ngeom, 12)));
    sc.prim_vnormal.resize(sc.prim_vnormal.size() + 9, 0.0f);
    sc.prim_use_vnormal.push_back(0);
    sc.prim_transform.push_back(xf);
    sc.prim_material.push_back(m);
    sc.prim_light.push_back(light);
    return id;
  };
  auto bad = [&]() { ++sc.warnings; };
  // tri a b c on the current mesh; false for an index out of range
  auto add_tri = [&](int a, int b, int c) {
    auto &verts = meshes[cur_mesh];
    const int nv = int(verts.size());
    if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) return false;
    float g[9];
    const Vertex *vs[3] = {&verts[a], &verts[b], &verts[c]};
    for (int k = 0; k < 3; ++k) {
      g[3 * k] = vs[k]->pos.x;
      g[3 * k + 1] = vs[k]->pos.y;
      g[3 * k + 2] = vs[k]->pos.z;
    }
    int id = add_prim(MUONX_PRIM_TRI, g, 9, current_xform(), current_mat(), -1);
    if (sc.compute_vertex_normals) {
      // muon/parser.cc:630-637: accumulate the unnormalised face normal.
      V3 n = cross(vs[1]->pos - vs[0]->pos, vs[2]->pos - vs[0]->pos);
      verts[a].normal = verts[a].normal + n;
      verts[b].normal = verts[b].normal + n;
      verts[c].normal = verts[c].normal + n;
      sc.prim_use_vnormal[id] = 1;
    }
    tri_refs.push_back(TriRef{id, cur_mesh, {a, b, c}});
    return true;
  };

  // lines as std::getline splits them, trimmed; vertex and tri lines of
  // plain numbers skip the stream (large meshes are millions of them)
  const size_t tn = text.size();
  size_t next = 0;
  std::string line;
  while (next < tn) {
    size_t lb = next, le = text.find('\n', next);
    if (le == std::string::npos) le = tn;
    next = le + 1;
    while (lb < le && std::isspace(static_cast<unsigned char>(text[lb]))) ++lb;
    while (le > lb && std::isspace(static_cast<unsigned char>(text[le - 1]))) --le;
    if (lb == le || text[lb] == '#') continue;
    {
      const char *p = text.data() + lb, *pe = text.data() + le;
      const char *ce = p;
      while (ce < pe && !std::isspace(static_cast<unsigned char>(*ce))) ++ce;
      const size_t cl = size_t(ce - p);
      if (cl == 6 && std::memcmp(p, "vertex", 6) == 0) {
        float x[3];
        if (FastNumbers(ce, pe, x, 3)) {
          meshes[cur_mesh].push_back(Vertex{V3{x[0], x[1], x[2]}, V3{0, 0, 0}});
          continue;
        }
      } else if (cl == 3 && std::memcmp(p, "tri", 3) == 0) {
        int t[3];
        if (FastNumbers(ce, pe, t, 3)) {
          if (!add_tri(t[0], t[1], t[2])) bad();
          continue;
        }
      }
    }
    line.assign(text, lb, le - lb);
    std::istringstream iss(line);
    std::string cmd;
    iss >> cmd;

    if (cmd == "random_seed" || cmd == "film_size") {
      if (cmd == "random_seed") {
        unsigned int seed;
        iss >> seed;
        if (iss.fail()) { bad(); continue; }
        seedgen.seed(seed);
        // falls through into film_size on the rest of the line
      }
      int w, h;
      iss >> w >> h;
      if (iss.fail()) { bad(); continue; }
      sc.width = w;
      sc.height = h;
    } else if (cmd == "min_depth" || cmd == "max_depth" || cmd == "pixel_samples" || cmd == "light_samples") {
      int v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      if (cmd == "min_depth") sc.min_depth = v;
      else if (cmd == "max_depth") sc.max_depth = v;
      else if (cmd == "pixel_samples") sc.pixel_samples = v;
      else sc.light_samples = v;
    } else if (cmd == "output") {
      std::string v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      sc.output = v;
    } else if (cmd == "gamma") {
      float v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      sc.gamma = v;
    } else if (cmd == "integrator") {
      std::string t;
      iss >> t;
      if (iss.fail()) { bad(); continue; }
      if (t == "normals") sc.integrator = MUONX_INTEGRATOR_NORMALS;
      else if (t == "albedo") sc.integrator = MUONX_INTEGRATOR_ALBEDO;
      else if (t == "depth") sc.integrator = MUONX_INTEGRATOR_DEPTH;
      else if (t == "raytracer") sc.integrator = MUONX_INTEGRATOR_RAYTRACER;
      else if (t == "analyticdirect") sc.integrator = MUONX_INTEGRATOR_ANALYTICDIRECT;
      else if (t == "pathtracer") {
        sc.integrator = MUONX_INTEGRATOR_PATHTRACER;
        sc.path_seed = static_cast<uint32_t>(seedgen());
      } else bad();
    } else if (cmd == "light_stratify") {
      if (!ParseOnOff(iss, sc.light_stratify)) bad();
    } else if (cmd == "russian_roulette") {
      if (!ParseOnOff(iss, sc.russian_roulette)) bad();
    } else if (cmd == "compute_vertex_normals") {
      if (!ParseOnOff(iss, sc.compute_vertex_normals)) bad();
    } else if (cmd == "next_event_estimation") {
      std::string v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      if (v == "off") sc.nee = MUONX_NEE_OFF;
      else if (v == "on") sc.nee = MUONX_NEE_ON;
      else if (v == "mis") sc.nee = MUONX_NEE_MIS;
      else bad();
    } else if (cmd == "importance_sampling") {
      std::string v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      if (v == "hemisphere") sc.importance_sampling = MUONX_IS_HEMISPHERE;
      else if (v == "cosine") sc.importance_sampling = MUONX_IS_COSINE;
      else if (v == "brdf") sc.importance_sampling = MUONX_IS_BRDF;
      else bad();
    } else if (cmd == "camera") {
      float e[3], l[3], u[3], fov;
      iss >> e[0] >> e[1] >> e[2] >> l[0] >> l[1] >> l[2] >> u[0] >> u[1] >> u[2] >> fov;
      if (iss.fail()) { bad(); continue; }
      // muon/camera.cc:15-35
      float tfy = std::tan(Radians(fov) / 2.0f);
      float tfx = tfy * (static_cast<float>(sc.width) / sc.height);
      V3 w = normalize(V3{e[0], e[1], e[2]} - V3{l[0], l[1], l[2]});
      V3 cu = normalize(cross(V3{u[0], u[1], u[2]}, w));
      V3 cv = cross(w, cu);
      sc.has_camera = true;
      for (int i = 0; i < 3; ++i) sc.eye[i] = e[i];
      sc.cam_u[0] = cu.x; sc.cam_u[1] = cu.y; sc.cam_u[2] = cu.z;
      sc.cam_v[0] = cv.x; sc.cam_v[1] = cv.y; sc.cam_v[2] = cv.z;
      sc.cam_w[0] = w.x; sc.cam_w[1] = w.y; sc.cam_w[2] = w.z;
      sc.tan_fov_x = tfx;
      sc.tan_fov_y = tfy;
      stack.push_back(stack.back());
      update_cached();
    } else if (cmd == "load") {
      // muon/parser.cc:465-561: every triangle of the file's meshes, with the
      // current transform and material; a file that cannot be read is
      // logged and skipped
      std::string fname;
      iss >> fname;
      if (iss.fail()) { bad(); continue; }
      const std::string path = (base_dir.empty() || fname[0] == '/') ? fname : base_dir + "/" + fname;
      std::vector<LoadedMesh> loaded;
      std::string err;
      if (!LoadMeshFile(path, loaded, err)) { bad(); continue; }
      for (const LoadedMesh &lm : loaded) {
        const size_t base = gen.size();
        for (const auto &p : lm.pos) gen.push_back(Vertex{V3{p[0], p[1], p[2]}, V3{0, 0, 0}});
        for (const auto &t : lm.tris) {
          const size_t ia = base + size_t(t[0]), ib = base + size_t(t[1]), ic = base + size_t(t[2]);
          Vertex *vs[3] = {&gen[ia], &gen[ib], &gen[ic]};
          float g[9];
          for (int k = 0; k < 3; ++k) {
            g[3 * k] = vs[k]->pos.x;
            g[3 * k + 1] = vs[k]->pos.y;
            g[3 * k + 2] = vs[k]->pos.z;
          }
          int id = add_prim(MUONX_PRIM_TRI, g, 9, current_xform(), current_mat(), -1);
          if (sc.compute_vertex_normals) {
            const V3 n = cross(vs[1]->pos - vs[0]->pos, vs[2]->pos - vs[0]->pos);
            for (Vertex *v : vs) v->normal = v->normal + n;
            sc.prim_use_vnormal[id] = 1;
          }
          tri_refs.push_back(TriRef{id, -1, {int(ia), int(ib), int(ic)}});
        }
      }
    } else if (cmd == "sphere") {
      float g[4];
      iss >> g[0] >> g[1] >> g[2] >> g[3];
      if (iss.fail()) { bad(); continue; }
      add_prim(MUONX_PRIM_SPHERE, g, 4, current_xform(), current_mat(), -1);
    } else if (cmd == "start_mesh") {
      meshes.emplace_back();
      cur_mesh = int(meshes.size()) - 1;
    } else if (cmd == "end_mesh") {
      cur_mesh = 0;
    } else if (cmd == "vertex") {
      float x, y, z;
      iss >> x >> y >> z;
      if (iss.fail()) { bad(); continue; }
      meshes[cur_mesh].push_back(Vertex{V3{x, y, z}, V3{0, 0, 0}});
    } else if (cmd == "vertex_normal" || cmd == "tri_normal") {
      // no-ops in the reference
    } else if (cmd == "tri") {
      int a, b, c;
      iss >> a >> b >> c;
      if (iss.fail() || !add_tri(a, b, c)) { bad(); continue; }
    } else if (cmd == "translate" || cmd == "scale") {
      float x, y, z;
      iss >> x >> y >> z;
      if (iss.fail()) { bad(); continue; }
      stack.back() = Mul(stack.back(), cmd == "translate" ? Translate(x, y, z) : Scale(x, y, z));
      update_cached();
    } else if (cmd == "rotate") {
      float x, y, z, ang;
      iss >> x >> y >> z >> ang;
      if (iss.fail()) { bad(); continue; }
      stack.back() = Mul(stack.back(), Rotate(Radians(ang), x, y, z));
      update_cached();
    } else if (cmd == "push_transform") {
      stack.push_back(stack.back());
      update_cached();
    } else if (cmd == "pop_transform") {
      if (stack.size() <= 1) { bad(); continue; }  // undefined in the reference
      stack.pop_back();
      update_cached();
    } else if (cmd == "directional_light" || cmd == "point_light") {
      float x, y, z, r, g, b;
      iss >> x >> y >> z >> r >> g >> b;
      if (iss.fail()) { bad(); continue; }
      std::array<float, 16> L{};
      L[0] = float(cmd == "point_light" ? MUONX_LIGHT_POINT : MUONX_LIGHT_DIRECTIONAL);
      L[1] = r; L[2] = g; L[3] = b;
      L[4] = x; L[5] = y; L[6] = z;
      if (cmd == "point_light") { L[7] = attenuation[0]; L[8] = attenuation[1]; L[9] = attenuation[2]; }
      sc.lights.push_back(L);
    } else if (cmd == "attenuation") {
      float c0, c1, c2;
      iss >> c0 >> c1 >> c2;
      if (iss.fail()) { bad(); continue; }
      attenuation[0] = c0; attenuation[1] = c1; attenuation[2] = c2;
    } else if (cmd == "quad_light") {
      float v[12];
      for (float &x : v) iss >> x;
      if (iss.fail()) { bad(); continue; }
      V3 corner{v[0], v[1], v[2]}, e0{v[3], v[4], v[5]}, e1{v[6], v[7], v[8]};
      std::array<float, 16> L{};
      L[0] = float(MUONX_LIGHT_QUAD);
      L[1] = v[9]; L[2] = v[10]; L[3] = v[11];
      for (int i = 0; i < 9; ++i) L[4 + i] = v[i];
      int32_t light_id = int32_t(sc.lights.size());
      sc.lights.push_back(L);
      V3 va = corner, vb = corner + e0, vc = corner + e1, vd = corner + e0 + e1;
      MaterialState lm;
      for (int i = 0; i < 3; ++i) lm.ambient[i] = 0.0f;
      lm.shininess = 0.0f;
      lm.emission[0] = v[9]; lm.emission[1] = v[10]; lm.emission[2] = v[11];
      std::array<float, 48> ident{};
      const M4 I = Identity();
      for (int k = 0; k < 3; ++k) std::copy(I.begin(), I.end(), ident.begin() + 16 * k);
      int32_t xf = xforms.Get(ident);
      int32_t m = mats.Get(lm.Record());
      const float t0[9] = {va.x, va.y, va.z, vc.x, vc.y, vc.z, vb.x, vb.y, vb.z};
      const float t1[9] = {vb.x, vb.y, vb.z, vc.x, vc.y, vc.z, vd.x, vd.y, vd.z};
      add_prim(MUONX_PRIM_TRI, t0, 9, xf, m, light_id);
      add_prim(MUONX_PRIM_TRI, t1, 9, xf, m, light_id);
    } else if (cmd == "brdf") {
      mat_id = -1;
      std::string v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      if (v == "lambertian" || v == "phong") mat.brdf = MUONX_BRDF_PHONG;
      else if (v == "ggx") mat.brdf = MUONX_BRDF_GGX;
      else bad();
    } else if (cmd == "ambient" || cmd == "diffuse" || cmd == "specular" || cmd == "emission") {
      mat_id = -1;
      float r, g, b;
      iss >> r >> g >> b;
      if (iss.fail()) { bad(); continue; }
      float *dst = cmd == "ambient" ? mat.ambient : cmd == "diffuse" ? mat.diffuse
                 : cmd == "specular" ? mat.specular : mat.emission;
      dst[0] = r; dst[1] = g; dst[2] = b;
    } else if (cmd == "shininess" || cmd == "roughness") {
      mat_id = -1;
      float v;
      iss >> v;
      if (iss.fail()) { bad(); continue; }
      if (cmd == "shininess") mat.shininess = v;
      else mat.roughness = v;
    } else {
      bad();  // unknown command: logged and skipped by the reference
    }
  }

  // muon/parser.cc:862-871: normalise accumulated vertex normals.
  if (sc.compute_vertex_normals) {
    for (auto &mesh : meshes)
      for (auto &v : mesh)
        if (dot(v.normal, v.normal) > 0.0f) v.normal = normalize(v.normal);
  }
  for (const TriRef &t : tri_refs) {
    for (int k = 0; k < 3; ++k) {
      const Vertex &v = t.mesh < 0 ? gen[size_t(t.idx[k])] : meshes[t.mesh][t.idx[k]];
      sc.prim_vnormal[t.prim * 9 + 3 * k] = v.normal.x;
      sc.prim_vnormal[t.prim * 9 + 3 * k + 1] = v.normal.y;
      sc.prim_vnormal[t.prim * 9 + 3 * k + 2] = v.normal.z;
    }
  }
  if (build_bvh || acceleration == 0) {
    BuildBVH(sc, acceleration, partition);
  } else {
    sc.acceleration = 1;  // the BVH is built on the device at upload
  }
}

muonx_scene_desc HostScene::Desc() const {
  muonx_scene_desc d{};
  d.num_prims = uint32_t(num_prims());
  d.prim_type = prim_type.data();
  d.prim_geom = prim_geom.data();
  d.prim_vnormal = prim_vnormal.data();
  d.prim_use_vnormal = prim_use_vnormal.data();
  d.prim_transform = prim_transform.data();
  d.prim_material = prim_material.data();
  d.prim_light = prim_light.data();
  d.num_transforms = uint32_t(transforms.size());
  d.transforms = transforms.empty() ? nullptr : transforms[0].data();
  d.num_materials = uint32_t(materials.size());
  d.materials = materials.empty() ? nullptr : materials[0].data();
  d.num_lights = uint32_t(lights.size());
  d.lights = lights.empty() ? nullptr : lights[0].data();
  d.num_nodes = uint32_t(node_left.size());
  d.node_bounds = node_bounds.data();
  d.node_left = node_left.data();
  d.node_right = node_right.data();
  d.node_axis = node_axis.data();
  d.node_start = node_start.data();
  d.node_count = node_count.data();
  d.leaf_prims = leaf_prims.data();
  d.acceleration = acceleration;
  return d;
}

muonx_render_params HostScene::Params() const {
  muonx_render_params p{};
  p.width = width;
  p.height = height;
  p.spp = pixel_samples;
  p.min_depth = min_depth;
  p.max_depth = max_depth;
  p.integrator = integrator;
  p.nee = nee;
  p.importance_sampling = importance_sampling;
  p.russian_roulette = russian_roulette ? 1 : 0;
  p.light_samples = light_samples;
  p.light_stratify = light_stratify ? 1 : 0;
  p.seed = path_seed;
  p.gamma = gamma;
  for (int i = 0; i < 3; ++i) {
    p.eye[i] = eye[i];
    p.cam_u[i] = cam_u[i];
    p.cam_v[i] = cam_v[i];
    p.cam_w[i] = cam_w[i];
  }
  p.tan_fov_x = tan_fov_x;
  p.tan_fov_y = tan_fov_y;
  p.has_camera = has_camera ? 1 : 0;
  p.sample_mode = 0;
  return p;
}

}  // namespace host
}  // namespace muonx
