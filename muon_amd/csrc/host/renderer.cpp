#include "renderer.h"

#include <cfenv>
#include <chrono>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <vector>

#include "../../../include/muonx.h"

namespace muonx {
namespace {

int Fail(const char *what) {
  std::cerr << "muonx: " << what << ": " << muonx_last_error() << std::endl;
  return 1;
}

// Floating-point exception traps (FE_DIVBYZERO | FE_INVALID) switched on
// for host-side stretches only, so that the GPU runtime's own host code
// runs with the default environment.
class HostFpTraps {
 public:
  explicit HostFpTraps(bool enabled) : enabled_(enabled) { On(); }
  ~HostFpTraps() { Off(); }
  void On() {
    if (enabled_) feenableexcept(FE_DIVBYZERO | FE_INVALID);
  }
  void Off() {
    if (enabled_) fedisableexcept(FE_DIVBYZERO | FE_INVALID);
  }

 private:
  bool enabled_;
};

}  // namespace

// muon/renderer.cc:18-85 with the tile loop replaced by muonx_render.
int Renderer::Render() const {
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  muonx_options o{options_.acceleration, options_.partition_strategy, int32_t(options_.parallelism),
                  options_.show_stats ? 1 : 0, options_.default_seed};
  muonx_host_scene *hs = nullptr;
  // muon/debug.cc:7-17: --enable_feexcept traps division by zero and
  // invalid operations; here in the host-side parse, BVH build and film
  // conversion (the render itself runs on the GPU)
  HostFpTraps traps(options_.enable_feexcept);
  const int prc = muonx_host_scene_parse_file(scene_file_.c_str(), &o, &hs);
  traps.Off();
  if (prc != MUONX_OK) return Fail("parse");
  muonx_scene_desc desc;
  muonx_render_params params;
  muonx_host_scene_desc(hs, &desc);
  muonx_host_scene_params(hs, &params);
  muonx_ctx *ctx = nullptr;
  muonx_scene *scene = nullptr;
  if (muonx_ctx_create(options_.device, &ctx) != MUONX_OK) {
    muonx_host_scene_destroy(hs);
    return Fail("device");
  }
  if (options_.show_stats) {  // the path tracer counts its tests only when asked (muon's --stats)
    muonx_tuning t;
    muonx_ctx_get_tuning(ctx, &t);
    t.count_work = 1;
    muonx_ctx_set_tuning(ctx, &t);
  }
  if (muonx_scene_create(ctx, &desc, &scene) != MUONX_OK) {
    muonx_host_scene_destroy(hs);
    muonx_ctx_destroy(ctx);
    return Fail("scene upload");
  }
  const auto t1 = Clock::now();
  std::vector<float> accum(size_t(params.width) * params.height * 3);
  muonx_trace_stats st{};
  const int rc = muonx_render_host(scene, &params, accum.data(), &st);
  const auto t2 = Clock::now();
  int result = 0;
  if (rc != MUONX_OK) {
    result = Fail("render");
  } else {
    std::vector<uint8_t> rgb(accum.size());
    traps.On();
    muonx_film_to_rgb8(accum.data(), params.width, params.height, params.spp, params.gamma, rgb.data());
    traps.Off();
    // muon/renderer.cc:26-27: --output overrides the scene's output
    std::string out = options_.output;
    if (out.empty()) {
      const char *scene_out = nullptr;
      muonx_host_scene_output(hs, &scene_out);
      out = scene_out ? scene_out : "raytrace.png";
    }
    if (muonx_write_png(out.c_str(), rgb.data(), params.width, params.height) != MUONX_OK) result = Fail("write");
  }
  if (options_.show_stats && result == 0) {
    auto secs = [](Clock::duration d) { return std::chrono::duration<double>(d).count(); };
    std::cerr << std::fixed << std::setprecision(2);
    std::cerr << "Total time         : " << std::setw(12) << secs(t2 - t0) << " (sec)\n";
    std::cerr << "Build time         : " << std::setw(12) << secs(t1 - t0) << " (sec)\n";
    std::cerr << "Render time        : " << std::setw(12) << secs(t2 - t1) << " (sec)\n";
    std::cerr << "Primary rays       : " << std::setw(12) << st.primary_rays << "\n";
    std::cerr << "Secondary rays     : " << std::setw(12) << st.secondary_rays << "\n";
    std::cerr << "Shadow rays        : " << std::setw(12) << st.shadow_rays << "\n";
    // muon/stats.cc:41-88 (counted by the deterministic integrators)
    auto rate = [](uint64_t h, uint64_t t) { return t ? 100.0 * double(h) / double(t) : 0.0; };
    std::cerr << "Ray-object tests   : " << std::setw(12) << st.object_tests << "\n";
    std::cerr << "Ray-object hits    : " << std::setw(12) << st.object_hits << "\n";
    std::cerr << "Hit rate           : " << std::setw(12) << rate(st.object_hits, st.object_tests) << " %\n";
    std::cerr << "Bounds tests       : " << std::setw(12) << st.bounds_tests << "\n";
    std::cerr << "Bounds hits        : " << std::setw(12) << st.bounds_hits << "\n";
    std::cerr << "Bounds hit rate    : " << std::setw(12) << rate(st.bounds_hits, st.bounds_tests) << " %\n";
  }
  muonx_scene_destroy(scene);
  muonx_ctx_destroy(ctx);
  muonx_host_scene_destroy(hs);
  return result;
}

}  // namespace muonx
