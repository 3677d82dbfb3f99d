// muonx::Renderer — the host shell of muon::Renderer (muon/renderer.h:11-24)
// over the C ABI: parse, upload, render on the GPU, write the film.
#ifndef MUONX_HOST_RENDERER_H_
#define MUONX_HOST_RENDERER_H_

#include <cstdint>
#include <string>

namespace muonx {

// muon::Options (muon/options.h:11-23) plus the device ordinal.
struct Options {
  std::string output;
  int acceleration = 1;        // 0 linear, 1 bvh
  int partition_strategy = 2;  // 0 uniform, 1 midpoint, 2 sah
  uint32_t parallelism = 1;    // host-side threads (the render itself runs on the GPU)
  bool show_stats = true;
  int64_t default_seed = -1;
  int device = 0;
  bool enable_feexcept = false;  // muon/debug.cc:7-17 (--enable_feexcept)
};

class Renderer {
 public:
  // Stores `options` by reference, as muon::Renderer does: keep it alive.
  Renderer(std::string scene_file, const Options &options) : scene_file_(std::move(scene_file)), options_(options) {}
  // Returns 0 on success; errors are printed to stderr.
  int Render() const;

 private:
  std::string scene_file_;
  const Options &options_;
};

}  // namespace muonx

#endif
