// muonx CLI: muon's flags (muon/muon.cc:12-21) over the GPU renderer.
//   muonx --scene=path.muon [--output=out.png] [--acceleration=bvh|linear]
//         [--partition_strategy=sah|midpoint|uniform] [--parallelism=N]
//         [--stats=true|false] [--enable_feexcept[=true|false]] [--seed=N] [--device=N]
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "renderer.h"

namespace {

bool Flag(const char *arg, const char *name, std::string &value) {
  const size_t n = std::strlen(name);
  if (std::strncmp(arg, "--", 2) != 0 || std::strncmp(arg + 2, name, n) != 0) return false;
  const char *rest = arg + 2 + n;
  if (*rest == '=') {
    value = rest + 1;
    return true;
  }
  if (*rest == '\0') {
    value = "true";
    return true;
  }
  return false;
}

}  // namespace

int main(int argc, char **argv) {
  std::string scene;
  muonx::Options opts;
  for (int i = 1; i < argc; ++i) {
    std::string v;
    if (Flag(argv[i], "scene", v)) scene = v;
    else if (Flag(argv[i], "output", v)) opts.output = v;
    else if (Flag(argv[i], "acceleration", v)) {
      if (v == "bvh") opts.acceleration = 1;
      else if (v == "linear") opts.acceleration = 0;
      else { std::cerr << "unknown value for acceleration\n"; return 1; }
    } else if (Flag(argv[i], "partition_strategy", v)) {
      if (v == "sah") opts.partition_strategy = 2;
      else if (v == "midpoint") opts.partition_strategy = 1;
      else if (v == "uniform") opts.partition_strategy = 0;
      else { std::cerr << "unknown value for partition_strategy\n"; return 1; }
    } else if (Flag(argv[i], "parallelism", v)) opts.parallelism = uint32_t(std::strtoul(v.c_str(), nullptr, 10));
    else if (Flag(argv[i], "stats", v)) opts.show_stats = (v == "true" || v == "1");
    else if (Flag(argv[i], "nostats", v)) opts.show_stats = false;
    else if (Flag(argv[i], "enable_feexcept", v)) opts.enable_feexcept = (v == "true" || v == "1");
    else if (Flag(argv[i], "seed", v)) opts.default_seed = std::strtoll(v.c_str(), nullptr, 10);
    else if (Flag(argv[i], "device", v)) opts.device = std::atoi(v.c_str());
    else {
      std::cerr << "unknown flag: " << argv[i] << "\n";
      return 1;
    }
  }
  if (scene.empty()) {
    std::cerr << "A scene file is required\n";  // muon/muon.cc:36-39
    return 1;
  }
  muonx::Renderer r(scene, opts);
  return r.Render();
}
