// C ABI glue: error state, device contexts, host-scene ingest wrappers and
// film output.
#include <cmath>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>

#include "../common/vecmath.h"
#include "../host/host_scene.h"
#include "query.h"
#include "runtime.h"

namespace muonx {
namespace {
thread_local std::string g_last_error;
}
void SetError(const std::string &msg) { g_last_error = msg; }
void DestroyWorkspace(RenderWorkspace *ws);
}  // namespace muonx

using namespace muonx;

struct muonx_host_scene {
  host::HostScene sc;
};

extern "C" {

const char *muonx_last_error(void) { return g_last_error.c_str(); }

const char *muonx_version(void) { return "muonx 0.1 (gfx950)"; }

int muonx_ctx_create(int device, muonx_ctx **out) {
  if (!out) {
    SetError("muonx_ctx_create: null out");
    return MUONX_ERR_INVALID;
  }
  *out = nullptr;
  int n = 0;
  MX_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    SetError("muonx_ctx_create: device " + std::to_string(device) + " not present (" + std::to_string(n) +
             " devices)");
    return MUONX_ERR_INVALID;
  }
  MX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MX_HIP(hipGetDeviceProperties(&prop, device));
  const int rc = InitStaticScratch(device);
  if (rc != MUONX_OK) return rc;
  auto *c = new (std::nothrow) muonx_ctx();
  if (!c) return MUONX_ERR_NOMEM;
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  *out = c;
  return MUONX_OK;
}

void muonx_ctx_destroy(muonx_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  DestroyWorkspace(ctx->ws);
  if (ctx->spill_done) (void)hipEventDestroy(ctx->spill_done);
  delete ctx;
}

int muonx_ctx_set_tuning(muonx_ctx *ctx, const muonx_tuning *t) {
  if (!ctx || !t) {
    SetError("muonx_ctx_set_tuning: null argument");
    return MUONX_ERR_INVALID;
  }
  if (!(t->stack_entries == 0 || t->stack_entries == 8 || t->stack_entries == 16) || t->blocks_per_cu < 1 ||
      t->blocks_per_cu > 64 || t->path_budget_mb < 16 || !(t->bvh_width == 2 || t->bvh_width == 4) ||
      t->ray_fetch < 0 || t->ray_fetch > 64 || !(t->node_bits == 8 || t->node_bits == 32) ||
      !(t->count_work == 0 || t->count_work == 1)) {
    SetError("muonx_ctx_set_tuning: stack_entries must be 0, 8 or 16; blocks_per_cu 1..64; path_budget_mb >= 16; "
             "bvh_width 2 or 4; ray_fetch 0..64; node_bits 8 or 32; count_work 0 or 1");
    return MUONX_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->tuning = *t;
  return MUONX_OK;
}

int muonx_ctx_get_tuning(const muonx_ctx *ctx, muonx_tuning *t) {
  if (!ctx || !t) {
    SetError("muonx_ctx_get_tuning: null argument");
    return MUONX_ERR_INVALID;
  }
  *t = ctx->tuning;
  return MUONX_OK;
}

namespace {
int ParseInto(const std::string &text, const std::string &base_dir, const muonx_options *opts,
              muonx_host_scene **out) {
  muonx_options o{1, 2, 1, 1, -1, 0};
  if (opts) o = *opts;
  auto *hs = new (std::nothrow) muonx_host_scene();
  if (!hs) return MUONX_ERR_NOMEM;
  try {
    host::ParseScene(text, o.acceleration, o.partition_strategy, o.default_seed, hs->sc, base_dir,
                     o.device_bvh == 0);
  } catch (const std::exception &e) {
    delete hs;
    SetError(std::string("parse failed: ") + e.what());
    return MUONX_ERR_INVALID;
  }
  *out = hs;
  return MUONX_OK;
}
}  // namespace

int muonx_host_scene_parse_text(const char *text, size_t len, const muonx_options *opts, muonx_host_scene **out) {
  if (!text || !out) {
    SetError("muonx_host_scene_parse_text: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = nullptr;
  return ParseInto(std::string(text, len), std::string(), opts, out);
}

int muonx_host_scene_parse_file(const char *path, const muonx_options *opts, muonx_host_scene **out) {
  if (!path || !out) {
    SetError("muonx_host_scene_parse_file: null argument");
    return MUONX_ERR_INVALID;
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    SetError(std::string("cannot open scene file: ") + path);
    return MUONX_ERR_IO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  *out = nullptr;
  // `load` paths are relative to the scene file (muon/parser.cc:474-477)
  const std::string p(path);
  const size_t slash = p.find_last_of('/');
  const std::string dir = slash == std::string::npos ? std::string(".") : slash == 0 ? std::string("/") : p.substr(0, slash);
  return ParseInto(ss.str(), dir, opts, out);
}

void muonx_host_scene_destroy(muonx_host_scene *hs) { delete hs; }

int muonx_host_scene_desc(const muonx_host_scene *hs, muonx_scene_desc *out) {
  if (!hs || !out) {
    SetError("muonx_host_scene_desc: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = hs->sc.Desc();
  return MUONX_OK;
}

int muonx_host_scene_params(const muonx_host_scene *hs, muonx_render_params *out) {
  if (!hs || !out) {
    SetError("muonx_host_scene_params: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = hs->sc.Params();
  return MUONX_OK;
}

int muonx_host_scene_warnings(const muonx_host_scene *hs) { return hs ? hs->sc.warnings : -1; }

int muonx_host_scene_output(const muonx_host_scene *hs, const char **out) {
  if (!hs || !out) {
    SetError("muonx_host_scene_output: null argument");
    return MUONX_ERR_INVALID;
  }
  *out = hs->sc.output.c_str();
  return MUONX_OK;
}

// muon/film.cc:21-43
int muonx_film_to_rgb8(const float *accum, int32_t w, int32_t h, int32_t spp, float gamma, uint8_t *rgb8) {
  if (!accum || !rgb8 || w <= 0 || h <= 0 || spp <= 0) {
    SetError("muonx_film_to_rgb8: bad argument");
    return MUONX_ERR_INVALID;
  }
  const size_t n = size_t(w) * h * 3;
  for (size_t i = 0; i < n; ++i) {
    float v = accum[i] / static_cast<float>(spp);
    v = std::pow(v, 1.0f / gamma);
    v = gclamp(v, 0.0f, 1.0f);
    // NaN survives the clamp; x86-64 float->int conversion turns it into 0.
    rgb8[i] = (v == v) ? static_cast<uint8_t>(v * 255) : uint8_t(0);
  }
  return MUONX_OK;
}

}  // extern "C"
