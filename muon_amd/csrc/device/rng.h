// Counter-based random numbers for the GPU render path. Every draw is a pure
// function of (job seed, pixel, sample, draw index), so results do not
// depend on scheduling, queue order or which other jobs share a launch.
// The reference's per-tile / per-thread std::mt19937 streams
// (muon/sampling.cc:41-57, muon/integration.h:123) depend on thread count
// and tile order; only their distribution (uniform [0,1)) is kept.
//
// Mixer: "lowbias32" integer hash (C. Wellons, hash-prospector, 2018).
#ifndef MUONX_DEVICE_RNG_H_
#define MUONX_DEVICE_RNG_H_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace muonx {

__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__host__ __device__ __forceinline__ uint32_t path_key(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample) {
  uint32_t k = mix32(seed ^ 0x9e3779b9u);
  k = mix32(k ^ job);
  k = mix32(k ^ pixel);
  k = mix32(k + 0x85ebca6bu * sample);
  return k;
}

// Uniform float in [0, 1) with 24 random bits.
__host__ __device__ __forceinline__ float rng_uniform(uint32_t key, uint32_t &ctr) {
  uint32_t x = mix32(key ^ mix32(ctr * 0x9e3779b9u + 0x632be5abu));
  ++ctr;
  return float(x >> 8) * (1.0f / 16777216.0f);
}

}  // namespace muonx

#endif
