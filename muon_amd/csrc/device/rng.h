// Counter-based random numbers for the GPU render path. Every draw is a pure
// function of (job seed, job key, pixel, sample, draw index), so results do
// not depend on scheduling, queue order or which other jobs share a launch.
// The reference's per-tile / per-thread std::mt19937 streams
// (muon/sampling.cc:41-57, muon/integration.h:123) depend on thread count
// and tile order; only their distribution (uniform [0,1)) is kept.
//
// Mixer: "lowbias32" integer hash (C. Wellons, hash-prospector, 2018). A
// sample's key has two 32-bit lanes from two independent hash chains over
// (seed, job key, pixel, sample), and every draw mixes in both, so two
// samples share a random stream only if both lanes collide (a 64-bit key:
// about 1e-5 expected coincident pairs among the 1.1e9 samples of a
// 4096^2 x 64 spp image, against 1.4e8 pairs with one 32-bit lane).
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

struct PathKey {
  uint32_t a, b;
};

__host__ __device__ __forceinline__ PathKey path_key(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample) {
  uint32_t k = mix32(seed ^ 0x9e3779b9u);
  k = mix32(k ^ job);
  k = mix32(k ^ pixel);
  const uint32_t a = mix32(k + 0x85ebca6bu * sample);
  // second lane: another chain over the same inputs, in another order
  uint32_t h = mix32(seed + 0x68e31da4u);
  h = mix32(h ^ (job * 0xb5297a4du));
  h = mix32(h ^ (sample + 0x1b56c4e9u));
  const uint32_t b = mix32(h + 0x27d4eb2fu * pixel);
  return PathKey{a, b};
}

// Uniform float in [0, 1) with 24 random bits.
__host__ __device__ __forceinline__ float rng_uniform(PathKey key, uint32_t &ctr) {
  const uint32_t x = mix32(mix32(key.a ^ mix32(ctr * 0x9e3779b9u + 0x632be5abu)) ^ key.b);
  ++ctr;
  return float(x >> 8) * (1.0f / 16777216.0f);
}

}  // namespace muonx

#endif
