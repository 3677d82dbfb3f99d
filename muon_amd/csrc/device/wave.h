// Wavefront-level helpers (64 lanes): queue appends with one atomic per
// wavefront and counter reductions.
#ifndef MUONX_DEVICE_WAVE_H_
#define MUONX_DEVICE_WAVE_H_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace muonx {

// Appends the lanes with `want` to a queue: returns each such lane's index
// (lane order within the wavefront).
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t *counter) {
  const uint64_t active = __ballot(1);
  const uint64_t mask = __ballot(want);
  const uint32_t lane = __lane_id();
  const uint32_t leader = uint32_t(__ffsll((unsigned long long)active) - 1);
  uint32_t base = 0;
  if (lane == leader && mask) base = atomicAdd(counter, uint32_t(__popcll(mask)));
  base = uint32_t(__shfl(int(base), int(leader)));
  return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Reserves `cnt` consecutive entries per lane with one atomic per wavefront
// (exclusive prefix sum over the lanes); all 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_append_n(uint32_t cnt, uint32_t *counter) {
  const int lane = int(__lane_id());
  uint32_t incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = uint32_t(__shfl_up(int(incl), off));
    if (lane >= off) incl += v;
  }
  const uint32_t total = uint32_t(__shfl(int(incl), 63));
  uint32_t base = 0;
  if (lane == 63 && total) base = atomicAdd(counter, total);
  base = uint32_t(__shfl(int(base), 63));
  return base + incl - cnt;
}

// Block-wide versions: NW wavefronts (blockDim.x = 64 * NW) reserve their
// entries with one atomic per block, in block order. Every thread of the
// block must call them (they synchronise the block); `sh` is NW + 1 words of
// shared memory.
template <int NW>
__device__ __forceinline__ uint32_t block_append_n(uint32_t cnt, uint32_t *counter, uint32_t *sh) {
  const int lane = int(__lane_id());
  const int w = int(threadIdx.x) >> 6;
  uint32_t incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = uint32_t(__shfl_up(int(incl), off));
    if (lane >= off) incl += v;
  }
  if (lane == 63) sh[w] = incl;  // the wavefront's total
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) tot += sh[k];
    sh[NW] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = sh[NW];
  for (int k = 0; k < w; ++k) base += sh[k];
  __syncthreads();  // sh is reused by the next call
  return base + incl - cnt;
}

template <int NW>
__device__ __forceinline__ uint32_t block_append(bool want, uint32_t *counter, uint32_t *sh) {
  const uint64_t m = __ballot(want);
  const int lane = int(__lane_id());
  const int w = int(threadIdx.x) >> 6;
  if (lane == 0) sh[w] = uint32_t(__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) tot += sh[k];
    sh[NW] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = sh[NW];
  for (int k = 0; k < w; ++k) base += sh[k];
  __syncthreads();
  return base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
}

__device__ __forceinline__ void wave_count(uint32_t n, unsigned long long *ctr) {
  unsigned long long a = n;
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off);
  if (__lane_id() == 0 && a) atomicAdd(ctr, a);
}

// Work claiming for kernels that give each lane a new item as soon as its
// current one is finished. Items are handed out in chunks of `chunk`
// consecutive indices (64..max_chunk, about four chunks per wavefront of the
// grid; larger chunks leave more work in a few wavefronts' last chunks,
// smaller ones put more atomics on the one counter): wavefront w's first chunk is [w * chunk, (w + 1) * chunk) with no
// atomic, later chunks come from a global counter (one atomic per chunk,
// offset past the first chunks of all wavefronts). A wavefront whose first
// chunk lies past the item count exits without touching the counter, so a
// launch over few or no items costs no contended atomics. cbase / cend /
// chunk are the same in every lane.
struct WaveFetch {
  static constexpr uint32_t kMaxChunk = 256;
  uint32_t chunk, first, cbase, cend;
  __device__ __forceinline__ explicit WaveFetch(uint32_t items, uint32_t max_chunk = kMaxChunk) {
    const uint32_t wpb = (blockDim.x + 63u) >> 6;
    const uint32_t waves = gridDim.x * wpb;
    const uint32_t c = (items / (4u * waves)) & ~63u;
    chunk = c < 64u ? 64u : (c > max_chunk ? max_chunk : c);
    first = waves * chunk;
    cbase = (blockIdx.x * wpb + (threadIdx.x >> 6)) * chunk;
    cend = cbase + chunk;
  }
  // Called by every lane of the wavefront with the ballot of `need`; returns
  // the index claimed for this lane (meaningful where `need` is set; an
  // index >= the item count means none is left).
  __device__ __forceinline__ uint32_t claim(uint64_t needm, uint32_t *counter) {
    const uint32_t lane = __lane_id();
    const uint32_t cnt = uint32_t(__popcll(needm));
    const uint32_t rank = uint32_t(__popcll(needm & ((1ull << lane) - 1ull)));
    const uint32_t rem = cend - cbase;
    uint32_t k;
    if (rem >= cnt) {
      k = cbase + rank;
      cbase += cnt;
    } else {
      uint32_t nb = 0;
      if (lane == 0) nb = first + atomicAdd(counter, chunk);
      nb = uint32_t(__shfl(int(nb), 0));
      k = rank < rem ? cbase + rank : nb + (rank - rem);
      cbase = nb + (cnt - rem);
      cend = nb + chunk;
    }
    return k;
  }
};

}  // namespace muonx

#endif
