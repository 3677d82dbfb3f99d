// Streaming-read probe for the benchmark's roofline: reads a device buffer
// with 16-byte loads and no writes, `passes` times per launch, so that the
// time of a launch gives the read bandwidth of the level of the memory
// hierarchy the buffer fits in (the L2 for a few MiB per XCD, the 256 MB
// MALL for tens of MiB, HBM beyond). Built into its own library
// (libmuonx_probe.so), used only by bench.py; not part of the muonx C ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_read(const uint4 *p, size_t n, int passes, uint32_t *sink) {
  const size_t stride = size_t(gridDim.x) * kBlock;
  uint32_t acc = 0;
  for (int pass = 0; pass < passes; ++pass) {
    size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
      const uint4 a = p[i], b = p[i + stride], c = p[i + 2 * stride], d = p[i + 3 * stride];
      acc ^= (a.x ^ a.y ^ a.z ^ a.w) ^ (b.x ^ b.y ^ b.z ^ b.w) ^ (c.x ^ c.y ^ c.z ^ c.w) ^ (d.x ^ d.y ^ d.z ^ d.w);
    }
    for (; i < n; i += stride) {
      const uint4 a = p[i];
      acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    // every pass loads again: the compiler may not reuse the previous
    // pass's values (an empty statement that clobbers memory)
    asm volatile("" ::: "memory");
  }
  // a data-dependent store the compiler cannot remove; the caller fills
  // the buffer so that it is never taken
  if (acc == 0x9e3779b9u) sink[blockIdx.x & 255u] = acc;
}

}  // namespace

// Launches one read of `bytes` (a multiple of 16) from `buf`, `passes`
// times over, on `stream` with `blocks` blocks of 256 threads; `sink`: 256
// words of device memory. Returns 0, or the HIP error code.
extern "C" int probe_read_launch(const void *buf, size_t bytes, int passes, int blocks, void *sink, void *stream) {
  hipLaunchKernelGGL(k_read, dim3(blocks), dim3(kBlock), 0, static_cast<hipStream_t>(stream),
                     static_cast<const uint4 *>(buf), bytes / 16, passes, static_cast<uint32_t *>(sink));
  return int(hipGetLastError());
}
