// Index arithmetic of the persistent engine, shared by its kernels and its
// host code: constexpr functions of plain integers (device code calls them
// too) and the byte layout of the per-path buffers. No HIP types, so the
// host C++ compiler can check them on the CPU.
#ifndef MUONX_DEVICE_ENGINE_PLAN_H_
#define MUONX_DEVICE_ENGINE_PLAN_H_

#include <cstddef>
#include <cstdint>

namespace muonx {

// Samples the plan step appends behind the `next` continuing paths of the
// next queue: all admitted work not generated yet (total_items - cursor),
// up to the queue's free room, so the queue never holds more than
// `capacity` paths.
constexpr uint32_t FillCount(unsigned long long total_items, unsigned long long cursor, uint32_t capacity,
                             uint32_t next) {
  const unsigned long long avail = total_items > cursor ? total_items - cursor : 0ull;
  const uint32_t room = capacity > next ? capacity - next : 0u;
  return uint32_t(avail < room ? avail : room);
}

// A path's job word: its job slot in the low 16 bits and the job's scene
// slot in the high 16, so that the kernels fetch the job record and the
// scene view in parallel instead of reading the scene slot from the job
// record first. The engine keeps max_jobs and the scene slots below 2^16.
constexpr uint32_t kMaxPathSlotIndex = 1u << 16;
constexpr uint32_t PackPathJob(uint32_t job_slot, uint32_t scene_slot) { return job_slot | (scene_slot << 16); }
constexpr uint32_t PathJobSlot(uint32_t word) { return word & 0xffffu; }
constexpr uint32_t PathSceneSlot(uint32_t word) { return word >> 16; }

// Sequence number of the admitted job whose work holds global item `item`:
// the last of the jobs [lo, hi) whose first item is <= item. First items
// grow with the sequence number; job_first is a ring indexed by
// seq % max_jobs, valid for the admitted window (hi - lo <= max_jobs).
// Requires lo < hi and job_first[lo % max_jobs] <= item.
constexpr unsigned long long FindJobSeq(const unsigned long long *job_first, uint32_t max_jobs,
                                        unsigned long long lo, unsigned long long hi, unsigned long long item) {
  while (hi - lo > 1) {
    const unsigned long long m = lo + (hi - lo) / 2;
    if (job_first[m % max_jobs] <= item) lo = m;
    else hi = m;
  }
  return lo;
}

// Where a job-relative sample item lies: pixel = local % npix, sample =
// sample_base + local / npix (a slice covers every pixel for consecutive
// samples).
struct ItemPos {
  uint32_t pixel, sample;
};
constexpr ItemPos ItemPosition(unsigned long long local, uint32_t npix, uint32_t sample_base) {
  return ItemPos{uint32_t(local % npix), sample_base + uint32_t(local / npix)};
}

// Index of shadow ray k of path i (the shadow arrays hold kmax rays per path
// in k-major order). Engine::Create keeps capacity * kmax <= 2^31, so the
// index and the queue entries that store it fit 32 bits.
constexpr uint32_t ShadowSlot(uint32_t k, uint32_t capacity, uint32_t i) { return k * capacity + i; }
constexpr unsigned long long kMaxShadowSlots = 1ull << 31;

// Byte layout of the engine's buffers: arrays of `per_path` elements of
// `elem_bytes` per path (per path and shadow ray for the shadow buffer), in
// carve order, each starting on a 256-byte boundary.
struct ArraySpec {
  uint32_t elem_bytes, per_path;
};
constexpr size_t kCarveAlign = 256;
constexpr size_t CarveSize(size_t count, size_t elem_bytes) {
  return (count * elem_bytes + kCarveAlign - 1) & ~(kCarveAlign - 1);
}

// Path buffer: two path states (job, pixel, RNG counter, depth, radiance x3,
// throughput x3, the two words of the sample's random-stream key), the two
// ray queues and the MIS rays (two float4 per path each), the hit primitive,
// the next-queue index, the MIS factor x3, the MIS hit, the shadow / MIS
// origin and the hit normal (one float4 each).
constexpr ArraySpec kPathArrays[] = {
    {4, 1}, {4, 1}, {4, 1}, {4, 1}, {4, 3}, {4, 3}, {4, 1}, {4, 1},  // state 0
    {4, 1}, {4, 1}, {4, 1}, {4, 1}, {4, 3}, {4, 3}, {4, 1}, {4, 1},  // state 1
    {16, 2}, {16, 2}, {16, 2},                                // ray queues 0, 1; MIS rays
    {4, 1}, {4, 1}, {4, 3}, {4, 1},                           // hit prim, next index, MIS factor, MIS hit
    {16, 1}, {16, 1}};                                        // shadow origin + material, hit normal
constexpr int kNumPathArrays = int(sizeof(kPathArrays) / sizeof(kPathArrays[0]));
// Shadow buffer, per shadow ray: direction + tmax (float4), contribution
// x3, visible and valid flags, the trace-queue entry.
constexpr ArraySpec kShadowArrays[] = {{16, 1}, {4, 3}, {1, 1}, {1, 1}, {4, 1}};
constexpr int kNumShadowArrays = int(sizeof(kShadowArrays) / sizeof(kShadowArrays[0]));

// Offsets of the arrays for `n` rows (paths, or paths x kmax) into `off`;
// returns the buffer's size.
inline size_t LayoutOffsets(const ArraySpec *spec, int count, size_t n, size_t *off) {
  size_t p = 0;
  for (int k = 0; k < count; ++k) {
    off[k] = p;
    p += CarveSize(n * spec[k].per_path, spec[k].elem_bytes);
  }
  return p;
}

// Device bytes of one path in flight with `kmax` shadow rays per vertex,
// without the 256-byte rounding of each array: what the path budget
// (tuning.path_budget_mb) is divided by.
constexpr uint64_t PathSlotBytes(int kmax) {
  uint64_t path = 0, shadow = 0;
  for (const ArraySpec &a : kPathArrays) path += uint64_t(a.elem_bytes) * a.per_path;
  for (const ArraySpec &a : kShadowArrays) shadow += uint64_t(a.elem_bytes) * a.per_path;
  return path + shadow * uint64_t(kmax);
}

// The job slots of jobs [lo, hi) in the ring of max_jobs slots (slot = seq
// % max_jobs, hi - lo <= max_jobs) as at most two contiguous runs, in job
// order: what the admission flush copies to the device. Returns the number
// of runs.
struct SlotRun {
  uint32_t first, count;
};
constexpr int SlotRuns(unsigned long long lo, unsigned long long hi, uint32_t max_jobs, SlotRun *out) {
  int n = 0;
  for (unsigned long long q = lo; q < hi && n < 2;) {
    const uint32_t s0 = uint32_t(q % max_jobs);
    const unsigned long long left = hi - q, room = max_jobs - s0;
    const uint32_t run = uint32_t(left < room ? left : room);
    out[n++] = SlotRun{s0, run};
    q += run;
  }
  return n;
}

}  // namespace muonx

#endif
