#!/usr/bin/env python
"""Benchmark: incoherent closest-hit rays on the displaced level-7 icosphere.

Workload (one step): 2^24 rays, origins uniform in the mesh's world bounds,
directions uniform on the sphere, traced for the closest hit against the
icosphere (327,680 tris) inside the Cornell enclosure (36 more tris). Rays
and the scene are resident in HBM before the timed region. value = rays /
second over K timed launches (HIP events on the launch stream).

Roofline (no dense contraction; memory-bound family): `achieved` is the
device tree's algorithmic bytes per ray,
  24 (ray in) + 8 (hit out) + 64 * N_node + 48 * N_tri,
with N_node / N_tri the quantized 4-wide nodes fetched and triangles tested
per ray (the traversal's own counters on the same rays), times the rays of
one launch over the mean launch time, against the 8 TB/s HBM peak (frac).
The scene (58 MB) is resident in the 256 MB MALL, so the same bytes are also
set against the MALL read rate measured in the run by a streaming-read
probe (mall_frac; read_measured_gbs holds its L2, MALL and HBM rates).
`traffic` is the kernel's measured traffic past L2 per launch (rocprofv3
FETCH_SIZE / WRITE_SIZE scaled by a copy of known size, from
profiles/traffic_closest_icosphere.json). reference_equivalent gives the
bytes the reference's binary SAH tree would move for the same rays,
32 + 16 + 32 per bounds test + 36 per object test (its counters measured by
the CPU restatement on a sample of the rays).

CPU baseline: the CPU restatement of the reference (same BVH, same
per-primitive object-space tests) on all host cores: a bounded sample of the
same rays, and renders of the Cornell box at full size, the sphere field at
512x512 and the icosphere at 128x128 (same spp, estimators and depth as the
GPU renders), each with its GPU/CPU ratio.

Also reported (outside the timed region), after muonx_autotune has chosen the
launch configuration for this scene and render (the microbench above runs
with the library's default configuration):
* render: one full path-traced render of the same scene (512x512, 64 spp,
  depth 8, BRDF IS + NEE + MIS); render.sphere_field: the 10,000-sphere field
  with 8 emissive spheres (1024x1024, 32 spp, depth 8, BRDF sampling, NEE off);
  render.cornell_256: the Cornell box at 256x256, 16 spp, depth 5 (the
  reference's CPU-runnable case).
* admission: the level-8 icosphere (1.31M tris) made resident with the
  host's SAH build and with the BVH built on the device, and the
  closest-hit rate of each tree.
* batched: 64 jobs, each on its own scene (a displaced icosphere or a
  sphere field, primitives log-uniform in [1k, 2M], about 17.6M in all,
  BVHs built on the device), resolution uniform in {128^2 .. 1024^2}, spp
  uniform in [4, 64], depth 8, seeded poses; rendered as one batch on the
  persistent engine and as 64 sequential plain muonx_render calls: images/s
  of both, whether every image is bit-identical between the two (so the
  per-image relMSE is equal), and the relMSE of up to 6 of the images
  against converged renders (16,384 spp). The same for 128 small jobs
  (128^2, spp 4..16) as batched.small_jobs; batched.converged_cross_check
  sets the GPU's converged reference against the CPU restatement: the mean
  relMSE of 16 oracle images and 16 GPU images at 256 spp against it.
* serving: the request loop under a Poisson stream of requests (resolution
  {256^2, 512^2, 1024^2}, spp 8..128, random scene and pose) over its own 8
  resident scenes (icospheres of 20k..1.3M tris, four of them at 1.3M, and
  a sphere field: about 5.7M primitives) at a fixed rate chosen as a load
  fraction of the batched throughput of the same request mix on those
  scenes (192 requests by default), replayed under each slice order
  (oldest first, least work first, highest response ratio); images/s, p50/p99
  latency (submit to film written), the relMSE of the four smallest
  requests against converged renders, and each request rendered again
  alone: its slowdown (served latency / alone latency, p50 and p99) and
  whether every served film equals the alone one bit for bit.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s (incoherent, 1 GPU) + p50/p99 render latency at fixed request rate"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_pool(ctx, log_fn, kinds=None):
    from muon_amd import scenes, workloads
    return workloads.build_pool(ctx, kinds or scenes.POOL, lambda m: log_fn("bench: " + m))


def bench_batched(ctx, pool, njobs, seed, log_fn, res_choices=(128, 256, 512, 1024), spp_range=(4, 64),
                  sequential=True, jobs=None):
    import numpy as np
    import torch
    from muon_amd.renderer import render_device
    from muon_amd.serving import engine_config, render_batch
    from muon_amd.workloads import batched_jobs
    # key 0 for every job: the plain path renders with key 0 too, so both
    # paths must produce the same images (the jobs differ by seed and pose)
    if jobs is None:
        jobs = batched_jobs(pool, njobs, seed, res_choices, spp_range)
    njobs = len(jobs)
    samples = sum(p.width * p.height * p.spp for _, p, _ in jobs)
    film = sum(p.width * p.height * 24 + 256 for _, p, _ in jobs)
    cfg = engine_config(max_jobs=max(64, 2 * njobs), film_budget_bytes=film + (1 << 20), iterations_per_poll=4)
    log_fn(f"bench: batched {njobs} jobs, {samples / 1e6:.0f} M samples")
    render_batch(ctx, jobs[:2], cfg)  # engine + graphs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    accs, results, st = render_batch(ctx, jobs, cfg)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    rays = st.primary_rays + st.secondary_rays + st.shadow_rays
    if not sequential:
        return None, samples / tb
    log_fn(f"bench: batched {tb:.2f} s; sequential plain renders")
    render_device(jobs[0][0], jobs[0][1])  # the plain path's engine
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq = []
    for ds, p, _ in jobs:
        a, _ = render_device(ds, p)
        seq.append(a)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    same = sum(int(torch.equal(a, b)) for a, b in zip(accs, seq))
    # per-image relMSE against converged renders, for up to 6 of the jobs at
    # 128^2 or 256^2 (equal for the batch and the plain renders: the images
    # are identical)
    small = [i for i, (_, p, _) in enumerate(jobs) if p.width * p.height <= 256 * 256][:6]
    quality = relmse_vs_converged(ctx, [(jobs[i][0], jobs[i][1], accs[i]) for i in small], 5000)
    lat = [r.latency_s for r in results]
    return {"jobs": njobs, "samples": samples, "batched_s": round(tb, 3), "sequential_s": round(ts, 3),
            "images_per_s": round(njobs / tb, 3), "sequential_images_per_s": round(njobs / ts, 3),
            "speedup": round(ts / tb, 3), "mrays_per_s": round(rays / tb / 1e6, 1),
            "msamples_per_s": round(samples / tb / 1e6, 1), "identical_to_plain": f"{same}/{njobs}",
            "job_latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "job_latency_p99_s": round(float(np.percentile(lat, 99)), 3),
            "relmse_vs_converged": quality}, samples / tb


def host_threads():
    """Host cores this process may run on (all of them for the CPU baseline)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_description():
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{model}, nproc {os.cpu_count()}, {host_threads()} usable"


def cpu_render_cases():
    """(key, scene text, sample) of the CPU restatement's render baselines:
    the reference's CPU-runnable case at full size; the sphere field (10,008
    spheres, BRDF sampling, NEE off, depth 8) at 512x512 instead of
    1024x1024; the level-7 icosphere (GGX, NEE + MIS, depth 8) at 128x128
    instead of 512x512 (the same scenes, spp and estimators on fewer
    pixels)."""
    from muon_amd import scenes
    return [("cornell_256_render", scenes.cornell_text(256, 256, 16, 5, seed=0),
             "Cornell box 256x256, 16 spp, depth 5 (full size)"),
            ("sphere_field_render", scenes.sphere_field_text(512, 512, 32, 8),
             "sphere field 512x512 of the 1024x1024, 32 spp, depth 8"),
            ("icosphere_render", scenes.icosphere_text(128, 128, 64, 8),
             "icosphere 128x128 of the 512x512, 64 spp, depth 8")]


def read_rates_gbs(dev, num_cus):
    """Measured read bandwidth of a streaming read kernel
    (muon_amd/csrc/bench/read_probe.hip, 16-byte loads) over buffers resident
    in the L2 (16 MiB: 2 MiB per XCD), the MALL (64 MiB) and HBM (8 GiB):
    {level: GB/s}, the best of 8, 16 and 32 blocks of 256 threads per CU and
    of 5 timed launches after 3 untimed ones. None if the probe library is
    missing."""
    import ctypes
    import torch
    path = os.path.join(ROOT, "muon_amd", "lib", "libmuonx_probe.so")
    if not os.path.exists(path):
        return None
    probe = ctypes.CDLL(path)
    probe.probe_read_launch.restype = ctypes.c_int
    probe.probe_read_launch.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p]
    stream = torch.cuda.current_stream(dev)
    sink = torch.zeros(256, dtype=torch.int32, device=dev)
    out = {}
    for level, nbytes, passes in (("l2", 16 << 20, 64), ("mall", 64 << 20, 8), ("hbm", 8 << 30, 1)):
        buf = torch.zeros(nbytes // 4, dtype=torch.int32, device=dev)
        best = None
        for bpc in (8, 16, 32):
            def launch():
                return probe.probe_read_launch(buf.data_ptr(), nbytes, passes, bpc * num_cus, sink.data_ptr(),
                                               stream.cuda_stream)
            for _ in range(3):
                assert launch() == 0
            for _ in range(5):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                assert launch() == 0
                b.record(stream)
                torch.cuda.synchronize()
                ms = a.elapsed_time(b)
                best = ms if best is None else min(best, ms)
        out[level] = round(nbytes * passes / (best / 1e3) / 1e9, 1)
        del buf
    return out


def roofline_of(rays, mean_ms, dev_nodes, dev_prims, ref_nodes, ref_objs, scene_bytes, dev):
    """Roofline of the closest-hit kernel.

    achieved: the device tree's algorithmic bytes per ray, 24 (origin,
    direction) + 8 (prim id, distance) + 64 per quantized 4-wide node fetched
    + 48 per triangle test, times the rays of one launch, over the mean
    launch time; frac against the 8 TB/s HBM peak. The scene is resident in
    the 256 MB MALL, so the same bytes are also set against the measured
    read rate of a MALL-resident buffer (mall_frac; read_measured_gbs holds
    the probe's L2, MALL and HBM rates) and the kernel's
    measured traffic past L2 (rocprofv3 FETCH_SIZE / WRITE_SIZE, from
    profiles/traffic_closest_icosphere.json) is reported as `traffic`.
    reference_equivalent: the bytes the reference's binary SAH tree would move
    for the same rays, 32 + 16 + 32 per bounds test + 36 per object test
    (its counters measured by the oracle on a sample of the rays)."""
    b_dev = 24 + 8 + 64 * dev_nodes + 48 * dev_prims
    achieved = b_dev * rays / (mean_ms / 1e3) / 1e9
    b_ref = 32 + 16 + 32 * ref_nodes + 36 * ref_objs
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_closest_icosphere.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    import torch
    rates = read_rates_gbs(dev, torch.cuda.get_device_properties(dev).multi_processor_count) or {}
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "bytes_per_ray": round(b_dev, 1), "device_nodes_per_ray": round(dev_nodes, 2),
           "device_prim_tests_per_ray": round(dev_prims, 2), "scene_bytes": int(scene_bytes),
           "read_measured_gbs": rates or None,
           "mall_frac": round(achieved / rates["mall"], 4) if rates else None,
           "reference_equivalent": {"bytes_per_ray": round(b_ref, 1), "gbs": round(b_ref * rays / (mean_ms / 1e3) / 1e9, 1),
                                    "bounds_tests_per_ray": round(ref_nodes, 2),
                                    "object_tests_per_ray": round(ref_objs, 2)}}
    if traffic:
        out["traffic_gbs"] = round(traffic / (mean_ms / 1e3) / 1e9, 1)
        out["traffic_frac"] = round(out["traffic_gbs"] / HBM_PEAK_GBS, 4)
    return out


CONVERGED_SPP = 16384


def relmse_vs_converged(ctx, items, key_base):
    """relMSE (mean over pixels and channels of
    (x - r)^2 / (r^2 + 1e-2) on the linear film / spp) of each rendered image
    against a converged render of the same job at 16,384 spp with another
    random stream. items: (DeviceScene, RenderParams, accum tensor). Returns
    a list of {"res", "spp", "relmse", "spp_for_1e-3"} (the spp at which the
    relMSE, falling as 1/spp, would reach 1e-3)."""
    import numpy as np
    from muon_amd import _lib
    from muon_amd.serving import render_batch
    out = []
    for k, (ds, p, acc) in enumerate(items):
        q = _lib.RenderParams.from_buffer_copy(p)
        q.spp = CONVERGED_SPP
        (ref,), _, _ = render_batch(ctx, [(ds, q, key_base + k)])
        rel = relmse(acc.cpu().numpy() / p.spp, ref.cpu().numpy() / CONVERGED_SPP)
        out.append({"res": f"{p.width}x{p.height}", "spp": int(p.spp), "relmse": round(rel, 6),
                    "spp_for_1e-3": int(np.ceil(p.spp * rel / 1e-3))})
        del ref
    return out


def converged_cross_check(ctx, orc, threads, log_fn, seeds=16, spp=256):
    """The converged GPU reference against the oracle: the Cornell box at
    64x64 (BRDF sampling + NEE + MIS, depth 5) rendered at 16,384 spp on the
    GPU, then `seeds` images at `spp` by the oracle (scene seeds 1..seeds)
    and by the GPU (render keys 1..seeds); the mean relMSE of the two sets
    against the GPU reference must agree (the oracle's estimator and random
    streams against the GPU's). One image's relMSE varies by about 8% from
    seed to seed, so the mean over 16 has a spread of about 3% per side."""
    import numpy as np
    from muon_amd import DeviceScene, HostScene, scenes
    from muon_amd.serving import render_batch
    from muon_amd import _lib
    text = scenes.cornell_text(64, 64, spp, 5, seed=0)
    hs = HostScene.from_text(text)
    ds = DeviceScene(ctx, hs)
    q = _lib.RenderParams.from_buffer_copy(hs.params)
    q.spp = CONVERGED_SPP
    (ref,), _, _ = render_batch(ctx, [(ds, q, 77)])
    r = ref.cpu().numpy() / CONVERGED_SPP
    accs, _, _ = render_batch(ctx, [(ds, hs.params, 100 + k) for k in range(seeds)])
    g = [relmse(a.cpu().numpy() / spp, r) for a in accs]
    o, secs = [], 0.0
    for k in range(seeds):
        acc, _, _, t = orc.load_text(scenes.cornell_text(64, 64, spp, 5, seed=1 + k)).render(threads)
        o.append(relmse(acc / spp, r))
        secs += float(t)
    out = {"scene": "Cornell box 64x64, depth 5", "reference_spp": CONVERGED_SPP, "spp": spp, "images": seeds,
           "gpu_relmse_mean": round(float(np.mean(g)), 6), "oracle_relmse_mean": round(float(np.mean(o)), 6),
           "gpu_relmse_sd": round(float(np.std(g)), 6), "oracle_relmse_sd": round(float(np.std(o)), 6),
           "oracle_seconds": round(secs, 2)}
    out["ratio"] = round(out["gpu_relmse_mean"] / out["oracle_relmse_mean"], 3)
    log_fn(f"bench: converged-reference cross-check {out}")
    return out


def bench_admission(ctx, log_fn):
    """Time to make a scene resident: the displaced level-8 icosphere
    (1,310,732 primitives) ingested with the host's SAH build (the
    reference's tree, as .muon text) and loaded from binary PLY with its BVH
    built on the device; the closest-hit rate of the two trees on 2^22
    incoherent rays."""
    import tempfile
    import torch
    from muon_amd import DeviceScene, HostScene, scenes
    out = {"scene": "displaced icosphere, level 8"}
    text = scenes.icosphere_text(256, 256, 16, 8, level=8, seed=0)
    t0 = time.perf_counter()
    hs = HostScene.from_text(text)
    t1 = time.perf_counter()
    ds = DeviceScene(ctx, hs)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["host_sah"] = {"parse_and_build_s": round(t1 - t0, 3), "upload_s": round(t2 - t1, 3),
                       "total_s": round(t2 - t0, 3), "prims": ds.info()["num_prims"]}
    with tempfile.TemporaryDirectory() as tmp:
        ptext = scenes.icosphere_text(256, 256, 16, 8, level=8, seed=0, mesh_file=f"{tmp}/ico8.ply")
        t0 = time.perf_counter()
        hd = HostScene.from_text(ptext, device_bvh=True)
        t1 = time.perf_counter()
        dd = DeviceScene(ctx, hd)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    info = dd.info()
    out["device_bvh"] = {"parse_s": round(t1 - t0, 3), "build_and_upload_s": round(t2 - t1, 3),
                         "build_ms": round(info["build_ms"], 1), "total_s": round(t2 - t0, 3),
                         "prims": info["num_prims"]}
    lo, hi = scenes.icosphere_world_bounds(level=8)
    o_np, d_np = scenes.microbench_rays(1 << 22, lo, hi, seed=1)
    dev = torch.device("cuda:0")
    o, d = torch.from_numpy(o_np).to(dev), torch.from_numpy(d_np).to(dev)
    for key, scene in (("host_sah", ds), ("device_bvh", dd)):
        for _ in range(2):
            scene.intersect_closest(o, d)
        stream = torch.cuda.current_stream(dev)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(5):
            scene.intersect_closest(o, d)
        b.record(stream)
        torch.cuda.synchronize()
        out[key]["closest_hit_mrays_per_s"] = round(5 * (1 << 22) / (a.elapsed_time(b) / 1e3) / 1e6, 1)
    return out


def relmse(x, r):
    import numpy as np
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def serve_stream(ctx, reqs, gaps, order):
    """Replays the request stream (request i submitted gaps[0..i] after the
    start, key i) on a new server with the given slice order. Returns the
    results, the films and the rays traced."""
    from muon_amd.serving import Server, engine_config
    cfg = engine_config(max_jobs=256, film_budget_bytes=1 << 30, iterations_per_poll=2, slice_order=order)
    outs = {}
    with Server(ctx, cfg) as srv:
        # warm the engine (graph capture, scene table) outside the measurement
        ds0, p0 = reqs[0]
        srv.wait(srv.submit(ds0, p0, 0), timeout_ms=-1)
        warm = srv.stats()
        t_next = time.perf_counter()
        tickets = []
        for i, (ds, p) in enumerate(reqs):
            t_next += gaps[i]
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(ds, p, i))
        results = []
        for i, t in enumerate(tickets):
            acc, r = srv.wait(t, timeout_ms=-1)
            results.append(r)
            outs[i] = acc
        stats = srv.stats()
    rays = sum(stats[k] - warm[k] for k in ("primary_rays", "secondary_rays", "shadow_rays"))
    return results, outs, rays


def bench_serving(ctx, pool, nreq, capacity_samples_s, load, seed, log_fn):
    """The served workload under each slice order (the library's default at
    the top level, the others under other_orders), each request then
    rendered alone for its slowdown and a bit-for-bit comparison with every
    served film."""
    import numpy as np
    import torch
    from muon_amd.serving import (SLICES_LEAST_WORK_FIRST, SLICES_OLDEST_FIRST, SLICES_RESPONSE_RATIO, Server,
                                  engine_config)
    from muon_amd.workloads import poisson_gaps, served_requests
    reqs = served_requests(pool, nreq, seed)
    mean_samples = float(np.mean([p.width * p.height * p.spp for _, p in reqs]))
    rate = load * capacity_samples_s / mean_samples
    gaps = poisson_gaps(nreq, rate, seed)
    log_fn(f"bench: serving {nreq} requests at {rate:.2f}/s (load {load})")
    runs = {}
    orders = (("oldest_first", SLICES_OLDEST_FIRST), ("least_work_first", SLICES_LEAST_WORK_FIRST),
              ("response_ratio", SLICES_RESPONSE_RATIO))
    default = next(name for name, order in orders if order == engine_config().slice_order)
    for name, order in orders:
        runs[name] = serve_stream(ctx, reqs, gaps, order)
        log_fn(f"bench: serving stream replayed, {name}")
    # each request again, alone on a server with the same key: its latency
    # without contention, and its film, which must equal both served ones
    # bit for bit
    alone, same = [], {name: 0 for name in runs}
    with Server(ctx, engine_config(max_jobs=256, film_budget_bytes=1 << 30, iterations_per_poll=2)) as srv:
        srv.wait(srv.submit(reqs[0][0], reqs[0][1], 0), timeout_ms=-1)
        for i, (ds, p) in enumerate(reqs):
            acc, r = srv.wait(srv.submit(ds, p, i), timeout_ms=-1)
            alone.append(r.latency_s)
            for name, (_, outs, _) in runs.items():
                same[name] += int(torch.equal(acc, outs[i]))
            del acc
    alone = np.array(alone)

    def summary(name):
        results, _, rays = runs[name]
        makespan = (max(r.done_ns for r in results) - min(r.submit_ns for r in results)) * 1e-9
        lat = np.array([r.latency_s for r in results])
        slow = lat / np.maximum(alone, 1e-9)
        return {"images_per_s": round(nreq / makespan, 3), "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 3), "mean_latency_s": round(float(lat.mean()), 3),
                "mean_queue_s": round(float(np.mean([r.queue_s for r in results])), 3),
                "slowdown_p50": round(float(np.percentile(slow, 50)), 2),
                "slowdown_p99": round(float(np.percentile(slow, 99)), 2),
                "identical_to_alone": f"{same[name]}/{nreq}", "mrays_per_s": round(rays / makespan / 1e6, 1)}

    # relMSE of the four smallest requests against converged renders
    outs = runs[default][1]
    small = sorted(range(nreq), key=lambda i: reqs[i][1].width * reqs[i][1].height * reqs[i][1].spp)[:4]
    rel = relmse_vs_converged(ctx, [(reqs[i][0], reqs[i][1], outs[i]) for i in small], 9000)
    out = {"requests": nreq, "rate_per_s": round(rate, 3), "load": load, "slice_order": default}
    out.update(summary(default))
    out.update({"alone_latency_p50_s": round(float(np.percentile(alone, 50)), 3),
                "alone_latency_max_s": round(float(alone.max()), 3),
                "other_orders": {name: summary(name) for name, _ in orders if name != default},
                "relmse_vs_converged": rel})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--no-sphere-field", action="store_true", help="render only the icosphere scene")
    ap.add_argument("--only-sphere-field", action="store_true",
                    help="of the renders, only the sphere field (for a profile of that workload)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and serving workloads")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=192)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 256 MiB device copy first (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from muon_amd import Context, DeviceScene, HostScene, scenes
    from muon_amd.renderer import render_device

    dev = torch.device("cuda:0")
    log("bench: building the level-7 icosphere scene")
    text = scenes.icosphere_text()
    hs = HostScene.from_text(text)
    ctx = Context(0)
    ds = DeviceScene(ctx, hs)
    info = ds.info()
    lo, hi = scenes.icosphere_world_bounds()
    o_np, d_np = scenes.microbench_rays(args.rays, lo, hi, seed=0)
    o = torch.from_numpy(o_np).to(dev)
    d = torch.from_numpy(d_np).to(dev)
    prim = torch.empty(args.rays, dtype=torch.int32, device=dev)
    t = torch.empty(args.rays, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log(f"bench: warmup {args.warmup}")
    for _ in range(args.warmup):
        ds.intersect_closest(o, d, None, prim, t)
    torch.cuda.synchronize()
    log(f"bench: timing {args.steps} launches of {args.rays} rays")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        ds.intersect_closest(o, d, None, prim, t)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    mean_ms = float(np.mean(launch_ms))
    total_s = sum(launch_ms) / 1e3
    mrays = args.rays * args.steps / total_s / 1e6
    hits = int((prim >= 0).sum().item())

    # device traversal counters for one launch
    ctr = torch.zeros(2, dtype=torch.int64, device=dev)
    ds.query(o, d, counters=ctr, prim=prim, t=t)
    torch.cuda.synchronize()
    dev_nodes, dev_prims = [int(x) for x in ctr.cpu().tolist()]

    # reference-BVH counters and CPU baseline from the CPU restatement
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    orc = _oracle.Oracle()
    oscene = orc.load_text(text)
    cpu_threads = host_threads()
    sample = 1 << 16
    _, _, _, _, st = oscene.intersect(o_np[:sample], d_np[:sample], threads=cpu_threads)
    n_node = st[4] / sample
    n_obj = st[2] / sample
    roofline = roofline_of(args.rays, mean_ms, dev_nodes / args.rays, dev_prims / args.rays, n_node, n_obj,
                           info["device_bytes"], dev)

    # the autotuner picks the launch configuration for this scene and the C1
    # render job; the render, batched and serving measurements below use it
    tune = None
    if not args.no_autotune:
        from muon_amd.scene import autotune
        log("bench: autotune")
        rep = autotune(ds, hs.params, 1 << 22)
        tune = {"tuning": rep["tuning"], "candidates": rep["candidates"], "best_ms": round(rep["best_ms"], 4),
                "default_ms": round(rep["previous_ms"], 4), "worst_ms": round(rep["worst_ms"], 4),
                "sample_rays": rep["sample_rays"]}
        log(f"bench: autotune {tune}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads
        done, secs, chunk = 0, 0.0, 1 << 18
        while secs < args.cpu_seconds and done < args.rays:
            n = min(chunk, args.rays - done)
            t0 = time.perf_counter()
            oscene.intersect(o_np[done:done + n], d_np[done:done + n], threads=threads)
            secs += time.perf_counter() - t0
            done += n
        cpu = {"value": round(done / secs / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"first {done} of the same microbench rays, closest hit, {threads} threads, {secs:.1f} s",
               "host": host_description()}
        log(f"bench: cpu baseline {cpu['value']} Mrays/s on {threads} threads ({cpu['host']})")
        if not args.no_render:
            # the CPU restatement rendering the configurations (render time as
            # the reference's Stats measures it: tiles only, no parse / build):
            # the reference's own CPU case at full size, the sphere field and
            # the icosphere at reduced films (bounded samples of the same
            # scenes, estimator and depth)
            for key, scene_text, label in cpu_render_cases():
                c = orc.load_text(scene_text)
                _, _, cst, csecs = c.render(threads)
                rays = int(cst[0] + cst[1] + cst[6])
                cpu[key] = {"seconds": round(float(csecs), 3), "threads": threads, "sample": label,
                            "reference_rays": int(cst[0] + cst[1]), "rays": rays,
                            "mrays_per_s": round(rays / csecs / 1e6, 3)}
                log(f"bench: cpu {key}: {label}: {csecs:.2f} s, {cpu[key]['mrays_per_s']} Mrays/s")
                del c

    def timed_render(scene, params):
        accum, rst = render_device(scene, params)  # warm the workspace
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        accum, rst = render_device(scene, params)
        torch.cuda.synchronize()
        rs = time.perf_counter() - t0
        rays_total = rst.primary_rays + rst.secondary_rays + rst.shadow_rays
        traced = rays_total - rst.culled_rays
        return {"seconds": round(rs, 4), "images_per_s": round(1.0 / rs, 3),
                "mrays_per_s": round(rays_total / rs / 1e6, 1), "primary_rays": rst.primary_rays,
                "secondary_rays": rst.secondary_rays, "shadow_rays": rst.shadow_rays,
                # MIS rays whose direction misses every emitter are settled
                # without a traversal; they stay in secondary_rays (the
                # reference traces them) and are excluded here
                "culled_mis_rays": rst.culled_rays, "traced_mrays_per_s": round(traced / rs / 1e6, 1),
                "mean_radiance": float(accum.mean().item()) / params.spp}

    render = None
    if not args.no_render and not args.only_sphere_field:
        log("bench: path-traced render 512x512 64spp depth 8")
        render = timed_render(ds, hs.params)
        log(f"bench: render {render}")
    elif not args.no_render:
        render = {}
    if not args.no_render and not args.no_sphere_field:
        log("bench: sphere field 1024x1024 32spp depth 8 (10,008 spheres, BRDF sampling, NEE off)")
        hs2 = HostScene.from_text(scenes.sphere_field_text())
        ds2 = DeviceScene(ctx, hs2)
        render["sphere_field"] = timed_render(ds2, hs2.params)
        log(f"bench: sphere field {render['sphere_field']}")
        del ds2
    if not args.no_render and not args.only_sphere_field:
        log("bench: Cornell box 256x256 16spp depth 5 (the reference's CPU case)")
        hs0 = HostScene.from_text(scenes.cornell_text(256, 256, 16, 5, seed=0))
        ds0 = DeviceScene(ctx, hs0)
        render["cornell_256"] = timed_render(ds0, hs0.params)
        if cpu and "cornell_256_render" in cpu:
            render["cornell_256"]["gpu_over_cpu"] = round(cpu["cornell_256_render"]["seconds"] /
                                                          render["cornell_256"]["seconds"], 1)
        # Mrays/s ratios of the full-size GPU renders over the CPU samples of
        # the same scenes
        if cpu:
            for rkey, ckey in (("icosphere", "icosphere_render"), ("sphere_field", "sphere_field_render")):
                g = render if rkey == "icosphere" else render.get(rkey)
                if g and ckey in cpu and "mrays_per_s" in g:
                    cpu[ckey]["gpu_over_cpu_mrays"] = round(g["mrays_per_s"] / cpu[ckey]["mrays_per_s"], 1)
        log(f"bench: cornell 256 {render['cornell_256']}")
        del ds0

    batched = serving = None
    admission = None
    if not args.no_batch:
        admission = bench_admission(ctx, log)
        log(f"bench: admission {admission}")
        # the batched configuration: one independent scene per job (a
        # displaced icosphere or a sphere field, primitives log-uniform in
        # [1k, 2M]), BVHs built on the device at admission
        from muon_amd import workloads
        specs = workloads.independent_scene_specs(args.batch_jobs, seed=3)
        t0 = time.perf_counter()
        pool = workloads.build_independent_pool(ctx, specs, lambda m: log("bench: " + m))
        pool_s = time.perf_counter() - t0
        jobs = workloads.independent_jobs(pool, seed=7)
        batched, cap = bench_batched(ctx, pool, len(jobs), 7, log, jobs=jobs)
        batched["scene_pool"] = {"scenes": len(pool), "prims": int(sum(i["num_prims"] for *_, i in pool)),
                                 "icospheres": sum(k == "ico" for k, *_ in pool),
                                 "sphere_fields": sum(k == "spheres" for k, *_ in pool),
                                 "admission_s": round(pool_s, 2),
                                 "device_build_ms": round(sum(i["build_ms"] for *_, i in pool), 1)}
        log(f"bench: batched {batched}")
        quality_check = converged_cross_check(ctx, orc, cpu_threads, log)
        small, _ = bench_batched(ctx, pool, 128, 8, log, res_choices=(128,), spp_range=(4, 16))
        log(f"bench: batched small jobs {small}")
        batched["small_jobs"] = {k: small[k] for k in ("jobs", "samples", "batched_s", "sequential_s", "images_per_s",
                                                        "sequential_images_per_s", "speedup", "identical_to_plain",
                                                        "relmse_vs_converged")}
        batched["converged_cross_check"] = quality_check
        del pool, jobs
        # serving: its own 8 resident scenes (about 5.7M primitives); the
        # request rate is a load fraction of the batched throughput measured
        # on these scenes with the request mix
        spool = build_pool(ctx, log, scenes.SERVE_POOL)
        _, cap_serve = bench_batched(ctx, spool, 24, 9, log, res_choices=(256, 512, 1024), spp_range=(8, 128),
                                     sequential=False)
        serving = bench_serving(ctx, spool, args.serve_requests, cap_serve, args.serve_load, 7, log)
        serving["capacity_msamples_per_s"] = round(cap_serve / 1e6, 1)
        serving["scenes"] = [f"{k}:{i['num_prims']}" for k, _, _, _, i in spool]
        serving["total_prims"] = int(sum(i["num_prims"] for _, _, _, _, i in spool))
        log(f"bench: serving {serving}")
        del spool

    out = {
        "metric": METRIC, "value": round(mrays, 1), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(mean_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "closest-hit microbench: 2^24 incoherent rays vs displaced level-7 icosphere "
                               "(327,680 tris) in the Cornell enclosure",
                   "rays_per_step": args.rays, "prims": info["num_prims"], "bvh_depth": info["bvh_depth"],
                   "hit_fraction": round(hits / args.rays, 4), "wall_s": round(wall, 3)},
        "roofline": roofline, "cpu_baseline": cpu, "autotune": tune, "render": render, "batched": batched,
        "serving": serving, "admission": admission,
    }
    if cpu:
        out["gpu_over_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
