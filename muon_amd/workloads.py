"""The batched and served workloads: resident scene pools and seeded job /
request mixes, shared by bench.py and the GPU tests so that both run the
same mixes.

  build_pool(ctx, specs)      resident scenes (HostScene + DeviceScene) for
                              (kind, size[, seed]) specs of scenes.POOL /
                              scenes.SERVE_POOL
  batched_jobs(pool, n, seed) n jobs of uneven size: resolution uniform in
                              {128^2, 256^2, 512^2, 1024^2}, spp uniform in
                              [4, 64], depth 8, a random scene of the pool and
                              a seeded camera pose
  served_requests(pool, n, seed)
                              the request mix of the served workload:
                              resolution {256^2, 512^2, 1024^2}, spp 8..128
"""
import time

import numpy as np

from . import _lib, scenes
from .scene import DeviceScene, HostScene


def build_pool(ctx, specs, log_fn=None):
    """Resident scenes for (kind, size[, seed]) specs: a list of
    (kind, size, HostScene, DeviceScene, info)."""
    pool = []
    for spec in specs:
        kind, size, seed = (tuple(spec) + (0,))[:3]
        t0 = time.perf_counter()
        hs = HostScene.from_text(scenes.pool_scene_text(kind, size, seed))
        ds = DeviceScene(ctx, hs)
        pool.append((kind, size, hs, ds, ds.info()))
        if log_fn:
            log_fn(f"pool scene {kind} {size}: {ds.info()['num_prims']} prims, {time.perf_counter() - t0:.1f} s")
    return pool


def make_job(pool, rng, res_choices, spp_lo, spp_hi, depth=8):
    """One path-tracer job on a random scene of the pool: (DeviceScene,
    RenderParams) with a seeded pose, square resolution from res_choices,
    spp uniform in [spp_lo, spp_hi], the given max depth and a random seed."""
    kind, size, hs, ds, _ = pool[int(rng.integers(len(pool)))]
    p = _lib.RenderParams.from_buffer_copy(hs.params)
    r = int(rng.choice(res_choices))
    scenes.random_pose(p, kind, rng, r, r)
    p.spp = int(rng.integers(spp_lo, spp_hi + 1))
    p.max_depth = depth
    p.seed = int(rng.integers(1 << 31))
    return ds, p


def batched_jobs(pool, n=64, seed=7, res_choices=(128, 256, 512, 1024), spp_range=(4, 64)):
    """The batched workload's job mix: a list of (DeviceScene, RenderParams,
    key 0); key 0 for every job, as the plain path renders with key 0, so the
    batch and the plain renders must give the same images."""
    rng = np.random.default_rng(seed)
    return [make_job(pool, rng, list(res_choices), spp_range[0], spp_range[1]) + (0,) for _ in range(n)]


def served_requests(pool, n=48, seed=7):
    """The served workload's request mix: a list of (DeviceScene,
    RenderParams); request i renders with key i."""
    rng = np.random.default_rng(seed + 1)
    return [make_job(pool, rng, [256, 512, 1024], 8, 128) for _ in range(n)]


def poisson_gaps(n, rate, seed=7):
    """Inter-arrival times of a Poisson stream at `rate` per second."""
    rng = np.random.default_rng(seed + 2)
    return rng.exponential(1.0 / rate, n)
