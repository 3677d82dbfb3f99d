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
  independent_scene_specs(n, seed)
                              the batched configuration's scenes: one per
                              job, a displaced icosphere or a sphere field
                              with primitive count log-uniform in [1k, 2M]
  build_independent_pool(ctx, specs)
                              those scenes resident, their BVHs built on the
                              device (icospheres loaded from binary PLY,
                              sphere fields handed over as arrays)
  independent_jobs(pool, seed)
                              one job per scene: resolution uniform in
                              {128^2 .. 1024^2}, spp uniform in [4, 64], depth 8
"""
import math
import os
import tempfile
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


# ---------------------------------------------------------------- independent scenes
C3_MIN_PRIMS, C3_MAX_PRIMS = 1_000, 2_000_000


def independent_scene_specs(n=64, seed=3, lo=C3_MIN_PRIMS, hi=C3_MAX_PRIMS):
    """One scene per job: a displaced icosphere or a sphere field (each with
    probability 1/2), primitive count log-uniform in [lo, hi], its own
    displacement / placement seed. An icosphere takes the subdivision level
    whose 20 * 4^L triangles lie nearest the drawn count in log (levels 3..8,
    1,280 .. 1,310,720 triangles); a sphere field takes the count itself."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        count = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        if rng.random() < 0.5:
            level = int(np.clip(round(math.log(count / 20.0, 4)), 3, 8))
            specs.append(("ico", level, 100 + i))
        else:
            specs.append(("spheres", int(round(count)), 100 + i))
    return specs


def sphere_field_arrays(n, seed, width=256, height=256, spp=16, max_depth=8, device_bvh=True):
    """The density-scaled sphere field of scenes.sphere_field_text(n, seed,
    scaled=True) (the same spheres, materials and emitters; the emitters
    come first) built from
    arrays instead of text: the parser reads the settings, camera and
    emitters, and the n spheres with their material records are appended to
    its descriptor. Each record is the parser's material state at that
    sphere (muon/parser.cc:769-856: a command clones the current material and
    sets one field, so a Lambert or GGX sphere keeps the exponent of the last
    Phong sphere before it, and Lambert and Phong spheres the roughness of
    the last GGX one). Returns (HostScene, descriptor fields for
    DeviceScene.from_arrays)."""
    half, ne = scenes.sphere_field_extent(n, True)
    rng, c, r, kind, col, expo, rough = scenes.sphere_field_draw(n, seed, half)
    text = "\n".join(scenes.sphere_field_prologue(width, height, spp, max_depth, n, seed, half, ne) +
                     scenes.sphere_field_emitters(rng, half, ne)) + "\n"
    hs = HostScene.from_text(text, device_bvh=device_bvh)
    a = hs.arrays()
    f32 = np.float32
    lam, pho, ggx = kind == 0, kind == 1, kind == 2
    idx = np.arange(n)
    last_pho = np.maximum.accumulate(np.where(pho, idx, -1))
    last_ggx = np.maximum.accumulate(np.where(ggx, idx, -1))
    mats = np.zeros((n, 16), f32)  # ambient, diffuse, specular, emission, shininess, roughness, brdf
    mats[lam, 3:6] = col[lam].astype(f32)
    mats[pho, 3:6] = (0.4 * col[pho]).astype(f32)
    mats[pho, 6:9] = f32(0.4)
    mats[ggx, 3:6] = (0.1 * col[ggx]).astype(f32)
    mats[ggx, 6:9] = f32(0.6)
    mats[:, 12] = np.where(last_pho >= 0, expo[np.maximum(last_pho, 0)].astype(f32), f32(1.0))
    mats[:, 13] = np.where(last_ggx >= 0, rough[np.maximum(last_ggx, 0)].astype(f32), f32(0.0))
    mats[:, 14] = np.where(ggx, 2, 1)  # `brdf lambertian` builds Phong (muon/parser.cc:165-178)
    geom = np.zeros((n, 12), f32)
    geom[:, :3] = c.astype(f32)
    geom[:, 3] = r.astype(f32)
    n0, m0 = hs.desc.num_prims, hs.desc.num_materials
    fields = dict(
        num_prims=n0 + n, num_materials=m0 + n, prim_vnormal=None, prim_use_vnormal=None,
        prim_type=np.concatenate([a["prim_type"], np.ones(n, np.uint8)]),
        prim_geom=np.concatenate([a["prim_geom"], geom]).reshape(-1),
        prim_transform=np.concatenate([a["prim_transform"], np.full(n, a["prim_transform"][0], np.int32)]),
        prim_material=np.concatenate([a["prim_material"], m0 + np.arange(n, dtype=np.int32)]),
        prim_light=np.concatenate([a["prim_light"], np.full(n, -1, np.int32)]),
        materials=np.concatenate([a["materials"], mats]).reshape(-1))
    return hs, fields


def sphere_field_scene(ctx, n, seed, width=256, height=256, spp=16, max_depth=8, device_bvh=True):
    """sphere_field_arrays resident on the device: (HostScene, DeviceScene)."""
    hs, fields = sphere_field_arrays(n, seed, width, height, spp, max_depth, device_bvh)
    return hs, DeviceScene.from_arrays(ctx, hs, **fields)


def build_independent_pool(ctx, specs, log_fn=None, device_bvh=True):
    """Resident scenes for independent_scene_specs: a list of (kind, size,
    HostScene, DeviceScene, info)."""
    pool = []
    with tempfile.TemporaryDirectory() as tmp:
        for kind, size, seed in specs:
            t0 = time.perf_counter()
            if kind == "ico":
                path = os.path.join(tmp, f"ico_{size}_{seed}.ply")
                hs = HostScene.from_text(scenes.icosphere_text(256, 256, 16, 8, level=size, seed=seed,
                                                               mesh_file=path), device_bvh=device_bvh)
                ds = DeviceScene(ctx, hs)
                os.unlink(path)
            else:
                hs, ds = sphere_field_scene(ctx, size, seed, device_bvh=device_bvh)
            info = ds.info()
            pool.append((kind, size, hs, ds, info))
            if log_fn:
                log_fn(f"scene {kind} {size} seed {seed}: {info['num_prims']} prims, "
                       f"{time.perf_counter() - t0:.2f} s (device build {info['build_ms']:.1f} ms)")
    return pool


def independent_jobs(pool, seed=7, res_choices=(128, 256, 512, 1024), spp_range=(4, 64), depth=8):
    """One job per resident scene, in order: (DeviceScene, RenderParams,
    key 0) with a seeded pose, square resolution from res_choices, spp
    uniform in spp_range, the given max depth and a random seed."""
    rng = np.random.default_rng(seed)
    jobs = []
    for kind, size, hs, ds, _ in pool:
        p = _lib.RenderParams.from_buffer_copy(hs.params)
        r = int(rng.choice(list(res_choices)))
        scenes.random_pose(p, kind, rng, r, r, scenes.sphere_field_extent(size, True)[0] if kind == "spheres" else 10.0)
        p.spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        p.max_depth = depth
        p.seed = int(rng.integers(1 << 31))
        jobs.append((ds, p, 0))
    return jobs
