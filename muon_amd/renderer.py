"""Host-side mirror of muon's render surface (muon/renderer.h:11-24,
muon/options.h:11-23, muon/film.h:18-47, muon/stats.h:11-64) over the
C ABI. Renderer(scene_file, options).Render() parses the scene, uploads it
to the GPU, renders it there and writes the PNG, as muon::Renderer does."""
import ctypes as C
import dataclasses
import sys
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .scene import Context, DeviceScene, HostScene

ACCELERATION = {"linear": 0, "bvh": 1}
PARTITION = {"uniform": 0, "midpoint": 1, "sah": 2}


@dataclasses.dataclass
class Options:
    """muon::Options: output overrides the scene's output; parallelism only
    applies to host-side work (the render runs on the GPU)."""
    output: str = ""
    acceleration: str = "bvh"
    partition_strategy: str = "sah"
    parallelism: int = 1
    show_stats: bool = True
    default_seed: int = -1
    device: int = 0


@dataclasses.dataclass
class TraceStats:
    primary_rays: int = 0
    secondary_rays: int = 0
    shadow_rays: int = 0
    render_seconds: float = 0.0
    build_seconds: float = 0.0
    culled_rays: int = 0  # of secondary_rays: MIS rays settled without a traversal
    # muon/stats.h:11-45 test counters (deterministic integrators; 0 for the path tracer)
    object_tests: int = 0
    object_hits: int = 0
    bounds_tests: int = 0
    bounds_hits: int = 0

    def __str__(self):
        # muon/stats.cc:41-88
        def rate(h, t):
            return 100.0 * h / t if t else 0.0
        lines = [
            f"{'Build time':18s} : {self.build_seconds:12.2f} (sec)",
            f"{'Render time':18s} : {self.render_seconds:12.2f} (sec)",
            f"{'Primary rays':18s} : {self.primary_rays:12d}",
            f"{'Secondary rays':18s} : {self.secondary_rays:12d}",
            f"{'Shadow rays':18s} : {self.shadow_rays:12d}",
            f"{'Ray-object tests':18s} : {self.object_tests:12d}",
            f"{'Ray-object hits':18s} : {self.object_hits:12d}",
            f"{'Hit rate':18s} : {rate(self.object_hits, self.object_tests):12.2f} %",
            f"{'Bounds tests':18s} : {self.bounds_tests:12d}",
            f"{'Bounds hits':18s} : {self.bounds_hits:12d}",
            f"{'Bounds hit rate':18s} : {rate(self.bounds_hits, self.bounds_tests):12.2f} %",
        ]
        return "\n".join(lines)


class Film:
    """Linear float accumulator [H, W, 3] (sum over samples) and its 8-bit
    output (muon/film.cc:21-43)."""

    def __init__(self, accum, spp, gamma):
        self.accum = np.ascontiguousarray(accum, dtype=np.float32)
        self.spp = spp
        self.gamma = gamma

    @property
    def height(self):
        return self.accum.shape[0]

    @property
    def width(self):
        return self.accum.shape[1]

    def rgb8(self):
        out = np.zeros(self.accum.shape, np.uint8)
        check(lib.muonx_film_to_rgb8(self.accum.ctypes.data, self.width, self.height, self.spp, self.gamma,
                                     out.ctypes.data))
        return out

    def write_png(self, path):
        img = self.rgb8()
        check(lib.muonx_write_png(path.encode(), img.ctypes.data, self.width, self.height))


def render_device(dev_scene, params, accum=None):
    """Renders on the GPU; returns (accum torch tensor [H,W,3], TraceStats)."""
    import torch
    H, W = params.height, params.width
    if accum is None:
        accum = torch.empty((H, W, 3), dtype=torch.float32, device=f"cuda:{dev_scene.ctx.device}")
    st = _lib.TraceStatsC()
    s = torch.cuda.current_stream(accum.device).cuda_stream
    check(lib.muonx_render(dev_scene.h, C.byref(params), accum.data_ptr(), C.byref(st), s))
    stats = TraceStats(st.primary_rays, st.secondary_rays, st.shadow_rays, st.render_ns * 1e-9,
                       culled_rays=st.culled_rays, object_tests=st.object_tests, object_hits=st.object_hits,
                       bounds_tests=st.bounds_tests, bounds_hits=st.bounds_hits)
    return accum, stats


class Renderer:
    """muon::Renderer(scene_file, options).Render()."""

    def __init__(self, scene_file, options=None):
        self.scene_file = scene_file
        self.options = options or Options()

    def Render(self):
        o = self.options
        t0 = time.perf_counter()
        hs = HostScene.from_file(self.scene_file, ACCELERATION[o.acceleration], PARTITION[o.partition_strategy],
                                 o.default_seed)
        ctx = Context(o.device)
        if o.show_stats:  # the path tracer counts its tests only when asked (muon's --stats)
            ctx.set_tuning(count_work=1)
        ds = DeviceScene(ctx, hs)
        t1 = time.perf_counter()
        accum, stats = render_device(ds, hs.params)
        film = Film(accum.cpu().numpy(), hs.params.spp, hs.params.gamma)
        stats.build_seconds = t1 - t0
        # muon/renderer.cc:26-27: --output overrides the scene's output
        self.output = o.output or hs.output
        film.write_png(self.output)
        if o.show_stats:
            print(stats, file=sys.stderr)
        self.film, self.stats = film, stats
        return film
