"""Host scene ingest and device-resident scenes over the C ABI."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _view(ptr, count, dtype):
    if not ptr or count == 0:
        return np.zeros(0, dtype)
    ctype = {np.float32: C.c_float, np.int32: C.c_int32, np.uint8: C.c_uint8}[dtype]
    arr = (ctype * count).from_address(ptr)
    return np.ctypeslib.as_array(arr).copy()


class HostScene:
    """A parsed .muon scene with its host-built SAH BVH
    (muon::Parser::Parse + acceleration::BVH::Init)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)
        self.desc = _lib.SceneDesc()
        check(lib.muonx_host_scene_desc(self.h, C.byref(self.desc)))
        self.params = _lib.RenderParams()
        check(lib.muonx_host_scene_params(self.h, C.byref(self.params)))

    @classmethod
    def from_file(cls, path, acceleration=1, partition_strategy=2, default_seed=-1, device_bvh=False):
        """device_bvh: skip the host SAH build; the device builds a BVH when
        the scene is uploaded (fast admission; fast traversal mode only)."""
        opts = _lib.Options(acceleration, partition_strategy, 1, 1, default_seed, int(device_bvh))
        out = C.c_void_p()
        check(lib.muonx_host_scene_parse_file(path.encode(), C.byref(opts), C.byref(out)))
        return cls(out.value)

    @classmethod
    def from_text(cls, text, acceleration=1, partition_strategy=2, default_seed=-1, device_bvh=False):
        opts = _lib.Options(acceleration, partition_strategy, 1, 1, default_seed, int(device_bvh))
        out = C.c_void_p()
        b = text.encode()
        check(lib.muonx_host_scene_parse_text(b, len(b), C.byref(opts), C.byref(out)))
        return cls(out.value)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.muonx_host_scene_destroy(self.h)
            self.h = None

    @property
    def warnings(self):
        return lib.muonx_host_scene_warnings(self.h)

    @property
    def output(self):
        """The scene's output file name (muon/scene.h:36): the `output`
        command's argument, or "raytrace.png"."""
        out = C.c_char_p()
        check(lib.muonx_host_scene_output(self.h, C.byref(out)))
        return out.value.decode()

    def arrays(self):
        """Copies of the descriptor arrays (for tests and tools)."""
        d = self.desc
        n, nn = d.num_prims, d.num_nodes
        return dict(
            prim_type=_view(d.prim_type, n, np.uint8),
            prim_geom=_view(d.prim_geom, 12 * n, np.float32).reshape(n, 12),
            prim_vnormal=_view(d.prim_vnormal, 9 * n, np.float32).reshape(n, 9),
            prim_use_vnormal=_view(d.prim_use_vnormal, n, np.uint8),
            prim_transform=_view(d.prim_transform, n, np.int32),
            prim_material=_view(d.prim_material, n, np.int32),
            prim_light=_view(d.prim_light, n, np.int32),
            transforms=_view(d.transforms, 48 * d.num_transforms, np.float32).reshape(-1, 3, 4, 4),
            materials=_view(d.materials, 16 * d.num_materials, np.float32).reshape(-1, 16),
            lights=_view(d.lights, 16 * d.num_lights, np.float32).reshape(-1, 16),
            node_bounds=_view(d.node_bounds, 6 * nn, np.float32).reshape(nn, 6),
            node_left=_view(d.node_left, nn, np.int32),
            node_right=_view(d.node_right, nn, np.int32),
            node_axis=_view(d.node_axis, nn, np.int32),
            node_start=_view(d.node_start, nn, np.int32),
            node_count=_view(d.node_count, nn, np.int32),
            leaf_prims=_view(d.leaf_prims, n, np.int32),
        )


class Context:
    """One device context (muonx_ctx)."""

    def __init__(self, device=0):
        out = C.c_void_p()
        check(lib.muonx_ctx_create(device, C.byref(out)))
        self.h = out
        self.device = device

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.muonx_ctx_destroy(self.h)
            self.h = None

    def get_tuning(self):
        t = _lib.Tuning()
        check(lib.muonx_ctx_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in t._fields_}

    def set_tuning(self, **kw):
        cur = self.get_tuning()
        cur.update(kw)
        check(lib.muonx_ctx_set_tuning(self.h, C.byref(_lib.Tuning(**cur))))


def autotune(dev_scene, job=None, rays=0):
    """muonx_autotune: times launch configurations on this scene (and the
    render `job`, a RenderParams, if given), installs the fastest in the
    scene's context and returns the report as a dict."""
    r = _lib.AutotuneReport()
    check(lib.muonx_autotune(dev_scene.h, C.byref(job) if job is not None else None, rays, C.byref(r)))
    out = {k: getattr(r, k) for k, _ in r._fields_ if k != "tuning"}
    out["tuning"] = {k: getattr(r.tuning, k) for k, _ in r.tuning._fields_}
    return out


class DeviceScene:
    """A scene resident in HBM (muonx_scene)."""

    def __init__(self, ctx, host_scene, desc=None):
        out = C.c_void_p()
        check(lib.muonx_scene_create(ctx.h, C.byref(desc if desc is not None else host_scene.desc), C.byref(out)))
        self.h = out
        self.ctx = ctx  # keeps the context alive
        self.params = host_scene.params

    @classmethod
    def from_arrays(cls, ctx, host_scene, **replace):
        """A device scene from the host scene's descriptor with some fields
        replaced: counts (ints, e.g. num_prims) and arrays (numpy arrays of the
        descriptor's element type and length, or None for an optional array),
        e.g. materials=... with other BRDF tags, or more primitives. The
        arrays are read during the call only."""
        desc = _lib.SceneDesc.from_buffer_copy(host_scene.desc)
        keep = []
        for name, val in replace.items():
            if val is None or isinstance(val, int):
                setattr(desc, name, val)
                continue
            arr = np.ascontiguousarray(val)
            keep.append(arr)
            setattr(desc, name, arr.ctypes.data)
        return cls(ctx, host_scene, desc)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.muonx_scene_destroy(self.h)
            self.h = None

    def info(self):
        i = _lib.SceneInfo()
        check(lib.muonx_scene_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    # Device-pointer queries. o, d: torch float32 (n,3) on the GPU.
    def query(self, o, d, tmax=None, any_hit=False, exact=False, counters=None, stream=None, prim=None, t=None,
              occ=None, noalloc=False):
        """muonx_intersect (noalloc: muonx_intersect_noalloc, whose scratch
        is a static device array, so the call allocates no device memory)."""
        import torch
        n = o.shape[0]
        if any_hit and occ is None:
            occ = torch.empty(n, dtype=torch.uint8, device=o.device)
        if not any_hit:
            if prim is None:
                prim = torch.empty(n, dtype=torch.int32, device=o.device)
            if t is None:
                t = torch.empty(n, dtype=torch.float32, device=o.device)
        q = _lib.Query(o.data_ptr(), d.data_ptr(), None if tmax is None else tmax.data_ptr(), n, int(any_hit),
                       int(exact), None if prim is None else prim.data_ptr(), None if t is None else t.data_ptr(),
                       None if occ is None else occ.data_ptr(), None if counters is None else counters.data_ptr())
        s = torch.cuda.current_stream(o.device).cuda_stream if stream is None else stream
        fn = lib.muonx_intersect_noalloc if noalloc else lib.muonx_intersect
        check(fn(self.h, C.byref(q), s))
        return occ if any_hit else (prim, t)

    def intersect_closest(self, o, d, tmax=None, prim=None, t=None, stream=None, exact=False):
        return self.query(o, d, tmax, False, exact, None, stream, prim, t)

    def intersect_any(self, o, d, tmax=None, occ=None, stream=None, exact=False):
        return self.query(o, d, tmax, True, exact, None, stream, occ=occ)
