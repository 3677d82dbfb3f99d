"""ctypes bindings for oracle/liboracle.so, the CPU restatement of the
reference renderer. Test infrastructure: used only by tests/, smoke() and
bench.py's CPU baseline, always as the oracle results are compared with."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_c = ctypes
_P = _c.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data


class Scene:
    def __init__(self, lib, handle):
        self._lib = lib
        self.h = handle
        ints = np.zeros(16, np.int32)
        fl = np.zeros(4, np.float32)
        lib.oracle_info(self.h, ints.ctypes.data, fl.ctypes.data)
        (self.width, self.height, self.spp, self.num_prims, self.integrator, self.nee,
         self.importance_sampling, self.min_depth, self.max_depth, self.russian_roulette,
         self.light_samples, self.light_stratify, self.num_lights, self.has_camera,
         self.warnings, self.compute_vertex_normals) = [int(x) for x in ints]
        self.gamma = float(fl[0])

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.oracle_free(self.h)
            self.h = None

    def render(self, parallelism=1):
        acc = np.zeros((self.height, self.width, 3), np.float32)
        rgb = np.zeros((self.height, self.width, 3), np.uint8)
        st = np.zeros(7, np.uint64)
        secs = self._lib.oracle_render(self.h, parallelism, _ptr(acc), _ptr(rgb), _ptr(st))
        return acc, rgb, st, secs

    def intersect(self, o, d, threads=8):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        st = np.zeros(7, np.uint64)
        self._lib.oracle_intersect(self.h, n, _ptr(o), _ptr(d), _ptr(prim), _ptr(t), _ptr(pos), _ptr(nrm),
                                   _ptr(st), threads)
        return prim, t, pos, nrm, st

    def occluded(self, o, d, tmax=None, threads=8):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        occ = np.zeros(n, np.uint8)
        st = np.zeros(7, np.uint64)
        tm = None if tmax is None else np.ascontiguousarray(tmax, np.float32)
        self._lib.oracle_occluded(self.h, n, _ptr(o), _ptr(d), _ptr(tm), _ptr(occ), _ptr(st), threads)
        return occ, st

    def prim_distance(self, prim_id, o, d):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        out = _c.c_float(0.0)
        r = self._lib.oracle_prim_intersect(self.h, int(prim_id), _ptr(o), _ptr(d), _c.byref(out))
        return float(out.value) if r == 1 else None

    def prim_edge(self, prim_id, o, d):
        """Smallest normalised barycentric coordinate of the ray's hit on the
        plane of triangle `prim_id` (negative outside the triangle, 1 for a
        sphere); None for an unknown id or a ray parallel to the plane."""
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        out = _c.c_float(0.0)
        r = self._lib.oracle_prim_edge(self.h, int(prim_id), _ptr(o), _ptr(d), _c.byref(out))
        return float(out.value) if r == 0 else None

    def camera(self):
        out = np.zeros(14, np.float32)
        if self._lib.oracle_camera(self.h, out.ctypes.data) != 0:
            return None
        return out

    def pathtracer_seed(self):
        return int(self._lib.oracle_pathtracer_seed(self.h))

    def bvh(self):
        n = int(self._lib.oracle_bvh_num_nodes(self.h))
        b = np.zeros((n, 6), np.float32)
        cnt = np.zeros(n, np.int64)
        start = np.zeros(n, np.int64)
        axis = np.zeros(n, np.int32)
        left = np.zeros(n, np.int64)
        right = np.zeros(n, np.int64)
        if n:
            self._lib.oracle_bvh_dump(self.h, _ptr(b), _ptr(cnt), _ptr(start), _ptr(axis), _ptr(left), _ptr(right))
        order = np.zeros(self.num_prims, np.int32)
        self._lib.oracle_prim_order(self.h, _ptr(order))
        return dict(bounds=b, count=cnt, start=start, axis=axis, left=left, right=right, prim_order=order)

    def prims(self):
        n = self.num_prims
        typ = np.zeros(n, np.int32)
        xf = np.zeros((n, 3, 4, 4), np.float32)
        geom = np.zeros((n, 12), np.float32)
        mat = np.zeros((n, 16), np.float32)
        light = np.zeros(n, np.int32)
        self._lib.oracle_prim_dump(self.h, _ptr(typ), _ptr(xf), _ptr(geom), _ptr(mat), _ptr(light))
        return dict(type=typ, xform=xf, geom=geom, mat=mat, light=light)


class Oracle:
    def __init__(self):
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        lib = _c.CDLL(LIB)
        lib.oracle_load.restype = _P
        lib.oracle_load.argtypes = [_c.c_char_p, _c.c_int, _c.c_int, _c.c_longlong]
        lib.oracle_load_text.restype = _P
        lib.oracle_load_text.argtypes = [_c.c_char_p, _c.c_int, _c.c_int, _c.c_longlong]
        lib.oracle_load_text_lambertian.restype = _P
        lib.oracle_load_text_lambertian.argtypes = [_c.c_char_p, _c.c_longlong]
        lib.oracle_free.argtypes = [_P]
        lib.oracle_info.argtypes = [_P, _P, _P]
        lib.oracle_camera.argtypes = [_P, _P]
        lib.oracle_pathtracer_seed.argtypes = [_P]
        lib.oracle_pathtracer_seed.restype = _c.c_uint32
        lib.oracle_render.restype = _c.c_double
        lib.oracle_render.argtypes = [_P, _c.c_int, _P, _P, _P]
        lib.oracle_intersect.argtypes = [_P, _c.c_int64, _P, _P, _P, _P, _P, _P, _P, _c.c_int]
        lib.oracle_occluded.argtypes = [_P, _c.c_int64, _P, _P, _P, _P, _P, _c.c_int]
        lib.oracle_prim_intersect.argtypes = [_P, _c.c_int32, _P, _P, _P]
        lib.oracle_prim_edge.argtypes = [_P, _c.c_int32, _P, _P, _P]
        lib.oracle_bvh_num_nodes.restype = _c.c_int64
        lib.oracle_bvh_num_nodes.argtypes = [_P]
        lib.oracle_bvh_dump.argtypes = [_P] * 7
        lib.oracle_prim_order.argtypes = [_P, _P]
        lib.oracle_prim_dump.argtypes = [_P] * 6
        lib.oracle_num_tiles.argtypes = [_c.c_int] * 4
        lib.oracle_mt19937.argtypes = [_c.c_uint32, _c.c_int64, _P]
        lib.oracle_uniform.argtypes = [_c.c_uint32, _c.c_int64, _P]
        self.lib = lib

    def load(self, path, accel=1, partition=2, seed=0):
        return Scene(self.lib, self.lib.oracle_load(path.encode(), accel, partition, seed))

    def load_text(self, text, accel=1, partition=2, seed=0):
        return Scene(self.lib, self.lib.oracle_load_text(text.encode(), accel, partition, seed))

    def load_text_lambertian(self, text, seed=0):
        """load_text with `brdf lambertian` building the Lambertian BRDF
        (the reference's parser builds Phong for it)."""
        return Scene(self.lib, self.lib.oracle_load_text_lambertian(text.encode(), seed))

    def num_tiles(self, w, h, spp, par):
        return int(self.lib.oracle_num_tiles(w, h, spp, par))

    def mt19937(self, seed, n):
        out = np.zeros(n, np.uint32)
        self.lib.oracle_mt19937(seed, n, out.ctypes.data)
        return out

    def uniform(self, seed, n):
        out = np.zeros(n, np.float32)
        self.lib.oracle_uniform(seed, n, out.ctypes.data)
        return out


def image_mae(a, b):
    """ImageMagick `compare -metric MAE` normalised value for 8-bit RGB."""
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).mean() / 255.0)
