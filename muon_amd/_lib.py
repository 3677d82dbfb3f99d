"""ctypes bindings of include/muonx.h (libmuonx.so)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(HERE, "lib", "libmuonx.so")


class MuonxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"muonx error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [("acceleration", C.c_int32), ("partition_strategy", C.c_int32), ("parallelism", C.c_int32),
                ("show_stats", C.c_int32), ("default_seed", C.c_int64), ("device_bvh", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_prims", C.c_uint32), ("prim_type", C.c_void_p), ("prim_geom", C.c_void_p),
        ("prim_vnormal", C.c_void_p), ("prim_use_vnormal", C.c_void_p), ("prim_transform", C.c_void_p),
        ("prim_material", C.c_void_p), ("prim_light", C.c_void_p),
        ("num_transforms", C.c_uint32), ("transforms", C.c_void_p),
        ("num_materials", C.c_uint32), ("materials", C.c_void_p),
        ("num_lights", C.c_uint32), ("lights", C.c_void_p),
        ("num_nodes", C.c_uint32), ("node_bounds", C.c_void_p), ("node_left", C.c_void_p),
        ("node_right", C.c_void_p), ("node_axis", C.c_void_p), ("node_start", C.c_void_p),
        ("node_count", C.c_void_p), ("leaf_prims", C.c_void_p), ("acceleration", C.c_int32),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("min_depth", C.c_int32),
        ("max_depth", C.c_int32), ("integrator", C.c_int32), ("nee", C.c_int32),
        ("importance_sampling", C.c_int32), ("russian_roulette", C.c_int32), ("light_samples", C.c_int32),
        ("light_stratify", C.c_int32), ("seed", C.c_uint32), ("gamma", C.c_float),
        ("eye", C.c_float * 3), ("cam_u", C.c_float * 3), ("cam_v", C.c_float * 3), ("cam_w", C.c_float * 3),
        ("tan_fov_x", C.c_float), ("tan_fov_y", C.c_float), ("has_camera", C.c_int32),
        ("sample_mode", C.c_int32),
    ]


class TraceStatsC(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("secondary_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("render_ns", C.c_uint64), ("culled_rays", C.c_uint64), ("object_tests", C.c_uint64),
                ("object_hits", C.c_uint64), ("bounds_tests", C.c_uint64), ("bounds_hits", C.c_uint64)]


class Query(C.Structure):
    _fields_ = [("ray_o", C.c_void_p), ("ray_d", C.c_void_p), ("tmax", C.c_void_p), ("n", C.c_uint32),
                ("any_hit", C.c_int32), ("mode", C.c_int32), ("prim_id", C.c_void_p), ("t", C.c_void_p),
                ("occluded", C.c_void_p), ("counters", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("stack_entries", C.c_int32), ("while_while", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("path_budget_mb", C.c_int32), ("bvh_width", C.c_int32), ("ray_fetch", C.c_int32),
                ("node_bits", C.c_int32), ("count_work", C.c_int32)]


class AutotuneReport(C.Structure):
    _fields_ = [("tuning", Tuning), ("candidates", C.c_uint32), ("sample_rays", C.c_uint32),
                ("best_ms", C.c_float), ("worst_ms", C.c_float), ("previous_ms", C.c_float),
                ("render_ms", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("device_bytes", C.c_uint64), ("num_prims", C.c_uint32), ("num_nodes", C.c_uint32),
                ("bvh_depth", C.c_uint32), ("num_lights", C.c_uint32), ("num_quad_lights", C.c_uint32),
                ("device_built", C.c_uint32), ("build_ms", C.c_float)]


class EngineConfig(C.Structure):
    _fields_ = [("path_slots", C.c_uint32), ("max_jobs", C.c_uint32), ("shadow_per_vertex", C.c_int32),
                ("iterations_per_poll", C.c_uint32), ("film_budget_bytes", C.c_uint64), ("use_graph", C.c_int32),
                ("slice_order", C.c_int32)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", RenderParams), ("key", C.c_uint32), ("accum_rgb", C.c_void_p)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("submit_ns", C.c_uint64), ("admit_ns", C.c_uint64),
                ("done_ns", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("pending", C.c_uint64),
                ("in_flight", C.c_uint64), ("iterations", C.c_uint64), ("primary_rays", C.c_uint64),
                ("secondary_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("film_bytes_in_use", C.c_uint64),
                ("film_budget_bytes", C.c_uint64), ("path_slots", C.c_uint32), ("shadow_per_vertex", C.c_int32),
                ("cancelled", C.c_uint64)]


MUONX_PENDING = 1
MUONX_ERR_CANCELLED = -7

P = C.c_void_p
# (name, restype, argtypes): every entry point declared in include/muonx.h
SIGNATURES = [
    ("muonx_last_error", C.c_char_p, []),
    ("muonx_version", C.c_char_p, []),
    ("muonx_host_scene_parse_file", C.c_int, [C.c_char_p, C.POINTER(Options), C.POINTER(P)]),
    ("muonx_host_scene_parse_text", C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(Options), C.POINTER(P)]),
    ("muonx_host_scene_destroy", None, [P]),
    ("muonx_host_scene_desc", C.c_int, [P, C.POINTER(SceneDesc)]),
    ("muonx_host_scene_params", C.c_int, [P, C.POINTER(RenderParams)]),
    ("muonx_host_scene_warnings", C.c_int, [P]),
    ("muonx_host_scene_output", C.c_int, [P, C.POINTER(C.c_char_p)]),
    ("muonx_ctx_create", C.c_int, [C.c_int, C.POINTER(P)]),
    ("muonx_ctx_destroy", None, [P]),
    ("muonx_ctx_set_tuning", C.c_int, [P, C.POINTER(Tuning)]),
    ("muonx_ctx_get_tuning", C.c_int, [P, C.POINTER(Tuning)]),
    ("muonx_autotune", C.c_int, [P, C.POINTER(RenderParams), C.c_uint32, C.POINTER(AutotuneReport)]),
    ("muonx_scene_create", C.c_int, [P, C.POINTER(SceneDesc), C.POINTER(P)]),
    ("muonx_scene_destroy", None, [P]),
    ("muonx_scene_get_info", C.c_int, [P, C.POINTER(SceneInfo)]),
    ("muonx_intersect_closest", C.c_int, [P, P, P, P, C.c_uint32, P, P, P]),
    ("muonx_intersect_any", C.c_int, [P, P, P, P, C.c_uint32, P, P]),
    ("muonx_intersect", C.c_int, [P, C.POINTER(Query), P]),
    ("muonx_intersect_noalloc", C.c_int, [P, C.POINTER(Query), P]),
    ("muonx_render", C.c_int, [P, C.POINTER(RenderParams), P, C.POINTER(TraceStatsC), P]),
    ("muonx_render_host", C.c_int, [P, C.POINTER(RenderParams), P, C.POINTER(TraceStatsC)]),
    ("muonx_render_batch", C.c_int, [P, C.POINTER(Job), C.c_uint32, C.POINTER(EngineConfig),
                                     C.POINTER(JobResult), C.POINTER(TraceStatsC)]),
    ("muonx_server_create", C.c_int, [P, C.POINTER(EngineConfig), C.POINTER(P)]),
    ("muonx_server_submit", C.c_int, [P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    ("muonx_server_wait", C.c_int, [P, C.c_uint64, C.c_int64, C.POINTER(JobResult)]),
    ("muonx_server_cancel", C.c_int, [P, C.c_uint64]),
    ("muonx_server_get_stats", C.c_int, [P, C.POINTER(ServerStats)]),
    ("muonx_server_destroy", None, [P]),
    ("muonx_film_to_rgb8", C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, C.c_float, P]),
    ("muonx_write_png", C.c_int, [C.c_char_p, P, C.c_int32, C.c_int32]),
]


def _load():
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7)
    # but a different file name. Loading it first makes libmuonx bind to that
    # already-loaded runtime instead of pulling a second HIP/HSA runtime into
    # the process, which then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(library_path):
        raise ImportError(
            f"{library_path} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(library_path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    if rc != 0:
        raise MuonxError(rc, lib.muonx_last_error().decode(errors="replace"))
    return rc
