"""Batched and served rendering over the persistent engine of libmuonx
(muonx_render_batch, muonx_server_*). The reference renders one scene per
process (muon/muon.cc:48-49); these entry points run many path-tracer jobs
of resident scenes through one shared pool of path slots.

A job is (scene, params, key): equal params and key give the same image
whatever else is in flight."""
import ctypes as C
import dataclasses

from . import _lib
from ._lib import check, lib


SLICES_OLDEST_FIRST = 0      # a running render's next slice: oldest render first (default)
SLICES_LEAST_WORK_FIRST = 1  # ... or the render with the least work left (lower median, longer tail)
SLICES_RESPONSE_RATIO = 2    # ... or the highest ratio of work handed out since its submission to its work left


def engine_config(path_slots=0, max_jobs=0, shadow_per_vertex=0, iterations_per_poll=0, film_budget_bytes=0,
                  use_graph=True, slice_order=SLICES_OLDEST_FIRST):
    return _lib.EngineConfig(path_slots, max_jobs, shadow_per_vertex, iterations_per_poll, film_budget_bytes,
                             1 if use_graph else 0, slice_order)


@dataclasses.dataclass
class JobResult:
    status: int
    submit_ns: int
    admit_ns: int
    done_ns: int

    @property
    def latency_s(self):
        return (self.done_ns - self.submit_ns) * 1e-9

    @property
    def queue_s(self):
        return (self.admit_ns - self.submit_ns) * 1e-9


def _result(r):
    return JobResult(r.status, r.submit_ns, r.admit_ns, r.done_ns)


def make_job(dev_scene, params, key=0, accum=None):
    """Returns (Job struct, accum tensor [H, W, 3] float32 on the scene's
    device). The tensor must stay alive until the job has finished."""
    import torch
    if accum is None:
        dev = torch.device(f"cuda:{dev_scene.ctx.device}")
        accum = torch.empty((params.height, params.width, 3), dtype=torch.float32, device=dev)
        # the engine writes on its own stream: earlier users of this memory
        # on torch's stream must be finished
        torch.cuda.current_stream(dev).synchronize()
    job = _lib.Job(dev_scene.h, params, key & 0xFFFFFFFF, accum.data_ptr())
    return job, accum


def render_batch(ctx, jobs, config=None):
    """jobs: list of (dev_scene, params, key). Returns (list of accum
    tensors, list of JobResult, TraceStatsC)."""
    import torch
    structs, accums = [], []
    for ds, params, key in jobs:
        j, a = make_job(ds, params, key)
        structs.append(j)
        accums.append(a)
    torch.cuda.synchronize()  # films are written on the engine's own stream
    n = len(structs)
    arr = (_lib.Job * max(n, 1))(*structs)
    res = (_lib.JobResult * max(n, 1))()
    st = _lib.TraceStatsC()
    cfg = C.byref(config) if config is not None else None
    check(lib.muonx_render_batch(ctx.h, arr, n, cfg, res, C.byref(st)))
    return accums, [_result(res[i]) for i in range(n)], st


class Server:
    """Request loop with continuous batching (muonx_server_*)."""

    def __init__(self, ctx, config=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        cfg = C.byref(config) if config is not None else None
        check(lib.muonx_server_create(ctx.h, cfg, C.byref(self.h)))
        self._outputs = {}

    def submit(self, dev_scene, params, key=0, accum=None):
        """Queues a job; returns its ticket. The output tensor is kept until
        wait() collects the ticket."""
        job, accum = make_job(dev_scene, params, key, accum)
        t = C.c_uint64()
        check(lib.muonx_server_submit(self.h, C.byref(job), C.byref(t)))
        self._outputs[t.value] = accum
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """Returns (accum, JobResult), or None on timeout; (None, JobResult)
        with status MUONX_ERR_CANCELLED for a job cancelled before it
        started rendering."""
        r = _lib.JobResult()
        rc = lib.muonx_server_wait(self.h, ticket, timeout_ms, C.byref(r))
        if rc == _lib.MUONX_PENDING:
            return None
        accum = self._outputs.pop(ticket, None)
        if rc == _lib.MUONX_ERR_CANCELLED:
            return None, _result(r)
        check(rc)
        return accum, _result(r)

    def cancel(self, ticket):
        """Drops a job that has not started rendering (muonx_server_cancel);
        a job already rendering finishes normally."""
        check(lib.muonx_server_cancel(self.h, ticket))

    def stats(self):
        s = _lib.ServerStats()
        check(lib.muonx_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.h:
            lib.muonx_server_destroy(self.h)
            self.h = C.c_void_p()
            self._outputs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
