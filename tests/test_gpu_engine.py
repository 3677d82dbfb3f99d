"""Persistent engine: batched and served path tracing.

A job's film depends only on (scene, params, key): fixed-point film sums
and random numbers keyed by (seed, key, pixel, sample). So every job must
come out bit-identical whether it runs alone (muonx_render), in a batch of
uneven jobs (muonx_render_batch), through the request loop while other
requests come and go (muonx_server_*), or through a pool of path slots far
smaller than the job (slot regeneration)."""
import ctypes as C
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene, MuonxError, scenes  # noqa: E402
from muon_amd import _lib  # noqa: E402
from muon_amd.renderer import render_device  # noqa: E402
from muon_amd.serving import Server, engine_config, render_batch  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return Context(0)


@pytest.fixture(scope="module")
def resident(ctx):
    """Three resident scenes of different kinds and sizes."""
    out = {}
    out["cornell"] = HostScene.from_text(scenes.cornell_text(64, 64, 4, 5))
    out["ico"] = HostScene.from_text(scenes.icosphere_text(48, 40, 4, 8, level=3))
    out["spheres"] = HostScene.from_text(scenes.sphere_field_text(40, 56, 4, 6, n=300))
    return {k: (hs, DeviceScene(ctx, hs)) for k, hs in out.items()}


def params_of(hs, **kw):
    p = _lib.RenderParams.from_buffer_copy(hs.params)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def alone(ds, p):
    acc, _ = render_device(ds, p)
    torch.cuda.synchronize()
    return acc.cpu().numpy()


def job_list(resident):
    """Uneven jobs: sizes, spp, depth, seeds and integrator options differ."""
    jobs = []
    specs = [
        ("cornell", dict(width=64, height=64, spp=4, max_depth=5, seed=0)),
        ("ico", dict(width=48, height=40, spp=3, max_depth=8, seed=5)),
        ("spheres", dict(width=40, height=56, spp=2, max_depth=6, seed=9)),
        ("cornell", dict(width=32, height=24, spp=7, max_depth=3, seed=11, nee=1)),
        ("cornell", dict(width=50, height=30, spp=2, max_depth=4, seed=2, nee=0, russian_roulette=1)),
        ("ico", dict(width=16, height=16, spp=9, max_depth=2, seed=1)),
    ]
    for name, kw in specs:
        hs, ds = resident[name]
        jobs.append((ds, params_of(hs, **kw), 0))
    return jobs


def test_batch_equals_alone(ctx, resident):
    """Every job's image in the batch is its image alone, bit for bit. The
    jobs without NEE (the sphere field, the Cornell box with nee 0) render
    alone with the single shade kernel and in this batch, which holds NEE
    jobs, with the split NEE / BSDF kernels, so this also checks that the two
    kernel sets give the same bits."""
    jobs = job_list(resident)
    ref = [alone(ds, p) for ds, p, _ in jobs]
    accs, results, st = render_batch(ctx, jobs)
    torch.cuda.synchronize()
    for i, (a, r) in enumerate(zip(accs, ref)):
        got = a.cpu().numpy()
        assert got.shape == r.shape
        assert np.array_equal(got.view(np.uint32), r.view(np.uint32)), f"job {i}"
        assert results[i].status == 0
        assert results[i].submit_ns <= results[i].admit_ns <= results[i].done_ns
    assert st.primary_rays >= sum(p.width * p.height * p.spp for _, p, _ in jobs) * 0.5


def test_small_slot_pool_regenerates_paths(ctx, resident):
    """1024 path slots for ~40k samples: paths are refilled many times."""
    jobs = job_list(resident)[:3]
    ref = [alone(ds, p) for ds, p, _ in jobs]
    cfg = engine_config(path_slots=1024, max_jobs=4, shadow_per_vertex=2, iterations_per_poll=3)
    accs, results, _ = render_batch(ctx, jobs, cfg)
    torch.cuda.synchronize()
    for a, r in zip(accs, ref):
        assert np.array_equal(a.cpu().numpy().view(np.uint32), r.view(np.uint32))
    # a job slot pool of 4 with 3 jobs, then a second batch on the same engine
    accs2, _, _ = render_batch(ctx, jobs[::-1], cfg)
    torch.cuda.synchronize()
    for a, r in zip(accs2, ref[::-1]):
        assert np.array_equal(a.cpu().numpy().view(np.uint32), r.view(np.uint32))


def test_keys_change_the_noise_not_the_estimate(ctx, resident):
    hs, ds = resident["cornell"]
    p = params_of(hs, width=32, height=32, spp=8, max_depth=5, seed=0)
    accs, _, _ = render_batch(ctx, [(ds, p, 0), (ds, p, 1), (ds, p, 0)])
    a0, a1, a2 = (a.cpu().numpy().astype(np.float64) for a in accs)
    assert np.array_equal(a0, a2)
    assert not np.array_equal(a0, a1)
    # independent estimates of one image: the mean per-pixel difference is
    # zero up to its standard error
    d = (a0 - a1).sum(axis=2).ravel()
    assert abs(d.mean()) < 6.0 * d.std() / np.sqrt(d.size)


@pytest.mark.parametrize("order", [0, 1, 2], ids=["oldest-first", "least-work-first", "response-ratio"])
def test_server_results_equal_alone_under_load(ctx, resident, order):
    """Requests arrive while others run; the film budget admits only a few
    at a time, so later requests wait for films to free up. Every order of
    continuing slices gives every request its image alone."""
    jobs = job_list(resident) * 3
    ref = [alone(ds, p) for ds, p, _ in jobs]
    biggest = max(p.width * p.height for _, p, _ in jobs) * 24
    cfg = engine_config(path_slots=8192, max_jobs=8, shadow_per_vertex=4, iterations_per_poll=2,
                        film_budget_bytes=2 * biggest + 4096, slice_order=order)
    with Server(ctx, cfg) as srv:
        tickets = []
        for i, (ds, p, key) in enumerate(jobs):
            tickets.append(srv.submit(ds, p, key))
            if i % 4 == 3:
                time.sleep(0.002)
        outs = [srv.wait(t, timeout_ms=60000) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(jobs) and st["submitted"] == len(jobs)
    assert st["film_bytes_in_use"] == 0
    for i, (out, r) in enumerate(zip(outs, ref)):
        assert out is not None, f"request {i} timed out"
        acc, res = out
        assert res.status == 0 and res.submit_ns <= res.admit_ns <= res.done_ns
        assert np.array_equal(acc.cpu().numpy().view(np.uint32), r.view(np.uint32)), f"request {i}"


def test_server_cancel(ctx, resident):
    """The film budget holds one 512x512 film: B and C wait while A renders.
    Cancelling B drops it (its film is never written, wait reports
    MUONX_ERR_CANCELLED); cancelling A, already rendering, changes nothing;
    C runs after A. A and C equal their renders alone."""
    hs, ds = resident["ico"]
    # A renders for tens of milliseconds, so B is still waiting when cancelled
    p = params_of(hs, width=512, height=512, spp=128, max_depth=8)
    q = params_of(hs, width=512, height=512, spp=4, max_depth=8, seed=3)
    ref_a, ref_c = alone(ds, p), alone(ds, q)
    cfg = engine_config(max_jobs=4, film_budget_bytes=512 * 512 * 24 + 512 * 512 * 12)
    sentinel = torch.full((512, 512, 3), 7.0, dtype=torch.float32, device="cuda:0")
    out_c = torch.empty((512, 512, 3), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    with Server(ctx, cfg) as srv:
        ta = srv.submit(ds, p)
        t_end = time.monotonic() + 30.0
        while srv.stats()["in_flight"] == 0 and time.monotonic() < t_end:  # until A is admitted
            time.sleep(0.0005)
        tb = srv.submit(ds, q, accum=sentinel)
        tc = srv.submit(ds, q, accum=out_c)
        srv.cancel(tb)
        srv.cancel(ta)  # rendering: completes normally
        acc_b, res_b = srv.wait(tb, timeout_ms=60000)
        assert acc_b is None and res_b.status == _lib.MUONX_ERR_CANCELLED
        acc_a, res_a = srv.wait(ta, timeout_ms=60000)
        acc_c, res_c = srv.wait(tc, timeout_ms=60000)
        st = srv.stats()
        with pytest.raises(MuonxError) as e:
            srv.cancel(tb)  # collected
        assert e.value.code == -1
        with pytest.raises(MuonxError) as e:
            srv.cancel(10 ** 9)  # never issued
        assert e.value.code == -1
    torch.cuda.synchronize()
    assert res_a.status == 0 and res_c.status == 0
    assert np.array_equal(acc_a.cpu().numpy().view(np.uint32), ref_a.view(np.uint32))
    assert np.array_equal(acc_c.cpu().numpy().view(np.uint32), ref_c.view(np.uint32))
    assert bool((sentinel == 7.0).all())
    assert st["cancelled"] == 1 and st["completed"] == 2 and st["film_bytes_in_use"] == 0


def test_server_rejects_more_than_65536_job_slots(ctx):
    """A path's job word holds its job slot in 16 bits."""
    with pytest.raises(Exception, match="max_jobs"):
        with Server(ctx, engine_config(path_slots=4096, max_jobs=65537)):
            pass


def test_server_rejects_unknown_slice_order(ctx):
    with pytest.raises(Exception, match="slice_order"):
        with Server(ctx, engine_config(path_slots=4096, slice_order=7)):
            pass


def test_server_rejects_what_can_never_fit(ctx, resident):
    hs, ds = resident["cornell"]
    cfg = engine_config(path_slots=4096, max_jobs=4, shadow_per_vertex=1, film_budget_bytes=64 * 64 * 24)
    with Server(ctx, cfg) as srv:
        with pytest.raises(MuonxError) as e:
            srv.submit(ds, params_of(hs, width=128, height=128, spp=1))
        assert e.value.code == -6
        with pytest.raises(MuonxError) as e:
            srv.submit(ds, params_of(hs, width=32, height=32, spp=1, light_samples=4, light_stratify=1))
        assert e.value.code == -6  # 4 shadow rays per vertex, the engine holds 1
        with pytest.raises(MuonxError) as e:
            srv.submit(ds, params_of(hs, width=32, height=32, spp=1, integrator=0))
        assert e.value.code == -1
        # a valid request still runs after the rejections
        t = srv.submit(ds, params_of(hs, width=32, height=32, spp=2))
        acc, res = srv.wait(t, timeout_ms=60000)
        assert res.status == 0 and float(acc.sum()) > 0
        with pytest.raises(MuonxError):
            srv.wait(t, timeout_ms=0)  # already collected


def test_server_wait_timeout_and_destroy_drains(ctx, resident):
    hs, ds = resident["ico"]
    p = params_of(hs, width=64, height=64, spp=16, max_depth=8)
    ref = alone(ds, p)
    srv = Server(ctx, engine_config(path_slots=2048, max_jobs=4))
    t = srv.submit(ds, p, 0)
    out = srv.wait(t, timeout_ms=0)
    if out is None:  # still running: collect it properly
        out = srv.wait(t, timeout_ms=60000)
    assert np.array_equal(out[0].cpu().numpy().view(np.uint32), ref.view(np.uint32))
    # destroy with a request outstanding finishes it first
    t2 = srv.submit(ds, p, 0)
    acc2 = srv._outputs[t2]
    srv.close()
    torch.cuda.synchronize()
    assert np.array_equal(acc2.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_new_scene_at_a_recycled_address_is_not_confused(resident):
    """A context's engine caches the scenes it has seen. A scene created after
    another was destroyed can get the same host address; it must still be
    rendered from its own BVH (regression: the cache was keyed by address)."""
    import gc
    ctx2 = Context(0)
    hs_a = HostScene.from_text(scenes.cornell_text(32, 32, 2, 4))
    ds_a = DeviceScene(ctx2, hs_a)
    alone(ds_a, params_of(hs_a))
    addr_a = ds_a.h.value
    del ds_a
    gc.collect()
    hs_b, _ = resident["ico"]
    p = params_of(hs_b, width=32, height=32, spp=2, max_depth=5)
    ds_b = DeviceScene(ctx2, hs_b)
    got = alone(ds_b, p)
    fresh = alone(DeviceScene(Context(0), hs_b), p)
    assert np.array_equal(got.view(np.uint32), fresh.view(np.uint32)), ("recycled" if ds_b.h.value == addr_a else "")


def test_scene_destroy_waits_for_server_jobs(ctx, resident):
    hs, _ = resident["cornell"]
    p = params_of(hs, width=64, height=64, spp=32, max_depth=5)
    ds = DeviceScene(ctx, hs)
    ref = alone(ds, p)
    with Server(ctx, engine_config(path_slots=4096, max_jobs=4)) as srv:
        t = srv.submit(ds, p, 0)
        acc = srv._outputs[t]
        _lib.lib.muonx_scene_destroy(ds.h)  # blocks until the job is done
        ds.h = None
        out = srv.wait(t, timeout_ms=60000)
    assert out is not None and out[1].status == 0
    assert np.array_equal(acc.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_wide_bvh_renders_the_same_image(resident):
    """Path tracing over 4-wide nodes: the same closest hits except at
    distance ties, so the images agree up to the rare tie pixels."""
    ctx2 = Context(0)
    hs, _ = resident["ico"]
    p = params_of(hs, width=64, height=64, spp=8, max_depth=8)
    ds = DeviceScene(ctx2, hs)
    ctx2.set_tuning(bvh_width=2)
    a2 = alone(ds, p)
    ctx2.set_tuning(bvh_width=4)
    a4 = alone(ds, p)
    diff = np.abs(a2 - a4).max(axis=2) > 1e-4 * np.maximum(np.abs(a2).max(axis=2), 1.0)
    assert diff.mean() < 0.01, diff.mean()


def test_context_destroyed_without_rendering():
    """A context that never rendered has no workspace (regression: its
    destruction dereferenced the missing workspace)."""
    import gc
    c = Context(0)
    del c
    gc.collect()
    c = Context(0)
    c.set_tuning(bvh_width=4)
    del c
    gc.collect()


def test_autotune_installs_the_fastest_configuration(resident):
    """muonx_autotune times the traversal candidates and the path-queue
    capacity, installs the fastest, and queries keep their results."""
    from muon_amd.scene import autotune
    c = Context(0)
    hs, _ = resident["ico"]
    ds = DeviceScene(c, hs)
    rng = np.random.default_rng(3)
    lo, hi = scenes.icosphere_world_bounds(level=3)
    o, d = scenes.microbench_rays(1 << 16, lo, hi, seed=4)
    dev = torch.device("cuda:0")
    to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    p0, t0 = ds.intersect_closest(to, td)
    torch.cuda.synchronize()
    p0, t0 = p0.cpu().numpy(), t0.cpu().numpy()
    job = params_of(hs, width=48, height=48, spp=2, max_depth=4)
    rep = autotune(ds, job, 1 << 18)
    assert rep["candidates"] >= 24 and rep["sample_rays"] == 1 << 18
    assert 0 < rep["best_ms"] <= rep["worst_ms"] and rep["best_ms"] <= rep["previous_ms"]
    assert rep["render_ms"] > 0
    assert c.get_tuning() == rep["tuning"]
    p1, t1 = ds.intersect_closest(to, td)
    torch.cuda.synchronize()
    p1, t1 = p1.cpu().numpy(), t1.cpu().numpy()
    same = p1 == p0
    # node width and visit order can only change which of equal-distance hits wins
    assert same.mean() > 0.999
    assert np.allclose(t1[~same], t0[~same], rtol=1e-5)
    img = alone(ds, job)
    assert np.isfinite(img).all() and img.sum() > 0


@pytest.mark.parametrize("order", [0, 1, 2], ids=["oldest-first", "least-work-first", "response-ratio"])
def test_server_large_job_completes_under_a_request_stream(ctx, resident, order):
    """One large render (128 slices of a 4096-path engine) while a stream
    keeps up to 8 small requests waiting: new requests' first slices go
    first and least-work-first (and, until it has waited long, the
    response ratio) prefers the small ones, but a render passed over 16
    times by newer work goes ahead (tests/cpu/engine_plan_check.cpp
    checks that bound), so the large render finishes while the stream still
    runs, with its image alone."""
    hs, ds = resident["cornell"]
    big = params_of(hs, width=64, height=64, spp=32, max_depth=5, seed=4)
    small = params_of(hs, width=16, height=16, spp=4, max_depth=5, seed=5)
    ref = alone(ds, big)
    cfg = engine_config(path_slots=4096, max_jobs=64, shadow_per_vertex=4, iterations_per_poll=2, slice_order=order)
    small_tickets, done_big = [], None
    with Server(ctx, cfg) as srv:
        tb = srv.submit(ds, big, 0)
        t0 = time.time()
        while time.time() - t0 < 60.0:
            pending = srv.stats()["pending"]
            while pending < 8:
                small_tickets.append(srv.submit(ds, small, len(small_tickets) + 1))
                pending += 1
            out = srv.wait(tb, timeout_ms=0)
            if out is not None:
                done_big = out
                break
        assert done_big is not None, "the large render did not finish under the request stream"
        for t in small_tickets:
            assert srv.wait(t, timeout_ms=60000) is not None
    assert len(small_tickets) >= 16
    assert np.array_equal(done_big[0].cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.fixture
def counting(ctx):
    ctx.set_tuning(count_work=1)
    yield
    ctx.set_tuning(count_work=0)


def test_path_tracer_work_counters_off_by_default(ctx, resident):
    ds, p, key = job_list(resident)[0]
    assert ctx.get_tuning()["count_work"] == 0
    _, _, st = render_batch(ctx, [(ds, p, key)])
    assert st.primary_rays > 0 and (st.object_tests, st.object_hits, st.bounds_tests, st.bounds_hits) == (0, 0, 0, 0)


def test_path_tracer_work_counters(ctx, resident, counting):
    """With tuning.count_work the path tracer's TraceStats test counters are
    the device traversal's work (child boxes tested and passed, primitive
    tests and hits): per ray a pure function of the ray, so a batch's
    counters equal the sum of its jobs rendered alone, and a plain render's
    equal its one-job batch's; images do not change."""
    jobs = job_list(resident)[:3]
    ctx.set_tuning(count_work=0)
    ref = alone(jobs[0][0], jobs[0][1])  # not counted
    ctx.set_tuning(count_work=1)
    fields = ("primary_rays", "secondary_rays", "shadow_rays", "object_tests", "object_hits", "bounds_tests",
              "bounds_hits")
    singles = []
    for ds, p, key in jobs:
        _, _, st = render_batch(ctx, [(ds, p, key)])
        singles.append(st)
        assert st.object_tests > 0 and 0 < st.object_hits <= st.object_tests
        assert 0 < st.bounds_hits <= st.bounds_tests
        assert st.bounds_tests >= st.primary_rays
    _, _, stb = render_batch(ctx, jobs)
    for f in fields:
        assert getattr(stb, f) == sum(getattr(s, f) for s in singles), f
    ds, p, _ = jobs[0]
    acc, plain = render_device(ds, p)
    torch.cuda.synchronize()
    for f in fields:
        assert getattr(plain, f) == getattr(singles[0], f), f
    assert np.array_equal(acc.cpu().numpy().view(np.uint32), ref.view(np.uint32))
