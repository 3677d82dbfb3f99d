"""muonx_intersect_noalloc (static device scratch, no allocation) against
muonx_intersect on the same rays, through the C ABI.

Bar: bit-exact. Primitive ids, distance bits and any-hit flags equal those of
muonx_intersect, which the parity tests check against the oracle; the two
calls differ only in where the traversal stacks spill and the claim counter
live, and in the grid size. Two calls in a row, the second with a smaller
(ragged) batch, so the second reads scratch the first left behind."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene, scenes  # noqa: E402


@pytest.fixture(scope="module")
def ico():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    ctx = Context(0)
    hs = HostScene.from_text(scenes.icosphere_text())
    return ctx, DeviceScene(ctx, hs)


# (bvh_width, stack_entries, ray_fetch, node_bits): the default quantized
# 4-wide walk with per-lane fetch, binary nodes with an 8-entry LDS ring (the
# binary tree is deeper than 8, so stacks spill into the scratch), and 4-wide
# float nodes without per-lane fetch
TUNINGS = {"bvh4q-fetch": (4, 8, 48, 8), "bvh2-ring8-spill": (2, 8, 48, 32), "bvh4-gridstride": (4, 16, 0, 32)}


def _run(ds, o, d, tmax, any_hit, exact, noalloc):
    dev = torch.device("cuda:0")
    n = o.shape[0]
    prim = torch.full((n,), -7, dtype=torch.int32, device=dev)
    t = torch.full((n,), -7.0, dtype=torch.float32, device=dev)
    occ = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    if any_hit:
        ds.query(o, d, tmax, True, exact, occ=occ, noalloc=noalloc)
        torch.cuda.synchronize()
        return occ.cpu().numpy()
    ds.query(o, d, tmax, False, exact, prim=prim, t=t, noalloc=noalloc)
    torch.cuda.synchronize()
    return np.stack([prim.cpu().numpy().view(np.uint32), t.cpu().numpy().view(np.uint32)])


@pytest.mark.parametrize("tuning", list(TUNINGS), ids=list(TUNINGS))
def test_noalloc_equals_intersect(ico, tuning):
    ctx, ds = ico
    old = ctx.get_tuning()
    w, se, f, nb = TUNINGS[tuning]
    ctx.set_tuning(bvh_width=w, stack_entries=se, ray_fetch=f, node_bits=nb)
    try:
        lo, hi = scenes.icosphere_world_bounds()
        dev = torch.device("cuda:0")
        for n, seed in ((1 << 20, 3), ((1 << 12) + 37, 4)):
            o_np, d_np = scenes.microbench_rays(n, lo, hi, seed=seed)
            rng = np.random.default_rng(seed)
            tm_np = rng.uniform(0.01, 2.0, n).astype(np.float32)
            o = torch.from_numpy(o_np).to(dev)
            d = torch.from_numpy(d_np).to(dev)
            tm = torch.from_numpy(tm_np).to(dev)
            for any_hit in (False, True):
                for exact in (False, True):
                    tmax = tm if any_hit else None
                    ref = _run(ds, o, d, tmax, any_hit, exact, noalloc=False)
                    got = _run(ds, o, d, tmax, any_hit, exact, noalloc=True)
                    assert np.array_equal(got, ref), (n, any_hit, exact)
                    if not any_hit:
                        assert (ref[0].view(np.int32) >= 0).mean() > 0.5
    finally:
        ctx.set_tuning(**{k: old[k] for k in ("bvh_width", "stack_entries", "ray_fetch", "node_bits")})


def test_noalloc_rejects_bad_arguments(ico):
    import ctypes as C
    from muon_amd import _lib
    ctx, ds = ico
    q = _lib.Query(None, None, None, 16, 0, 0, None, None, None, None)
    assert _lib.lib.muonx_intersect_noalloc(ds.h, C.byref(q), None) == -1
    assert b"missing ray" in _lib.lib.muonx_last_error()
    assert _lib.lib.muonx_intersect_noalloc(None, C.byref(q), None) == -1
