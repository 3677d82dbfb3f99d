"""Edge cases of the batched queries (muonx_intersect) against the oracle:
empty and ragged batch sizes, axis-aligned rays (zero direction components,
the reference's slab-test skip, muon/bounds.cc:98-100), tmax limits on both
query kinds, and a scene whose BVH is a single leaf. Exact mode: ids and
distance bits equal the oracle's; fast mode: ids equal outside distance ties
and edge grazes, for every node format and ray-fetch setting."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene, scenes  # noqa: E402
from gpu_util import assert_anyhit_parity, assert_closest_parity, random_rays, scene_bounds  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return Context(0)


@pytest.fixture(params=[(2, 0, 32), (4, 0, 8), (4, 48, 8), (4, 48, 32), (2, 16, 32)],
                ids=["bvh2", "bvh4q", "bvh4q-fetch", "bvh4-fetch", "bvh2-fetch"])
def width(request, ctx):
    old = ctx.get_tuning()
    w, f, b = request.param
    ctx.set_tuning(bvh_width=w, ray_fetch=f, node_bits=b)
    yield w
    ctx.set_tuning(bvh_width=old["bvh_width"], ray_fetch=old["ray_fetch"], node_bits=old["node_bits"])


@pytest.fixture(scope="module")
def cornell(ctx, oracle):
    text = scenes.cornell_text(32, 32, 1, 5)
    hs = HostScene.from_text(text)
    return hs, DeviceScene(ctx, hs), oracle.load_text(text)


def _run(ds, o, d, tmax=None, exact=False):
    dev = torch.device("cuda:0")
    n = o.shape[0]
    to = torch.from_numpy(np.ascontiguousarray(o)).to(dev)
    td = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
    tt = None if tmax is None else torch.from_numpy(np.ascontiguousarray(tmax)).to(dev)
    # outputs pre-filled with a sentinel: every entry must be written
    prim = torch.full((n,), -7, dtype=torch.int32, device=dev)
    t = torch.full((n,), -7.0, dtype=torch.float32, device=dev)
    ds.intersect_closest(to, td, tt, prim=prim, t=t, exact=exact)
    tm = tt if tt is not None else torch.full((n,), float("inf"), device=dev)
    occ = torch.full((n,), 7, dtype=torch.uint8, device=dev)
    ds.intersect_any(to, td, tm, occ=occ, exact=exact)
    torch.cuda.synchronize()
    return prim.cpu().numpy(), t.cpu().numpy(), occ.cpu().numpy()


# The Cornell box's two boxes rest on its floor: their bottom faces and the
# floor tie for rays that start inside a box and leave through its bottom
# (about 1.1 % of rays spread uniformly through the room, counted with the
# oracle), and the visit order decides which one is reported.
CORNELL_TIE_FRAC = 2e-2


def _check_closest(orc, o, d, prim, t, pr, tr, exact, max_graze_frac=1e-4):
    if exact:
        assert np.array_equal(prim, pr)
        assert np.array_equal(t[pr >= 0].view(np.uint32), tr[pr >= 0].view(np.uint32))
    else:
        assert_closest_parity(orc, o, d, prim, t, pr, tr, max_tie_frac=CORNELL_TIE_FRAC,
                              max_graze_frac=max_graze_frac)


@pytest.mark.parametrize("n", [0, 1, 63, 65, 257, 4099])
@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
def test_ragged_batch_sizes(cornell, width, n, exact):
    hs, ds, orc = cornell
    rng = np.random.default_rng(n)
    lo, hi = scene_bounds(hs)
    o, d = random_rays(rng, n, lo, hi)
    prim, t, occ = _run(ds, o, d, exact=exact)
    assert prim.shape == (n,)
    if n == 0:
        return
    pr, tr, _, _, _ = orc.intersect(o, d)
    _check_closest(orc, o, d, prim, t, pr, tr, exact)
    oref, _ = orc.occluded(o, d, np.full(n, np.inf, np.float32))
    if exact:
        assert np.array_equal(occ, oref)
    else:
        assert_anyhit_parity(occ, oref)


@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
def test_axis_aligned_rays(cornell, width, exact):
    """Directions along +-x, +-y, +-z (two zero components) and in the
    coordinate planes (one zero component), from a grid of origins inside the
    box, including origins on the planes of the walls' boxes."""
    hs, ds, orc = cornell
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (0.6, 0.8, 0), (0, -0.6, 0.8), (0.8, 0, -0.6)]
    g = np.linspace(-0.95, 0.95, 9, dtype=np.float32)
    oy = np.linspace(0.05, 1.95, 9, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, oy, g, indexing="ij")
    org = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    org = np.concatenate([org, np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 2.0, 0.0], [-1.0, 1.0, -1.0]],
                                        np.float32)])
    o = np.repeat(org, len(dirs), axis=0).astype(np.float32)
    d = np.tile(np.array(dirs, np.float32), (org.shape[0], 1))
    prim, t, occ = _run(ds, o, d, exact=exact)
    pr, tr, _, _, _ = orc.intersect(o, d)
    assert (pr >= 0).mean() > 0.5
    # rays running inside wall planes from origins on them end on the walls'
    # edges: grazes are the point of this test, not a rare event
    _check_closest(orc, o, d, prim, t, pr, tr, exact, max_graze_frac=1e-2)
    oref, _ = orc.occluded(o, d, np.full(o.shape[0], np.inf, np.float32))
    if exact:
        bad = np.nonzero(occ != oref)[0]
        info = [(o[i].tolist(), d[i].tolist(), int(occ[i]), int(oref[i]), int(pr[i]), float(tr[i]), int(prim[i]))
                for i in bad[:6]]
        assert len(bad) == 0, (len(bad), info)
    else:
        # origins on the walls' planes graze boxes more often than random rays
        assert_anyhit_parity(occ, oref, max_frac=1e-3)


@pytest.mark.parametrize("exact", [True, False], ids=["exact", "fast"])
def test_tmax_limits(cornell, width, exact):
    """tmax of 0, negative, below and above each ray's hit distance: the
    closest hit is kept exactly when its distance is below tmax; any-hit
    follows HasIntersection (muon/acceleration.cc:144-202)."""
    hs, ds, orc = cornell
    rng = np.random.default_rng(3)
    lo, hi = scene_bounds(hs)
    o, d = random_rays(rng, 4096, lo, hi)
    pr, tr, _, _, _ = orc.intersect(o, d)
    hit = pr >= 0
    tmax = np.full(o.shape[0], np.inf, np.float32)
    q = rng.integers(0, 4, o.shape[0])
    tmax[q == 0] = 0.0
    tmax[q == 1] = -1.0
    tmax[(q == 2) & hit] = tr[(q == 2) & hit] * np.float32(0.5)
    tmax[(q == 3) & hit] = tr[(q == 3) & hit] * np.float32(2.0)
    prim, t, occ = _run(ds, o, d, tmax, exact=exact)
    expect = np.where(hit & (tr < tmax), pr, -1)
    assert (prim[tmax <= 0] == -1).all() and (occ[tmax <= 0] == 0).all()
    _check_closest(orc, o, d, prim, t, expect, tr, exact)
    oref, _ = orc.occluded(o, d, tmax)
    if exact:
        assert np.array_equal(occ, oref)
    else:
        assert_anyhit_parity(occ, oref, max_frac=1e-3)


@pytest.mark.parametrize("accel,geom", [(0, "tris+sphere"), (1, "one sphere")], ids=["linear", "one-prim"])
def test_single_leaf_scene(ctx, oracle, width, accel, geom):
    """BVHs whose root is one leaf: the linear structure (--acceleration=
    linear: every primitive in one leaf, muon/acceleration.cc:14-45) over two
    triangles and a sphere, and a one-sphere scene."""
    text = "camera 0 0 5 0 0 0 0 1 0 45\n"
    if geom == "one sphere":
        text += "sphere 0 0 -1 0.8\n"
    else:
        text += ("vertex -1 -1 0\nvertex 1 -1 0\nvertex 1 1 0\nvertex -1 1 0\ntri 0 1 2\ntri 0 2 3\n"
                 "sphere 0 0 -1 0.5\n")
    hs = HostScene.from_text(text, acceleration=accel)
    assert hs.desc.num_nodes == 1
    ds = DeviceScene(ctx, hs)
    orc = oracle.load_text(text, accel=accel)
    rng = np.random.default_rng(9)
    o, d = random_rays(rng, 2048, (-1.5, -1.5, -2.0), (1.5, 1.5, 1.5))
    pr, tr, _, _, _ = orc.intersect(o, d)
    assert (pr >= 0).sum() > 100
    for exact in (True, False):
        prim, t, occ = _run(ds, o, d, exact=exact)
        _check_closest(orc, o, d, prim, t, pr, tr, exact)
