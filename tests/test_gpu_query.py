"""GPU closest-hit / any-hit parity with the oracle (CPU restatement of
muon/acceleration.cc + muon/objects.cc), through the C ABI.

Bar, exact traversal mode: primitive ids, distance bits and any-hit flags
equal to the oracle's on every ray.
Bar, fast traversal mode: primitive ids equal except (i) distance ties
within 1e-5 relative — coincident surfaces (a face lying on another face)
reach 2e-6 because each distance is computed through its own transform —
and (ii) closer hits within 1e-6 (normalised barycentric) of a triangle
edge, which the reference's rounded slab test can reject; the GPU's id must
be a primitive the oracle's own test hits, and excused rays stay below 1e-4
of the batch (tests/gpu_util.py). Where ids agree, distance bits agree.
Any-hit: every ray occluded for the oracle is occluded on the GPU; extra
occlusions (edge grazes) below 1e-4 of the batch."""
import os

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


def _query(ds, o, d, tmax=None, exact=False):
    dev = torch.device("cuda:0")
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    tt = None if tmax is None else torch.from_numpy(tmax).to(dev)
    prim, t = ds.intersect_closest(to, td, tt, exact=exact)
    occ = ds.intersect_any(to, td, tt if tt is not None else torch.full((o.shape[0],), float("inf"), device=dev),
                           exact=exact)
    torch.cuda.synchronize()
    return prim.cpu().numpy(), t.cpu().numpy(), occ.cpu().numpy()


def _check(orc, o, d, prim, t, tmax=None, occ=None, max_tie_frac=1e-4):
    pr, tr, _, _, _ = orc.intersect(o, d)
    _, exact, hits = assert_closest_parity(orc, o, d, prim, t, pr, tr, max_tie_frac=max_tie_frac)
    if occ is not None:
        oref, _ = orc.occluded(o, d, tmax)
        assert_anyhit_parity(occ, oref)
    return exact, hits


# Scenes with coincident faces, where two primitives tie at the same distance
# and the visit order decides: the table legs of `transforms` end on the
# table top's underside, the boxes of `cornell_normals` rest on its floor.
COINCIDENT_TIE_FRAC = {"transforms": 2e-3, "cornell_normals": 2e-3}

SCENES = ["intersection", "transforms", "spheres_raytrace", "cornell_raytrace", "cornell_normals", "ggx", "mis",
          "sphere"]


@pytest.fixture(params=[(2, 0, 32), (4, 0, 32), (4, 0, 8), (2, 16, 32), (4, 16, 32), (4, 48, 8)],
                ids=["bvh2", "bvh4", "bvh4q", "bvh2-fetch", "bvh4-fetch", "bvh4q-fetch"])
def width(request, ctx):
    """Node format of the fast traversal mode (the exact mode is always
    binary): binary, 4-wide float boxes, 4-wide 8-bit boxes; with and
    without per-lane ray fetch."""
    old = ctx.get_tuning()
    w, f, b = request.param
    ctx.set_tuning(bvh_width=w, ray_fetch=f, node_bits=b)
    yield w
    ctx.set_tuning(bvh_width=old["bvh_width"], ray_fetch=old["ray_fetch"], node_bits=old["node_bits"])


@pytest.mark.parametrize("name", SCENES)
def test_reference_scene_queries(ctx, width, oracle, ref_dir, name):
    path = os.path.join(ref_dir, name + ".muon")
    hs = HostScene.from_file(path, default_seed=1)
    ds = DeviceScene(ctx, hs)
    orc = oracle.load(path, seed=1)
    rng = np.random.default_rng(11)
    lo, hi = scene_bounds(hs)
    o, d = random_rays(rng, 1 << 16, lo, hi)
    tmax = rng.uniform(0.1, 30.0, o.shape[0]).astype(np.float32)
    pr, tr, _, _, _ = orc.intersect(o, d)
    oref, _ = orc.occluded(o, d, tmax)
    assert (pr >= 0).sum() > 1000
    # exact mode: the reference's node visits, so ids and distance bits agree everywhere
    prim, t, _ = _query(ds, o, d, exact=True)
    assert np.array_equal(prim, pr)
    assert np.array_equal(t[pr >= 0].view(np.uint32), tr[pr >= 0].view(np.uint32))
    _, _, occ = _query(ds, o, d, tmax, exact=True)
    assert np.array_equal(occ, oref)
    # fast mode: equal outside ties / edge hits, distance bits equal where ids agree
    prim, t, _ = _query(ds, o, d)
    exact_frac, _ = _check(orc, o, d, prim, t, max_tie_frac=COINCIDENT_TIE_FRAC.get(name, 1e-4))
    assert exact_frac == 1.0
    _, _, occ = _query(ds, o, d, tmax)
    assert_anyhit_parity(occ, oref)


def test_camera_rays_bit_exact(ctx, oracle, ref_dir):
    """Primary rays of the reference's normals scene hit the same primitive at
    the same distance bit for bit (binary nodes visit children in the
    reference's order, so even the scene's coplanar faces tie the same way)."""
    old = ctx.get_tuning()["bvh_width"]
    ctx.set_tuning(bvh_width=2)
    path = os.path.join(ref_dir, "cornell_normals.muon")
    hs = HostScene.from_file(path)
    ds = DeviceScene(ctx, hs)
    orc = oracle.load(path)
    p = hs.params
    W, H = p.width, p.height
    xs = (np.arange(W, dtype=np.float32) + np.float32(0.5))
    ys = (np.arange(H, dtype=np.float32) + np.float32(0.5))
    X, Y = np.meshgrid(xs, ys)
    hw, hh = np.float32(W / 2.0), np.float32(H / 2.0)
    a = np.float32(p.tan_fov_x) * ((X - hw) / hw)
    b = np.float32(p.tan_fov_y) * ((hh - Y) / hh)
    u, v, w = [np.array(x, np.float32) for x in (p.cam_u, p.cam_v, p.cam_w)]
    dirs = (a[..., None] * u + b[..., None] * v) - w
    dirs = dirs.reshape(-1, 3).astype(np.float32)
    inv = np.float32(1.0) / np.sqrt(np.float32((dirs[:, 0] * dirs[:, 0] + dirs[:, 1] * dirs[:, 1]) + dirs[:, 2] * dirs[:, 2]))
    dirs = (dirs * inv[:, None]).astype(np.float32)
    org = np.tile(np.array(p.eye, np.float32), (dirs.shape[0], 1))
    try:
        prim, t, _ = _query(ds, org, dirs)
    finally:
        ctx.set_tuning(bvh_width=old)
    pr, tr, _, _, _ = orc.intersect(org, dirs)
    assert np.array_equal(prim, pr)
    assert np.array_equal(t.view(np.uint32), tr.view(np.uint32))


@pytest.fixture(scope="module")
def icosphere(ctx, oracle):
    text = scenes.icosphere_text()
    hs = HostScene.from_text(text)
    ds = DeviceScene(ctx, hs)
    orc = oracle.load_text(text)
    return hs, ds, orc


def test_icosphere_microbench_parity(icosphere, width):
    """Level-7 displaced icosphere (327,680 tris): the first 2^18 rays of the
    microbench ray sequence, closest-hit ids bit-exact vs the oracle."""
    hs, ds, orc = icosphere
    lo, hi = scenes.icosphere_world_bounds()
    o, d = scenes.microbench_rays(1 << 18, lo, hi, seed=0)
    prim, t, occ = _query(ds, o, d)
    exact, hits = _check(orc, o, d, prim, t, tmax=np.full(o.shape[0], np.inf, np.float32), occ=occ)
    assert hits > (1 << 17)
    assert exact > 0.9999


def test_icosphere_full_microbench_parity(icosphere):
    """All 2^24 microbench rays (the benchmark's batch) with the default
    launch configuration: closest-hit ids equal the oracle's on every ray
    outside ties and edge grazes (excused below 1e-4 of the rays), distance
    bits equal where ids agree; every hit a valid primitive at positive
    finite distance."""
    hs, ds, orc = icosphere
    lo, hi = scenes.icosphere_world_bounds()
    o, d = scenes.microbench_rays(1 << 24, lo, hi, seed=0)
    dev = torch.device("cuda:0")
    prim_t, t_t = ds.intersect_closest(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev))
    torch.cuda.synchronize()
    prim, t = prim_t.cpu().numpy(), t_t.cpu().numpy()
    del prim_t, t_t
    hit = prim >= 0
    assert hit.mean() > 0.5
    assert (prim[hit] < hs.desc.num_prims).all() and (t[hit] > 0).all() and np.isfinite(t[hit]).all()
    pr, tr, _, _, _ = orc.intersect(o, d, threads=16)
    excused, exact, hits = assert_closest_parity(orc, o, d, prim, t, pr, tr)
    print(f"2^24 rays: {hits} oracle hits, {int((prim != pr).sum())} id mismatches, {excused} excused, "
          f"distance bits equal on {exact:.6f} of agreeing hits")
    assert exact > 0.9999


def _stacked_leaf_text(n_stack, n_rand=200, seed=3):
    """n_stack copies of one triangle (identical bounds: the SAH build keeps
    them in one leaf) beside n_rand small random triangles."""
    rng = np.random.default_rng(seed)
    verts, tris = [], []
    for _ in range(n_stack):
        b = len(verts)
        verts += [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)]
        tris.append((b, b + 1, b + 2))
    for _ in range(n_rand):
        c = rng.uniform(-3, 3, 3)
        b = len(verts)
        verts += [tuple(float(x) for x in c + rng.uniform(-0.3, 0.3, 3)) for _ in range(3)]
        tris.append((b, b + 1, b + 2))
    lines = ["camera 0 0 5 0 0 0 0 1 0 45", f"maxverts {len(verts)}"]
    lines += ["vertex %r %r %r" % v for v in verts]
    lines += ["tri %d %d %d" % t for t in tris]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("n_stack", [100, 150], ids=["leaf100-counted", "leaf150-flagged"])
def test_large_leaf(ctx, width, oracle, n_stack):
    """A leaf of 100 primitives (the wide tree's leaf references carry counts
    up to 128) and one of 150 (leaf references fall back to the end-of-leaf
    flag): fast-mode ids equal the oracle's outside ties, any-hit flags equal
    but for edge grazes (fraction below 1e-4, as for the reference scenes)."""
    text = _stacked_leaf_text(n_stack)
    hs = HostScene.from_text(text)
    assert int(hs.arrays()["node_count"].max()) == n_stack
    ds = DeviceScene(ctx, hs)
    orc = oracle.load_text(text)
    rng = np.random.default_rng(17)
    lo, hi = scene_bounds(hs)
    o, d = random_rays(rng, 1 << 15, lo, hi)
    # half of the rays aimed at points of the stacked triangle
    k = o.shape[0] // 2
    uv = rng.random((k, 2), dtype=np.float32) * np.float32(0.5)
    tgt = np.stack([uv[:, 0], uv[:, 1], np.zeros(k, np.float32)], axis=1)
    dd = tgt - o[:k]
    d[:k] = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    tmax = np.full(o.shape[0], np.inf, np.float32)
    prim, t, occ = _query(ds, o, d, tmax)
    _, hits = _check(orc, o, d, prim, t)
    assert hits > 1000
    stack_hits = int(((prim >= 0) & (prim < n_stack)).sum())
    assert stack_hits > 1000
    oref, _ = orc.occluded(o, d, tmax)
    assert (occ != oref).mean() < 1e-4
