"""Scenes uploaded without a host-built tree: the device builds a linear
BVH (Morton-ordered primitives, binary radix tree, bottom-up boxes) and the
scene is traversed as a host-built one in the fast mode.

Bar: the fast-mode parity bar of the host-built trees against the oracle
(tests/test_gpu_query.py): ids equal outside ties and edge grazes, any-hit
flags never missing an occluder; the exact mode and the integrators that
need the reference's tree are refused; path-traced images equal the
host-tree scene's except where distance ties resolve differently."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from muon_amd import Context, DeviceScene, HostScene, MuonxError, scenes  # noqa: E402
from muon_amd.renderer import render_device  # noqa: E402
from gpu_util import assert_anyhit_parity, assert_closest_parity, random_rays, scene_bounds  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return Context(0)


def _query(ds, o, d, tmax):
    dev = torch.device("cuda:0")
    to, td, tt = (torch.from_numpy(x).to(dev) for x in (o, d, tmax))
    prim, t = ds.intersect_closest(to, td)
    occ = ds.intersect_any(to, td, tt)
    torch.cuda.synchronize()
    return prim.cpu().numpy(), t.cpu().numpy(), occ.cpu().numpy()


# scenes with coincident faces tie more often (see tests/test_gpu_query.py)
TIE_FRAC = {"transforms": 2e-3, "cornell_normals": 2e-3}


@pytest.mark.parametrize("name", ["intersection", "transforms", "spheres_raytrace", "cornell_normals", "ggx",
                                  "dragon"])
def test_device_tree_queries_match_the_oracle(ctx, oracle, ref_dir, tmp_path, name):
    path = os.path.join(ref_dir, name + ".muon")
    if name == "dragon":
        import gzip
        path = str(tmp_path / "dragon.muon")
        with gzip.open(os.path.join(ref_dir, "dragon.muon.gz"), "rb") as f:
            open(path, "wb").write(f.read())
    hs = HostScene.from_file(path, default_seed=1, device_bvh=True)
    assert hs.desc.num_nodes == 0
    ds = DeviceScene(ctx, hs)
    info = ds.info()
    assert info["device_built"] == 1 and info["build_ms"] > 0 and info["num_prims"] == hs.desc.num_prims
    orc = oracle.load(path, seed=1)
    lo, hi = scene_bounds(HostScene.from_file(path, default_seed=1))
    rng = np.random.default_rng(23)
    o, d = random_rays(rng, 1 << 16, lo, hi)
    tmax = rng.uniform(0.1, 30.0, o.shape[0]).astype(np.float32)
    prim, t, occ = _query(ds, o, d, tmax)
    pr, tr, _, _, _ = orc.intersect(o, d)
    assert (pr >= 0).sum() > 1000
    assert_closest_parity(orc, o, d, prim, t, pr, tr, max_tie_frac=TIE_FRAC.get(name, 1e-4))
    oref, _ = orc.occluded(o, d, tmax)
    assert_anyhit_parity(occ, oref)


def test_device_tree_icosphere_microbench(ctx, oracle):
    """The level-7 icosphere (327,692 primitives) built on the device: the
    first 2^20 microbench rays match the oracle."""
    text = scenes.icosphere_text()
    ds = DeviceScene(ctx, HostScene.from_text(text, device_bvh=True))
    orc = oracle.load_text(text)
    lo, hi = scenes.icosphere_world_bounds()
    o, d = scenes.microbench_rays(1 << 20, lo, hi, seed=0)
    prim, t, occ = _query(ds, o, d, np.full(o.shape[0], np.inf, np.float32))
    pr, tr, _, _, _ = orc.intersect(o, d, threads=16)
    assert_closest_parity(orc, o, d, prim, t, pr, tr)
    oref, _ = orc.occluded(o, d, np.full(o.shape[0], np.inf, np.float32), threads=16)
    assert_anyhit_parity(occ, oref)


def test_one_primitive_scene(ctx, oracle):
    text = "camera 0 0 5 0 0 0 0 1 0 45\nsphere 0 0 -1 0.8\n"
    ds = DeviceScene(ctx, HostScene.from_text(text, device_bvh=True))
    orc = oracle.load_text(text)
    rng = np.random.default_rng(2)
    o, d = random_rays(rng, 4096, (-1.5, -1.5, -2.0), (1.5, 1.5, 1.5))
    prim, t, occ = _query(ds, o, d, np.full(o.shape[0], np.inf, np.float32))
    pr, tr, _, _, _ = orc.intersect(o, d)
    assert (pr >= 0).sum() > 100
    assert_closest_parity(orc, o, d, prim, t, pr, tr)


def test_exact_mode_and_reference_integrators_refused(ctx, ref_dir):
    hs = HostScene.from_file(os.path.join(ref_dir, "cornell_normals.muon"), device_bvh=True)
    ds = DeviceScene(ctx, hs)
    o = torch.zeros((4, 3), device="cuda:0")
    d = torch.tensor([[0.0, 0.0, -1.0]] * 4, device="cuda:0")
    with pytest.raises(MuonxError) as e:
        ds.intersect_closest(o, d, exact=True)
    assert e.value.code == -5
    with pytest.raises(MuonxError) as e:
        render_device(ds, hs.params)  # the normals integrator
    assert e.value.code == -5


def test_path_traced_images_match_the_host_tree(ctx):
    """The Cornell box with a GGX icosphere (level 4): the same render keys
    give the same paths on either tree except where a ray's closest hit is a
    distance tie, so the images are nearly identical."""
    text = scenes.icosphere_text(64, 64, 16, 8, level=4)
    a = DeviceScene(ctx, HostScene.from_text(text))
    hb = HostScene.from_text(text, device_bvh=True)
    b = DeviceScene(ctx, hb)
    acc_a, st_a = render_device(a, hb.params)
    acc_b, st_b = render_device(b, hb.params)
    torch.cuda.synchronize()
    x, y = acc_a.cpu().numpy() / 16, acc_b.cpu().numpy() / 16
    assert st_a.primary_rays == st_b.primary_rays
    assert abs(x.mean() - y.mean()) < 1e-3 * x.mean()
    same = np.mean(np.all(x == y, axis=2))
    assert same > 0.95, same
