"""Pins the oracle (CPU restatement of the reference) against the reference's
own fixtures and against closed forms.

The reference's test suite (test/BUILD, test/diff_test.sh) has two modes:
exact (ImageMagick AE must be 0) and statistical (|MAE(out,truth) -
MAE(golden,truth)| <= tol). Both are restated here on the same scenes.
"""
import os

import numpy as np
import pytest
from PIL import Image

from _oracle import image_mae

EXACT = ["intersection", "transforms", "spheres_raytrace", "cornell_raytrace", "analytic",
         "cornell_normals", "cornell_depth", "cornell_mesh", "partial_emission"]
# test/BUILD:51-119
MAE_TOL = {"sphere": 5e-4, "cornell_phong": 2e-3, "cornell_brdf": 2e-3, "ggx": 2e-3, "mis": 2e-3,
           "partial_direct": 5e-4, "partial_indirect": 1e-3}


def _png(ref_dir, name):
    return np.asarray(Image.open(os.path.join(ref_dir, name)).convert("RGB"))


@pytest.mark.parametrize("name", EXACT)
def test_exact_golden(oracle, ref_dir, name):
    sc = oracle.load(os.path.join(ref_dir, name + ".muon"))
    _, rgb, _, _ = sc.render(parallelism=1)
    gold = _png(ref_dir, name + ".png")
    assert rgb.shape == gold.shape
    assert int((rgb != gold).any(axis=2).sum()) == 0


def test_random_seed_golden_bit_exact(oracle, ref_dir):
    """test/random_seed.muon: a seeded single-thread GGX path trace whose
    golden is an exact (AE=0) match of the whole CPU stack: mt19937 tile and
    path streams, libm transcendentals, NEE, RR and BRDF sampling."""
    sc = oracle.load(os.path.join(ref_dir, "random_seed.muon"))
    _, rgb, st, _ = sc.render(parallelism=1)
    gold = _png(ref_dir, "random_seed.png")
    assert int((rgb != gold).any(axis=2).sum()) == 0
    assert st[0] == 640 * 140 * 64  # primary rays


@pytest.mark.parametrize("name", sorted(MAE_TOL))
def test_statistical_golden(oracle, ref_dir, name):
    sc = oracle.load(os.path.join(ref_dir, name + ".muon"), seed=12345)
    _, rgb, _, _ = sc.render(parallelism=os.cpu_count() or 4)
    truth = _png(ref_dir, name + "_truth.png")
    gold = _png(ref_dir, name + "_golden.png")
    m, g = image_mae(rgb, truth), image_mae(gold, truth)
    assert (g - m) ** 2 <= MAE_TOL[name] ** 2, (m, g)


def test_mt19937_known_answer(oracle):
    # C++ [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937 (seed 5489) produces 4123659995.
    assert int(oracle.mt19937(5489, 10000)[-1]) == 4123659995


def test_uniform_random_semantics(oracle):
    # generate_canonical<float,24> over one 32-bit draw: float(u)/2^32,
    # clamped below 1 (libstdc++ bits/random.tcc).
    u = oracle.mt19937(77, 5000).astype(np.float64)
    f = oracle.uniform(77, 5000)
    exp = (u.astype(np.float32) * np.float32(1.0 / 4294967296.0)).astype(np.float32)
    exp = np.where(exp >= 1.0, np.nextafter(np.float32(1), np.float32(0)), exp)
    assert np.array_equal(f, exp)
    assert (f >= 0).all() and (f < 1).all()


def test_num_tiles_formula(oracle):
    # muon/sampling.cc:9-26: min(max(W*H*spp/50000, 3*parallelism), H)
    for w, h, spp, par in [(640, 140, 64, 1), (256, 256, 16, 8), (10, 5, 1, 4), (1024, 1024, 32, 16)]:
        assert oracle.num_tiles(w, h, spp, par) == min(max(w * h * spp // 50000, 3 * par), h)


TRI = """
camera 0 0 5  0 0 0  0 1 0  45
vertex 0 0 0
vertex 1 0 0
vertex 0 1 0
tri 0 1 2
"""


def test_tri_known_answer(oracle):
    sc = oracle.load_text(TRI)
    o = np.array([[0.25, 0.25, 1.0], [0.25, 0.25, -1.0], [2.0, 2.0, 1.0], [0.5, 0.5, 1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1]], np.float32)
    prim, t, pos, nrm, _ = sc.intersect(o, d)
    assert prim.tolist() == [0, -1, -1, 0]  # behind the origin misses; on-edge (0.5,0.5) hits
    assert t[0] == 1.0 and t[3] == 1.0
    assert np.array_equal(nrm[0], np.array([0, 0, 1], np.float32))


def test_tiny_triangle_is_invisible(oracle):
    # |n.d| < 1e-4 with the unnormalised normal (muon/objects.cc:104-108):
    # a triangle with 2*area < 1e-4 is never hit, even head-on.
    sc = oracle.load_text(TRI.replace("vertex 1 0 0", "vertex 0.005 0 0").replace("vertex 0 1 0", "vertex 0 0.005 0"))
    prim, _, _, _, _ = sc.intersect(np.array([[0.001, 0.001, 1.0]], np.float32), np.array([[0, 0, -1]], np.float32))
    assert prim[0] == -1
    sc2 = oracle.load_text(TRI.replace("vertex 1 0 0", "vertex 0.02 0 0").replace("vertex 0 1 0", "vertex 0 0.02 0"))
    prim, _, _, _, _ = sc2.intersect(np.array([[0.001, 0.001, 1.0]], np.float32), np.array([[0, 0, -1]], np.float32))
    assert prim[0] == 0


def test_sphere_known_answers(oracle):
    sc = oracle.load_text("camera 0 0 5 0 0 0 0 1 0 45\nsphere 0 0 0 1\n")
    o = np.array([[0, 0, 5], [0, 0, 0], [0, 1, 5], [0, 2, 5]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1]], np.float32)
    prim, t, _, _, _ = sc.intersect(o, d)
    assert prim.tolist() == [0, 0, 0, -1]
    assert t[0] == 4.0       # near root
    assert t[1] == 1.0       # starting inside picks the far root
    assert t[2] == 5.0       # tangent: discriminant 0
    # Non-uniform scale turns the sphere into an ellipsoid (world-space distance).
    sc2 = oracle.load_text("camera 0 0 5 0 0 0 0 1 0 45\nscale 2 1 1\nsphere 0 0 0 1\n")
    prim, t, pos, _, _ = sc2.intersect(np.array([[-10, 0, 0]], np.float32), np.array([[1, 0, 0]], np.float32))
    assert prim[0] == 0 and abs(t[0] - 8.0) < 1e-5 and abs(pos[0, 0] + 2.0) < 1e-5


def _random_scene(rng, n_tri=60, n_sph=30):
    lines = ["camera 0 0 30 0 0 0 0 1 0 45"]
    for _ in range(n_tri):
        c = rng.uniform(-5, 5, 3)
        for _ in range(3):
            v = c + rng.uniform(-1, 1, 3)
            lines.append("vertex %.6f %.6f %.6f" % tuple(v))
    for i in range(n_tri):
        lines.append("tri %d %d %d" % (3 * i, 3 * i + 1, 3 * i + 2))
    lines.append("push_transform")
    lines.append("rotate 0.3 1 0.2 33")
    lines.append("scale 1.5 0.7 1.2")
    for _ in range(n_sph):
        c = rng.uniform(-5, 5, 3)
        lines.append("sphere %.6f %.6f %.6f %.6f" % (c[0], c[1], c[2], rng.uniform(0.1, 0.8)))
    lines.append("pop_transform")
    return "\n".join(lines) + "\n"


def test_bvh_equals_brute_force(oracle):
    """BVH closest-hit equals the Linear structure (muon/acceleration.cc:14-45)
    up to exact distance ties, and any-hit agrees exactly."""
    rng = np.random.default_rng(3)
    text = _random_scene(rng)
    bvh = oracle.load_text(text, accel=1, partition=2)
    lin = oracle.load_text(text, accel=0)
    n = 20000
    o = rng.uniform(-8, 8, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pb, tb, _, _, _ = bvh.intersect(o, d)
    pl, tl, _, _, _ = lin.intersect(o, d)
    assert (pb >= 0).sum() > 1000
    same = pb == pl
    # any disagreement must be an exact tie in distance
    assert np.array_equal(tb[~same], tl[~same])
    assert same.mean() > 0.999
    tmax = rng.uniform(0.5, 10, n).astype(np.float32)
    ob, _ = bvh.occluded(o, d, tmax)
    ol, _ = lin.occluded(o, d, tmax)
    assert np.array_equal(ob, ol)


def test_partition_strategies_agree(oracle):
    rng = np.random.default_rng(5)
    text = _random_scene(rng, 40, 20)
    n = 5000
    o = rng.uniform(-8, 8, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ref, tref, _, _, _ = oracle.load_text(text, partition=2).intersect(o, d)
    for part in (0, 1):
        p, t, _, _, _ = oracle.load_text(text, partition=part).intersect(o, d)
        assert np.array_equal(t, tref)
