"""Pins for the parts of the oracle that the reference's golden images do not
fix: the SAH tree shape, the TraceStats counters, the edge measure that the
fast-mode parity rule excuses mismatches by, and that rule itself.

Each expected value below is derived by hand from the reference's algorithm
(the derivations are in the comments), not computed by the code under test:

* SAH build, muon/acceleration.cc:241-458: centroid bounds, widest axis, 12
  buckets by floor(offset * 12) (offset 1 clamped into the last bucket),
  left and right cumulative size x surface area, cost 0.125 + sum / total
  surface, first strict minimum, leaf when the primitive count is below the
  cost, std::partition (libstdc++'s bidirectional swap loop) by bucket <=
  the split bucket.
* Intersect / HasIntersection counters, muon/acceleration.cc:65-202: a
  bounds test per node popped or descended to, a bounds hit when the slab
  test (muon/bounds.cc:88-125, which skips axes with a zero direction
  component) passes with the current best distance, object tests per leaf
  primitive, object hits per primitive Intersect that returns a hit.
"""
import numpy as np
import pytest

from muon_amd import HostScene
from gpu_util import allowed, classify_mismatches

# ---------------------------------------------------------------- scenes
# Six spheres of radius 0.5 on the x axis, parsed in the order x = 12, 0,
# 16, 4, 2, 14 (ids 0..5). Every bound, surface area and cost below is an
# exact float.
SIX_X = [12.0, 0.0, 16.0, 4.0, 2.0, 14.0]


def six_spheres_text(extra=""):
    lines = ["film_size 1 1", extra] + [f"sphere {x:g} 0 0 0.5" for x in SIX_X]
    return "\n".join(l for l in lines if l) + "\n"


# Root: centroids x in [0, 16] (space 16, y and z extents 0: axis x); bucket
# = floor(x / 16 * 12) = floor(0.75 x): x 0,2,4,12,14,16 -> 0,1,3,9,10,11.
# A box of x-length a and y, z length 1 has surface 2(2a + 1) = 4a + 2.
# Left costs (size x surface) after buckets 0..10: 1*6, 2*14, 2*14, 3*22 x6,
# 4*54, 5*62; right costs of buckets > i: 5*62, 4*54, 4*54, 3*22 x6, 2*14,
# 1*6. Sums: 316, 244, 244, 132 (i = 3..8), 244, 316; the first minimum is
# bucket 3, cost 0.125 + 132/70 = 2.01 < 6 primitives: split x <= 4.
# libstdc++ std::partition on [12,0,16,4,2,14]: swap 12<->2, then 16<->4:
# [2,0,4 | 16,12,14].
# [2,0,4]: space 4, buckets 3x -> 6, 0, 11; every split costs 34 (ties), the
# first (bucket 0) wins with 0.125 + 34/22 < 3: partition -> [0 | 2,4];
# [2,4]: buckets 0, 11, all splits 12, 0.125 + 12/14 < 2: [2 | 4].
# [16,12,14]: space 4, buckets 3(x-12) -> 11, 0, 6: as above, split bucket 0:
# partition swaps 16<->12 -> [12 | 16,14]; [16,14]: buckets 11, 0: swap ->
# [14 | 16].
# Leaf order x = 0,2,4,12,14,16 = ids 1,4,3,0,5,2. Pre-order nodes:
#  0 internal [-0.5,16.5]     1 internal [-0.5,4.5]   2 leaf x=0
#  3 internal [1.5,4.5]       4 leaf x=2              5 leaf x=4
#  6 internal [11.5,16.5]     7 leaf x=12             8 internal [13.5,16.5]
#  9 leaf x=14               10 leaf x=16
SIX_TREE = dict(
    xmin=[-0.5, -0.5, -0.5, 1.5, 1.5, 3.5, 11.5, 11.5, 13.5, 13.5, 15.5],
    xmax=[16.5, 4.5, 0.5, 4.5, 2.5, 4.5, 16.5, 12.5, 16.5, 14.5, 16.5],
    left=[1, 2, -1, 4, -1, -1, 7, -1, 9, -1, -1],
    right=[6, 3, -1, 5, -1, -1, 8, -1, 10, -1, -1],
    count=[0, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1],
    start=[None, None, 0, None, 1, 2, None, 3, None, 4, 5],
    leaf_ids=[1, 4, 3, 0, 5, 2],
)

# Two spheres of radius 10 at x = 0 and 0.5: buckets 0 and 11, every split
# costs 2400 + 2400 over a total surface of 2(20.5*20*2 + 400) = 2440, so
# 0.125 + 4800/2440 = 2.09 > 2 primitives: the root is one leaf of both, in
# parse order.
OVERLAP_TEXT = "film_size 1 1\nsphere 0 0 0 10\nsphere 0.5 0 0 10\n"
# Three spheres with one centre: centroid space 0, every offset 0, so all
# left costs are 3 * 216 (the radius-3 box) and no right cost is added:
# 0.125 + 648/216 = 3.125 > 3: one leaf of three.
CONCENTRIC_TEXT = "film_size 1 1\nsphere 0 0 0 1\nsphere 0 0 0 2\nsphere 0 0 0 3\n"


def _host_tree(text):
    a = HostScene.from_text(text).arrays()
    return a


def _oracle_tree(oracle, text):
    return oracle.load_text(text).bvh()


def test_sah_tree_six_spheres_host(oracle):
    a = _host_tree(six_spheres_text())
    t = SIX_TREE
    nb = a["node_bounds"]
    assert nb.shape == (11, 6)
    assert np.array_equal(nb[:, 0], np.float32(t["xmin"])) and np.array_equal(nb[:, 3], np.float32(t["xmax"]))
    assert (nb[:, [1, 2]] == -0.5).all() and (nb[:, [4, 5]] == 0.5).all()
    assert list(a["node_left"]) == t["left"] and list(a["node_right"]) == t["right"]
    assert list(a["node_count"]) == t["count"]
    for i, s in enumerate(t["start"]):
        if s is not None:
            assert a["node_start"][i] == s
        else:
            assert a["node_axis"][i] == 0  # every split is on x
    assert list(a["leaf_prims"]) == t["leaf_ids"]


def test_sah_tree_six_spheres_oracle(oracle):
    b = _oracle_tree(oracle, six_spheres_text())
    t = SIX_TREE
    assert b["bounds"].shape == (11, 6)
    assert np.array_equal(b["bounds"][:, 0], np.float32(t["xmin"]))
    assert np.array_equal(b["bounds"][:, 3], np.float32(t["xmax"]))
    assert list(b["left"]) == t["left"] and list(b["right"]) == t["right"]
    assert list(b["count"]) == t["count"]
    for i, s in enumerate(t["start"]):
        if s is not None:
            assert b["start"][i] == s
        else:
            assert b["axis"][i] == 0
    assert list(b["prim_order"]) == t["leaf_ids"]


@pytest.mark.parametrize("text,n", [(OVERLAP_TEXT, 2), (CONCENTRIC_TEXT, 3)], ids=["overlap", "concentric"])
def test_sah_leaf_decisions(oracle, text, n):
    a = _host_tree(text)
    assert list(a["node_count"]) == [n] and list(a["node_start"]) == [0]
    assert list(a["leaf_prims"]) == list(range(n))
    b = _oracle_tree(oracle, text)
    assert list(b["count"]) == [n] and list(b["prim_order"]) == list(range(n))


# ---------------------------------------------------------------- TraceStats
# Hand traces over the six-sphere tree (x splits; children[dir.x < 0] first).
# (origin, direction, closest id or -1, closest distance,
#  (bounds tests, bounds hits, object tests, object hits))
#  r1 from x=-5 along +x: root, n1, leaf x=0 (hit at 4.5), then n3 (t 6.5 >
#     4.5) and n6 (16.5) rejected: 5 tests, 3 hits, 1 object test and hit.
#  r2 from x=20 along -x: root, n6, n8, leaf 16 (hit at 3.5), leaf 14 (5.5),
#     leaf 12 (7.5), n1 (15.5) rejected: 7 tests, 4 hits.
#  r3 from (5,5,0) along +y: the root's y slab lies behind the origin: 1 test.
#  r4 from (2.45,5,0.45) along -y: the x and z direction components are zero,
#     so every box's x and z slabs are skipped and all 11 nodes pass; the ray
#     misses all six spheres (it passes 0.64 from the centre of the one at
#     x=2): 11 tests, 11 hits, 6 object tests, no hit.
TRACE_RAYS = [
    ((-5.0, 0.0, 0.0), (1.0, 0.0, 0.0), 1, 4.5, (5, 3, 1, 1)),
    ((20.0, 0.0, 0.0), (-1.0, 0.0, 0.0), 2, 3.5, (7, 4, 1, 1)),
    ((5.0, 5.0, 0.0), (0.0, 1.0, 0.0), -1, None, (1, 0, 0, 0)),
    ((2.45, 5.0, 0.45), (0.0, -1.0, 0.0), -1, None, (11, 11, 6, 0)),
]
# HasIntersection with max_distance: r1 with no limit stops at its first
# object hit (3 tests, 3 hits, 1, 1); r1 with max_distance 4 rejects the root
# (its entry 4.5 is not below 4); r4 walks everything as above.
ANY_RAYS = [(0, np.inf, 1, (3, 3, 1, 1)), (0, 4.0, 0, (1, 0, 0, 0)), (3, np.inf, 0, (11, 11, 6, 0))]


def _stats(st):
    # oracle layout: primary, secondary, object tests, object hits, bounds tests, bounds hits, shadow
    return (int(st[4]), int(st[5]), int(st[2]), int(st[3]))


@pytest.mark.parametrize("k", range(len(TRACE_RAYS)))
def test_trace_stats_closest_hand_counted(oracle, k):
    o, d, pid, dist, counts = TRACE_RAYS[k]
    sc = oracle.load_text(six_spheres_text())
    prim, t, _, _, st = sc.intersect(np.float32([o]), np.float32([d]), threads=1)
    assert int(prim[0]) == pid
    if dist is not None:
        assert float(t[0]) == dist
    assert _stats(st) == counts


@pytest.mark.parametrize("k", range(len(ANY_RAYS)))
def test_trace_stats_any_hand_counted(oracle, k):
    ray, tmax, occ_expect, counts = ANY_RAYS[k]
    o, d = TRACE_RAYS[ray][:2]
    sc = oracle.load_text(six_spheres_text())
    occ, st = sc.occluded(np.float32([o]), np.float32([d]), np.float32([tmax]), threads=1)
    assert int(occ[0]) == occ_expect
    assert _stats(st) == counts


# ---------------------------------------------------------------- edge measure
# Triangle a=(0,0,0) b=(1,0,0) c=(0,1,0), normal (0,0,1), |n|^2 = 1: a ray
# down -z through (px, py) has normalised barycentrics (px, py, 1-px-py);
# the second copy is drawn under `scale 2 2 2`, so the world point (2px, 2py)
# has the same barycentrics. A sphere reports 1.
EDGE_TEXT = ("film_size 1 1\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1 0\ntri 0 1 2\n"
             "push_transform\nscale 2 2 2\ntranslate 0 0 -3\ntri 0 1 2\npop_transform\nsphere 5 5 5 1\n")


@pytest.mark.parametrize("px,py,expect", [(0.2, 0.3, 0.2), (0.25, 0.5, 0.25), (0.5, 0.0, 0.0),
                                          (0.7, 0.6, -0.3), (0.1, 0.1, 0.1)])
def test_prim_edge_closed_form(oracle, px, py, expect):
    sc = oracle.load_text(EDGE_TEXT)
    m = sc.prim_edge(0, np.float32([px, py, 1.0]), np.float32([0, 0, -1]))
    assert m == pytest.approx(expect, abs=2e-7)
    m2 = sc.prim_edge(1, np.float32([2 * px, 2 * py, 1.0]), np.float32([0, 0, -1]))
    assert m2 == pytest.approx(expect, abs=2e-7)
    if expect >= 0:
        assert sc.prim_distance(0, np.float32([px, py, 1.0]), np.float32([0, 0, -1])) == pytest.approx(1.0)
        assert sc.prim_distance(1, np.float32([2 * px, 2 * py, 1.0]), np.float32([0, 0, -1])) == pytest.approx(7.0)
    else:
        assert sc.prim_distance(0, np.float32([px, py, 1.0]), np.float32([0, 0, -1])) is None


def test_prim_edge_failures(oracle):
    sc = oracle.load_text(EDGE_TEXT)
    o, down = np.float32([0.2, 0.2, 1.0]), np.float32([0, 0, -1])
    assert sc.prim_edge(2, o, down) == 1.0             # sphere
    assert sc.prim_edge(3, o, down) is None            # unknown id
    assert sc.prim_edge(-1, o, down) is None
    assert sc.prim_edge(0, o, np.float32([1, 0, 0])) is None  # parallel to the plane
    assert sc.prim_distance(3, o, down) is None


# ---------------------------------------------------------------- the parity rule
def _two_tri_batch(oracle):
    """Two triangles side by side (ids 0, 1) and rays down -z; the oracle's
    answers are the reference ones."""
    text = ("film_size 1 1\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1 0\nvertex 1 1 0\n"
            "tri 0 1 2\ntri 1 3 2\nsphere 5 5 -5 1\n")
    sc = oracle.load_text(text)
    rng = np.random.default_rng(0)
    xy = rng.uniform(0.05, 0.95, (64, 2)).astype(np.float32)
    o = np.concatenate([xy, np.ones((64, 1), np.float32)], axis=1)
    d = np.tile(np.float32([0, 0, -1]), (64, 1))
    pr, tr, _, _, _ = sc.intersect(o, d, threads=1)
    assert (pr >= 0).all()
    return sc, o, d, pr, tr


def test_classifier_accepts_identical_results(oracle):
    sc, o, d, pr, tr = _two_tri_batch(oracle)
    assert classify_mismatches(sc, o, d, pr.copy(), tr.copy(), pr, tr) == ([], [], [])


def test_classifier_flags_planted_wrong_ids(oracle):
    """A wrong existing id (the other triangle, which the ray misses by a
    wide margin), an unknown id, an out-of-range negative id, a miss where
    the oracle hits, and the sphere the ray never reaches are all bad."""
    sc, o, d, pr, tr = _two_tri_batch(oracle)
    g = pr.copy()
    planted = {0: 1 - int(pr[0]), 1: 999, 2: -5, 3: -1, 4: 2}
    for i, v in planted.items():
        g[i] = v
    bad, ties, grazes = classify_mismatches(sc, o, d, g, tr.copy(), pr, tr)
    assert sorted(bad) == sorted(planted) and ties == [] and grazes == []


def test_classifier_excuses_edge_graze_and_tie_only(oracle):
    """A ray through the shared edge x + y = 1 hits both triangles at the
    same distance: either id is a tie. A ray that grazes triangle 0's outer
    edge x = 0 from inside, reported as a closer hit where the oracle found
    nothing, is an edge graze; the same report 0.3 inside the triangle is
    bad."""
    text = ("film_size 1 1\nvertex 0 0 0\nvertex 1 0 0\nvertex 0 1 0\nvertex 1 1 0\n"
            "tri 0 1 2\ntri 1 3 2\n")
    sc = oracle.load_text(text)
    o = np.float32([[0.5, 0.5, 1.0], [0.0, 0.5, 1.0], [0.3, 0.5, 1.0]])
    d = np.tile(np.float32([0, 0, -1]), (3, 1))
    pr, tr, _, _, _ = sc.intersect(o, d, threads=1)
    assert pr[0] in (0, 1)
    g = pr.copy()
    g[0] = 1 - pr[0]
    ref_p, ref_t = pr.copy(), tr.copy()
    ref_p[1:] = -1  # pretend the reference's slab test rejected both
    g[1:] = 0
    bad, ties, grazes = classify_mismatches(sc, o, d, g, tr.copy(), ref_p, ref_t)
    assert ties == [0] and grazes == [1] and bad == [2]


def test_allowed_excused_count():
    assert allowed(1e-4, 1 << 24) == 1677
    assert allowed(1e-4, 4099) == 1


# ---------------------------------------------------------------- Lambertian BRDF
def test_oracle_lambertian_equals_phong_without_specular(oracle):
    """muon/materials.cc:17-47: the Lambertian BRDF is kd / pi with cosine
    sampling, which ignores the specular colour; Phong with a zero specular
    colour (reflectiveness 0, muon/materials.cc:113-123) is the same BRDF and
    sampling, with one more random draw per sample. So the oracle's
    Lambertian render with specular 0.5 and its Phong render with specular 0
    estimate the same image: their means over 32 seeds agree within 5
    standard errors, overall and per 2 x 2 block."""
    import importlib.util
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("mk", os.path.join(here, "golden", "make_lambertian_fixture.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    from muon_amd import scenes
    n = 32
    tiles = oracle.num_tiles(mk.W, mk.H, mk.SPP, mk.PAR)
    skip = mk.H // tiles + (mk.H - (mk.H // tiles) * tiles)
    lam = np.array([mk.block_means(oracle.load_text_lambertian(mk.scene_text(s)).render(mk.PAR)[0] / mk.SPP, skip)
                    for s in range(n)])
    phong = np.array([mk.block_means(oracle.load_text(scenes.cornell_text(mk.W, mk.H, mk.SPP, mk.DEPTH, seed=100 + s))
                                     .render(mk.PAR)[0] / mk.SPP, skip) for s in range(n)])
    z = (lam.mean(0) - phong.mean(0)) / np.sqrt(lam.var(0) / n + phong.var(0) / n)
    assert np.abs(z).max() < 5.0, z
    # and the same scene's Phong BRDF (specular 0.5 kept) is far brighter
    assert oracle.load_text(mk.scene_text(0)).render(mk.PAR)[0].mean() > 2 * lam.mean() * mk.SPP
