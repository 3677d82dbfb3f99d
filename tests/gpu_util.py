"""Helpers shared by the GPU tests."""
import numpy as np


def random_rays(rng, n, lo, hi, pad=0.5):
    lo = np.asarray(lo, np.float32) - pad
    hi = np.asarray(hi, np.float32) + pad
    o = (lo + rng.random((n, 3), dtype=np.float32) * (hi - lo)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    return o, d.astype(np.float32)


def scene_bounds(host_scene):
    b = host_scene.arrays()["node_bounds"]
    return b[0, :3], b[0, 3:]


def allowed(frac, n):
    """Largest excused count for a batch of n rays: a fraction `frac` of the
    rays, and at least one ray (small batches)."""
    return max(1, int(frac * n))


def classify_mismatches(orc_scene, o, d, prim_gpu, t_gpu, prim_ref, t_ref, rel=1e-6, edge=1e-6):
    """Splits the rays whose closest-hit ids differ (GPU fast mode vs the
    oracle) into (bad, ties, grazes) index lists.

    The fast mode runs the reference's primitive tests op for op and accepts
    a superset of the boxes the reference's slab test accepts, so it can
    differ from the reference only by (i) a distance tie, resolved by
    another visit order, or (ii) a closer hit on a primitive whose box the
    reference's rounded slab test rejected, which happens only where the hit
    lies within rounding of a box face, i.e. of the triangle's edges. A
    mismatch is excused only when the GPU's id is a valid primitive that the
    oracle's own primitive test hits and
      (i)  tie: the oracle's distance to it equals the oracle's best to `rel`;
      (ii) graze: it is no farther than the oracle's best (or the oracle
           found nothing) and the hit's smallest normalised barycentric m
           satisfies -edge < m < edge.
    Everything else is bad: an unknown or out-of-range id, an id whose
    primitive the oracle's test misses, a miss where the oracle hits, a hit
    farther than the oracle's best."""
    bad, ties, grazes = [], [], []
    n_prims = orc_scene.num_prims
    for i in np.nonzero(prim_gpu != prim_ref)[0]:
        pg, pr = int(prim_gpu[i]), int(prim_ref[i])
        kind = bad
        if 0 <= pg < n_prims:
            dg = orc_scene.prim_distance(pg, o[i], d[i])
            if dg is not None and np.isfinite(dg) and dg > 0:
                tr = float(t_ref[i])
                if pr >= 0 and abs(dg - tr) <= rel * max(abs(tr), 1e-30):
                    kind = ties
                elif pr < 0 or dg <= tr * (1 + rel):
                    m = orc_scene.prim_edge(pg, o[i], d[i])
                    if m is not None and -edge < m < edge:
                        kind = grazes
        kind.append(int(i))
    return bad, ties, grazes


def assert_closest_parity(orc_scene, o, d, prim_gpu, t_gpu, prim_ref, t_ref, rel=1e-5, edge=1e-6,
                          max_graze_frac=1e-4, max_tie_frac=1e-4):
    """Fast-mode closest-hit parity: no bad mismatch; edge grazes below
    `max_graze_frac` of the rays and ties below `max_tie_frac` (scenes with
    coincident faces state a larger tie bound); where ids agree the distances
    agree to `rel`. Returns (ties + grazes, fraction of agreeing hits whose
    distance bits are equal, number of oracle hits)."""
    bad, ties, grazes = classify_mismatches(orc_scene, o, d, prim_gpu, t_gpu, prim_ref, t_ref, rel=rel, edge=edge)
    n = len(prim_ref)
    assert not bad, f"{len(bad)} mismatches outside ties and edge grazes, first {bad[:5]}"
    assert len(grazes) <= allowed(max_graze_frac, n), f"{len(grazes)} edge grazes ({len(ties)} ties) of {n}"
    assert len(ties) <= allowed(max_tie_frac, n), f"{len(ties)} ties ({len(grazes)} edge grazes) of {n}"
    same = (prim_gpu == prim_ref) & (prim_ref >= 0)
    r = np.abs(t_gpu[same] - t_ref[same]) / np.maximum(np.abs(t_ref[same]), 1e-30)
    assert r.max(initial=0) <= rel
    exact = float(np.mean(t_gpu[same].view(np.uint32) == t_ref[same].view(np.uint32))) if same.any() else 1.0
    return len(ties) + len(grazes), exact, int((prim_ref >= 0).sum())


def assert_anyhit_parity(occ_gpu, occ_ref, max_frac=1e-4):
    """Fast-mode any-hit parity: the GPU visits a superset of the
    reference's boxes with the same primitive tests, so every ray the
    reference finds occluded is occluded on the GPU; the GPU may find an
    occluder the reference's rounded slab test missed (an edge graze), on
    at most `max_frac` of the rays."""
    missed = int(((occ_ref != 0) & (occ_gpu == 0)).sum())
    extra = int(((occ_ref == 0) & (occ_gpu != 0)).sum())
    assert missed == 0, f"{missed} rays occluded for the oracle but not on the GPU"
    assert extra <= allowed(max_frac, len(occ_ref)), f"{extra} extra occlusions of {len(occ_ref)}"
    return extra
