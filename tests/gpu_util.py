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


def classify_mismatches(orc_scene, o, d, prim_gpu, t_gpu, prim_ref, t_ref, rel=1e-6):
    """Returns the indices of closest-hit disagreements that are not distance
    ties: for a tie, the oracle's own distance to the primitive the GPU
    reported agrees with the oracle's best distance to `rel`."""
    bad = []
    for i in np.nonzero(prim_gpu != prim_ref)[0]:
        if prim_gpu[i] < 0 or prim_ref[i] < 0:
            # a hit on one side only: allowed only for grazing hits, which the
            # oracle confirms by a hit at the same distance on re-test
            bad.append(int(i))
            continue
        dg = orc_scene.prim_distance(int(prim_gpu[i]), o[i], d[i])
        if dg is None or abs(dg - t_ref[i]) > rel * max(abs(t_ref[i]), 1e-30):
            bad.append(int(i))
    return bad
