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


def classify_mismatches(orc_scene, o, d, prim_gpu, t_gpu, prim_ref, t_ref, rel=1e-6, edge=1e-6):
    """Indices of closest-hit disagreements outside the parity exclusions:
    a disagreement is excluded when it is a distance tie (the oracle's own
    distance to the primitive the GPU reported equals its best distance to
    `rel`) or when either reported hit lies within `edge` of a triangle edge
    (min normalised barycentric < edge)."""
    bad = []
    for i in np.nonzero(prim_gpu != prim_ref)[0]:
        near_edge = any(p >= 0 and orc_scene.prim_edge(int(p), o[i], d[i]) < edge
                        for p in (int(prim_gpu[i]), int(prim_ref[i])))
        if near_edge:
            continue
        if prim_gpu[i] >= 0 and prim_ref[i] >= 0:
            dg = orc_scene.prim_distance(int(prim_gpu[i]), o[i], d[i])
            if dg is not None and abs(dg - t_ref[i]) <= rel * max(abs(t_ref[i]), 1e-30):
                continue
        bad.append(int(i))
    return bad
