"""Writes tests/golden/spheres64_means.npz: the CPU side of an unbiasedness
check of the GPU path tracer on a sphere field, computed once with the
oracle (the CPU restatement of the reference renderer).

Scene: muon_amd.scenes.sphere_field_text(64, 64, 32 spp, depth 8, 1000
spheres, scene seed 0): Lambert / Phong / GGX spheres and 8 emitting
spheres, BRDF importance sampling, no NEE (the only estimator in which the
reference's sphere emitters light the scene). For render seeds s = 0..N-1
("random_seed s"), the oracle's film / spp is averaged over 2 x 4 blocks of
the rows below tile 0 (the reference seeds tile 0's pixel jitter with 0 for
every render seed, muon/sampling.cc:41-45) and 4 column bands: block_means[s,
row block, column block, channel].

The GPU test renders the same N seeds and compares the block means: both are
unbiased estimators of the same image, so every block's difference of means
must lie within a few standard errors (one render's block mean varies by a
few percent; N = 1024 seeds bring the standard error to about 0.1%).

Usage: python tests/golden/make_spheres_fixture.py [N] (about a minute on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle  # noqa: E402
from muon_amd import scenes  # noqa: E402

W = H = 64
SPP = 32
DEPTH = 8
SPHERES = 1000
PAR = 8


def scene_text(render_seed):
    return scenes.sphere_field_text(W, H, SPP, DEPTH, n=SPHERES, seed=0).replace(
        "random_seed 0", f"random_seed {render_seed}", 1)


def block_means(img, skip):
    """img: [H, W, 3] linear radiance; mean over 2 x 4 blocks of rows skip.. ."""
    rows = img[skip:]
    rb = np.array_split(np.arange(rows.shape[0]), 2)
    cb = np.array_split(np.arange(rows.shape[1]), 4)
    return np.array([[rows[r][:, c].reshape(-1, 3).mean(axis=0) for c in cb] for r in rb], np.float64)


def tile0_rows(o):
    tiles = o.num_tiles(W, H, SPP, PAR)
    return H // tiles + (H - (H // tiles) * tiles)


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    o = _oracle.Oracle()
    skip = tile0_rows(o)
    t0 = time.time()
    out = np.zeros((n, 2, 4, 3), np.float64)
    for s in range(n):
        acc, _, _, _ = o.load_text(scene_text(s)).render(PAR)
        out[s] = block_means(acc / SPP, skip)
    m = out.mean(axis=0)
    se = out.std(axis=0) / np.sqrt(n) / m
    print(f"{n} renders in {time.time() - t0:.0f} s; block means {m.mean():.5f}, relative standard error "
          f"{se.min():.4f}..{se.max():.4f}")
    np.savez_compressed(os.path.join(HERE, "spheres64_means.npz"), block_means=out,
                        meta=np.array([W, H, SPP, DEPTH, SPHERES, PAR, skip]))


if __name__ == "__main__":
    main()
