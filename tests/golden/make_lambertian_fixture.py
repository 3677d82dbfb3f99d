"""Writes tests/golden/lambertian_cornell32.npz: the CPU side of the check of
the GPU's Lambertian BRDF (muon/materials.cc:17-47), computed once with the
oracle (the CPU restatement of the reference renderer).

The reference's parser never builds this BRDF (`brdf lambertian` falls
through to Phong, muon/parser.cc:165-178); the oracle's test hook
load_text_lambertian builds it, and the GPU gets it through the scene
descriptor's material tag MUONX_BRDF_LAMBERTIAN. Scene: the Cornell box
(muon_amd.scenes.cornell_text, 32x32, 16 spp, max depth 5, BRDF sampling +
NEE + MIS) with every material Lambertian and a specular colour of 0.5 on
the walls and boxes, which the Lambertian BRDF ignores (a Phong BRDF would
add a specular lobe). For render seeds s = 0..N-1 the oracle's film / spp is
averaged over 2 x 2 blocks of the rows below tile 0:
block_means[s, row block, column block].

Usage: python tests/golden/make_lambertian_fixture.py [N]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

W = H = 32
SPP = 16
DEPTH = 5
PAR = 8


def scene_text(seed):
    from muon_amd import scenes
    t = scenes.cornell_text(W, H, SPP, DEPTH, seed=seed)
    assert "specular 0 0 0\n" in t
    return "brdf lambertian\n" + t.replace("specular 0 0 0\n", "specular 0.5 0.5 0.5\n")


def block_means(img, skip):
    rows = img[skip:]
    rb = np.array_split(np.arange(rows.shape[0]), 2)
    cb = np.array_split(np.arange(rows.shape[1]), 2)
    return np.array([[rows[r][:, c].mean() for c in cb] for r in rb], np.float64)


def main():
    import _oracle
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 64
    o = _oracle.Oracle()
    tiles = o.num_tiles(W, H, SPP, PAR)
    skip = H // tiles + (H - (H // tiles) * tiles)
    t0 = time.time()
    out = np.zeros((n, 2, 2), np.float64)
    phong = np.zeros((n, 2, 2), np.float64)
    for s in range(n):
        out[s] = block_means(o.load_text_lambertian(scene_text(s)).render(PAR)[0] / SPP, skip)
        phong[s] = block_means(o.load_text(scene_text(s)).render(PAR)[0] / SPP, skip)
    print(f"{n} seeds in {time.time() - t0:.0f} s; Lambertian / Phong mean {out.mean() / phong.mean():.3f}")
    np.savez_compressed(os.path.join(HERE, "lambertian_cornell32.npz"), block_means=out, phong_block_means=phong,
                        meta=np.array([W, H, SPP, DEPTH, PAR, skip]))


if __name__ == "__main__":
    main()
