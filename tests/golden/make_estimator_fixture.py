"""Writes tests/golden/estimators_cornell32.npz: the CPU side of the
estimator-option check of the GPU path tracer, computed once with the oracle
(the CPU restatement of the reference renderer).

Scene: the Cornell box (muon_amd.scenes.cornell_text, 32x32, 32 spp, max
depth 5) with a GGX and a Phong sphere added. Options: next-event estimation
off / on / MIS x hemisphere / cosine / BRDF importance sampling, plus Russian
roulette, 4 stratified light samples, 2 light samples and min_depth 2. For
each option and render seed s = 0..N-1 the oracle's film / spp is averaged
over 2 x 2 blocks of the rows below tile 0 (the reference seeds tile 0's
pixel jitter with 0 for every seed, muon/sampling.cc:41-45):
block_means[option, s, row block, column block].

Usage: python tests/golden/make_estimator_fixture.py [N] (about a minute on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

W = H = 32
SPP = 32
DEPTH = 5
PAR = 8

OPTIONS = [(nee, isamp, "") for nee in ("off", "on", "mis") for isamp in ("hemisphere", "cosine", "brdf")] + [
    ("mis", "brdf", "russian_roulette on"),
    ("mis", "brdf", "light_samples 4\nlight_stratify on"),
    ("on", "cosine", "light_samples 2"),
    ("mis", "brdf", "min_depth 2"),
]


def scene_text(seed, nee, isamp, extra):
    from muon_amd import scenes
    t = scenes.cornell_text(W, H, SPP, DEPTH, seed=seed)
    t = t.replace("next_event_estimation mis", f"next_event_estimation {nee}")
    t = t.replace("importance_sampling brdf", f"importance_sampling {isamp}")
    if extra:
        t = t.replace("russian_roulette off", "russian_roulette off\n" + extra)
    t += ("brdf ggx\nroughness 0.3\ndiffuse 0.1 0.1 0.1\nspecular 0.6 0.6 0.6\nsphere 0.35 0.85 0.2 0.25\n"
          "brdf phong\nshininess 40\ndiffuse 0.3 0.1 0.1\nspecular 0.4 0.4 0.4\nsphere -0.4 0.3 0.5 0.3\n")
    return t


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
    out = np.zeros((len(OPTIONS), n, 2, 2), np.float64)
    for k, opt in enumerate(OPTIONS):
        for s in range(n):
            out[k, s] = block_means(o.load_text(scene_text(s, *opt)).render(PAR)[0] / SPP, skip)
    print(f"{len(OPTIONS)} options x {n} seeds in {time.time() - t0:.0f} s")
    np.savez_compressed(os.path.join(HERE, "estimators_cornell32.npz"), block_means=out,
                        meta=np.array([W, H, SPP, DEPTH, PAR, skip]))


if __name__ == "__main__":
    main()
