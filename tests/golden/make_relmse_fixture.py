"""Writes tests/golden/relmse_cornell128.npz: the CPU side of the path-tracer
quality criterion, computed once with the oracle (the CPU restatement of the
reference renderer).

* conv: the Cornell box (muon_amd.scenes.cornell_text) at 128x128, max depth
  5, rendered by the oracle at 16384 spp, seed 99, divided by spp (linear).
* cpu_relmse[s]: relMSE of the oracle's 16-spp render with scene seed s
  (s = 0..N-1, 8 render threads) against conv, over the rows below tile 0
  (the reference seeds tile 0's pixel jitter with 0 for every scene seed,
  muon/sampling.cc:41-45, so those rows repeat across seeds).

The reference's CPU renders are not repeatable run to run (the tile each
thread takes decides which random stream it sees), so one 8-seed mean
varies by a few percent; N = 512 seeds bring the standard error of the mean
to about 0.35%.

Usage: python tests/golden/make_relmse_fixture.py [N] (about 10 minutes on 8 cores)
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

W = H = 128
SPP = 16
DEPTH = 5
PAR = 8
CONV_SPP = 16384


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 512
    o = _oracle.Oracle()
    tiles = o.num_tiles(W, H, SPP, PAR)
    skip = H // tiles + (H - (H // tiles) * tiles)
    t0 = time.time()
    conv, _, _, _ = o.load_text(scenes.cornell_text(W, H, CONV_SPP, DEPTH, seed=99)).render(PAR)
    conv = (conv / CONV_SPP).astype(np.float32)
    print(f"converged image: {time.time() - t0:.0f} s", flush=True)
    vals = []
    for s in range(n):
        acc, _, _, _ = o.load_text(scenes.cornell_text(W, H, SPP, DEPTH, seed=s)).render(PAR)
        vals.append(relmse((acc / SPP)[skip:], conv[skip:]))
    vals = np.array(vals)
    print(f"cpu relMSE mean {vals.mean():.5f}  sem {vals.std() / np.sqrt(n) / vals.mean():.4f} (relative)")
    np.savez_compressed(os.path.join(HERE, "relmse_cornell128.npz"), conv=conv, cpu_relmse=vals,
                        meta=np.array([W, H, SPP, DEPTH, PAR, CONV_SPP, skip]))


if __name__ == "__main__":
    main()
