"""The parser's number reading: vertex coordinates are the correctly rounded
float32 of their decimal text (what std::istream's operator>> gives, which
the reference parser uses, muon/parser.cc), whichever of its two readers
takes the token. The expected value is computed exactly: the token as a
rational, rounded to the nearest float32 with ties to even. Tokens include
decimal strings whose nearest double lies exactly halfway between two
float32 values, where rounding through a double would be wrong."""
from fractions import Fraction

import numpy as np

from muon_amd import HostScene


def _nearest_f32(tok):
    q = Fraction(tok)
    if q == 0:
        return np.float32(-0.0) if tok.lstrip().startswith("-") else np.float32(0.0)
    c = np.float32(float(q))
    best = None
    with np.errstate(over="ignore"):
        cands = (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf)))
    for cand in cands:
        if not np.isfinite(cand):
            continue
        err = abs(Fraction(float(cand)) - q)
        key = (err, int(np.array(cand, np.float32).view(np.uint32)) & 1)  # ties: even mantissa
        if best is None or key < best[0]:
            best = (key, cand)
    return np.float32(best[1])


def _tokens(rng):
    toks = []
    vals = rng.standard_normal(300).astype(np.float32) * np.float32(10.0) ** rng.integers(-6, 7, 300)
    for v in vals:
        toks += ["%.9g" % v, "%.17g" % v, "%.3f" % v, "%e" % v, "%.12E" % v]
    # decimal strings that parse to a double exactly halfway between floats
    for v in vals[:100]:
        v = np.float32(abs(v)) + np.float32(1e-3)
        mid = (float(v) + float(np.nextafter(v, np.float32(np.inf)))) / 2
        toks += ["%.17g" % mid, "%.16g" % mid, "%.25g" % mid]
    toks += ["0", "-0", "+1.5", "5.", ".25", "-.75", "1e5", "1E+05", "2.5e-3", "007", "123456789", "0.1",
             "16777217", "9007199254740993", "3.4028234e38", "1e-30", "4.9e-40", "1234567890123456789012"]
    return toks


def test_vertex_coordinates_are_correctly_rounded():
    rng = np.random.default_rng(11)
    toks = _tokens(rng)
    while len(toks) % 3:
        toks.append("1")
    lines = []
    for k in range(0, len(toks), 3):
        lines.append("vertex %s %s %s" % tuple(toks[k:k + 3]))
    nv = len(lines)
    lines += ["tri %d %d %d" % (k, k, k) for k in range(nv)]
    hs = HostScene.from_text("\n".join(lines) + "\n", acceleration=0)
    assert hs.warnings == 0
    g = hs.arrays()["prim_geom"][:, :3].reshape(-1)
    want = np.array([_nearest_f32(t) for t in toks], np.float32)
    bad = np.nonzero(g.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, [(toks[i], g[i], want[i]) for i in bad[:5]]


def test_malformed_numbers_follow_the_stream_reader():
    # trailing text after the three numbers is ignored; a bad token fails the line
    text = "vertex 1 2 3 extra\nvertex 1 2 3abc\nvertex 1,2,3\nvertex 1 2\nvertex 0x10 1 1\ntri 0 1 1\n"
    hs = HostScene.from_text(text, acceleration=0)
    # "1 2 3abc" reads 3 and stops there (no failure); "1,2,3", "1 2" and
    # "0x10 1 1" (0, then "x10" is not a number) fail
    assert hs.warnings == 3
    assert hs.desc.num_prims == 1
