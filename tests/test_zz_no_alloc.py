"""The no-allocation parse (puck_parse_viterbi_noalloc) through the C ABI: trees bit-identical to
puck_parse_viterbi's on two calls in a row, the second with a smaller input, on the exhaustive wsj_1
grammar and on wsj_1 -> wsj_6-shape coarse-to-fine; an input too large for the scratch array is
rejected with an error and leaves the handle usable."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(puck, parser, words, offsets):
    w = np.ascontiguousarray(words, np.int32)
    o = np.ascontiguousarray(offsets, np.int32)
    lens = np.diff(o)
    total = int(np.sum(2 * lens - 1))
    nodes = (puck.TreeNode * max(total, 1))()
    rc = puck.lib().puck_parse_viterbi_noalloc(parser._h, w.ctypes.data_as(puck._I32P), o.ctypes.data_as(puck._I32P),
                                               len(lens), nodes, total)
    if rc < 0:
        raise puck.PuckError(puck.lib().puck_last_error().decode())
    arr = np.ctypeslib.as_array(nodes).view(np.int32).reshape(-1, 4)[:total]
    out, s = [], 0
    for n in lens:
        out.append(arr[s:s + 2 * n - 1].copy())
        s += 2 * n - 1
    return out


def _same(a, b):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"sentence {i}"


def test_noalloc_matches_parse_exhaustive(puck, wsj1):
    p = puck.CLParser([wsj1], mem_bytes=2 << 30)
    big = wsj1.sample(seed=91, n=48, mode=0, min_len=3, max_len=30)
    small = wsj1.sample(seed=92, n=7, mode=0, min_len=1, max_len=12)
    for sents in (big, small):
        want = p.parse(sents)
        _same(_noalloc(puck, p, *sents), want)
    # the parser's own path still works after the no-alloc calls
    _same(p.parse(big), puck.CLParser([wsj1], mem_bytes=2 << 30).parse(big))


def test_noalloc_matches_parse_coarse_to_fine(puck, wsj1):
    g6 = puck.Grammar.load_text(puck.synth_wsj6_cached(seed=6), puck.VITERBI)
    p = puck.CLParser([wsj1, g6], mem_bytes=8 << 30)
    big = g6.sample(seed=93, n=6, mode=0, min_len=5, max_len=20)
    small = g6.sample(seed=94, n=3, mode=0, min_len=1, max_len=9)
    for sents in (big, small):
        want = p.parse(sents)
        _same(_noalloc(puck, p, *sents), want)


def test_noalloc_rejects_oversize_input(puck, wsj1):
    p = puck.CLParser([wsj1], mem_bytes=2 << 30)
    huge = wsj1.sample(seed=95, n=400, mode=0, min_len=30, max_len=40)
    with pytest.raises(puck.PuckError, match="no-alloc parse holds at most"):
        _noalloc(puck, p, *huge)
    small = wsj1.sample(seed=96, n=5, mode=0, min_len=2, max_len=10)
    _same(_noalloc(puck, p, *small), p.parse(small))
