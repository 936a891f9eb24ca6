"""HIP path vs the oracle (CPU restatement) on the same seeded inputs, through the C ABI.

Tolerances: Viterbi charts, trees, masks are bit-exact (integer/bit comparison of the fp32 values);
Real-semiring sentence log-likelihoods within 1e-4 relative.
"""
import numpy as np
import pytest

from conftest import cidx, sentence_slices, split_sentences, tri

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.int32), b.view(np.int32))


@pytest.fixture(scope="module")
def wsj1_sents(wsj1):
    # 200 PCFG sentences of length 5..20 (seeded): the small plain-method workload.
    words, offsets = wsj1.sample(seed=1, n=200, mode=0, min_len=5, max_len=20)
    return words, offsets


def test_viterbi_inside_chart_bit_exact(puck, wsj1, ref_wsj1, wsj1_sents):
    from oracle.oracle import Oracle
    words, offsets = wsj1_sents
    sub = (words[:offsets[24]], offsets[:25])
    p = puck.CLParser([wsj1], mem_bytes=2 << 30)
    ch = p.debug_charts(sub, layer=0, with_outside=False)
    orc = Oracle(ref_wsj1)
    for i, (c0, c1) in enumerate(sentence_slices(sub[1])):
        s = sub[0][sub[1][i]:sub[1][i + 1]]
        print("sentence", i, "len", len(s), flush=True)
        ref, _ = orc.inside(s)
        got = ch["inside"][c0:c1]
        bad = np.nonzero(~np.all(got.view(np.int32) == ref.view(np.int32), axis=1))[0]
        assert bad.size == 0, f"sentence {i}: cells {bad[:5]} differ"


def test_lone_split_pairs_bit_exact(puck, wsj1, ref_wsj1):
    """Spans with an odd number of splits (15 or more) fold split pairs in the rule kernel with a
    lone last split, whose missing partner item must contribute the semiring zero. Inside charts
    of sentences of length 16-21 bit-exact against the oracle, in the plain and persistent forms."""
    from oracle.oracle import Oracle
    words, offsets = wsj1.sample(seed=17, n=6, mode=0, min_len=16, max_len=21)
    lens = np.diff(offsets)
    assert (lens >= 16).all()
    orc = Oracle(ref_wsj1)
    refs = [orc.inside(words[offsets[i]:offsets[i + 1]])[0] for i in range(len(lens))]
    for persistent in (False, True):
        p = puck.CLParser([wsj1], mem_bytes=2 << 30, persistent=persistent)
        ch = p.debug_charts((words, offsets), layer=0, with_outside=False)
        for i, (c0, c1) in enumerate(sentence_slices(offsets)):
            print("persistent", persistent, "sentence", i, "len", lens[i], flush=True)
            got = ch["inside"][c0:c1]
            bad = np.nonzero(~np.all(got.view(np.int32) == refs[i].view(np.int32), axis=1))[0]
            assert bad.size == 0, f"persistent={persistent} sentence {i}: cells {bad[:5]} differ"


def test_viterbi_trees_bit_exact(puck, wsj1, ref_wsj1, wsj1_sents):
    from oracle.oracle import Oracle
    words, offsets = wsj1_sents
    p = puck.CLParser([wsj1], mem_bytes=2 << 30)
    trees = p.parse((words, offsets))
    orc = Oracle(ref_wsj1)
    for i, s in enumerate(split_sentences(words, offsets)):
        ref_in, _ = orc.inside(s)
        ref_tree = orc.viterbi(ref_in, len(s))
        assert np.array_equal(trees[i], ref_tree), f"sentence {i} (len {len(s)}) tree differs"


def test_viterbi_partitions_match_root(puck, wsj1, ref_wsj1, wsj1_sents):
    from oracle.oracle import Oracle
    words, offsets = wsj1_sents
    p = puck.CLParser([wsj1], mem_bytes=2 << 30)
    parts = p.partitions((words[:offsets[40]], offsets[:41]))
    orc = Oracle(ref_wsj1)
    for i, s in enumerate(split_sentences(words, offsets)[:40]):
        ref_in, _ = orc.inside(s)
        n = len(s)
        assert _same_bits(parts[i], ref_in[tri(n) + cidx(0, n, n), ref_wsj1.root])


def test_real_inside_outside_and_partitions(puck, wsj1_real, ref_wsj1_real, wsj1_sents):
    from oracle.oracle import Oracle
    words, offsets = wsj1_sents
    sub = (words[:offsets[16]], offsets[:17])
    p = puck.CLParser([wsj1_real], mem_bytes=2 << 30)
    ch = p.debug_charts(sub, layer=0, with_outside=True)
    parts = p.partitions(sub)
    orc = Oracle(ref_wsj1_real)
    for i, (c0, c1) in enumerate(sentence_slices(sub[1])):
        s = sub[0][sub[1][i]:sub[1][i + 1]]
        n = len(s)
        ref_in, ps = orc.inside(s)
        ref_out = orc.outside(ref_in, n)
        root = ref_in[tri(n) + cidx(0, n, n), ref_wsj1_real.root]
        ref_ll = np.float32(np.log(np.float64(root))) + np.float32(ps)
        assert abs(parts[i] - ref_ll) <= 1e-4 * abs(ref_ll), (i, parts[i], ref_ll)
        np.testing.assert_allclose(ch["inside"][c0:c1], ref_in, rtol=2e-4, atol=1e-30)
        np.testing.assert_allclose(ch["outside"][c0:c1], ref_out, rtol=2e-3, atol=1e-25)
