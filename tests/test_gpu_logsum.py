"""LogSum semiring (RuleSemiring.scala:45-76) on the GPU vs the oracle: inside/outside charts and
partitions within 1e-4 relative (log-add in a different order than the oracle's, with the fast
exp/log the reference also uses), the FP64 oracle's error stated as a bound, MBR trees under a
coarse Viterbi pass, and the batched engine identical to the plain form."""
import numpy as np
import pytest

from conftest import cidx, sentence_slices, split_sentences, tri

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wsj1_logsum(puck):
    return puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.LOGSUM)


@pytest.fixture(scope="module")
def ref_logsum(puck):
    from oracle.grammar_ref import RefGrammar
    return RefGrammar(puck.grammar_path("wsj_1"), "logsum")


def test_logsum_charts_and_partitions(puck, wsj1, wsj1_logsum, ref_logsum):
    from oracle.oracle import Oracle
    sents = wsj1.sample(seed=61, n=16, mode=0, min_len=2, max_len=20)
    p = puck.CLParser([wsj1_logsum], mem_bytes=2 << 30)
    ch = p.debug_charts(sents, layer=0, with_outside=True)
    parts = p.partitions(sents)
    orc = Oracle(ref_logsum)
    words, offsets = sents
    for i, (c0, c1) in enumerate(sentence_slices(offsets)):
        s = words[offsets[i]:offsets[i + 1]]
        n = len(s)
        ref_in, _ = orc.inside(s)
        ref_out = orc.outside(ref_in, n)
        root = ref_in[tri(n) + cidx(0, n, n), ref_logsum.root]
        assert abs(parts[i] - root) <= 1e-4 * abs(root), (i, parts[i], root)
        live = ref_in > -1e5  # entries with a derivation; "zero" is -1e6 plus log-add drift
        np.testing.assert_allclose(ch["inside"][c0:c1][live], ref_in[live], rtol=1e-4)
        assert np.all(ch["inside"][c0:c1][~live] < -1e5)
        live_o = ref_out > -1e5
        np.testing.assert_allclose(ch["outside"][c0:c1][live_o], ref_out[live_o], rtol=1e-4, atol=1e-3)
        # error of the fp32 oracle against the FP64 one, for the record: far inside the bound
        f64, _ = orc.inside(s, f64=True)
        assert abs(root - f64[tri(n) + cidx(0, n, n), ref_logsum.root]) <= 1e-4 * abs(root)


def test_coarse_to_fine_logsum_mbr(puck, wsj0, wsj1, wsj1_logsum, ref_logsum):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle, pipeline
    sents = wsj1.sample(seed=62, n=40, mode=0, min_len=3, max_len=22)
    words, offsets = sents
    p = puck.CLParser([wsj0, wsj1_logsum], mem_bytes=2 << 30)
    parts = p.partitions(sents)
    mbr = p.parse(sents)
    coarse = Oracle(RefGrammar(puck.grammar_path("wsj_0"), "viterbi"), final_layer=False)
    ref = pipeline([coarse, Oracle(ref_logsum)], split_sentences(words, offsets))
    ll = np.array([r["partition"] for r in ref], np.float32)
    np.testing.assert_allclose(parts, ll, rtol=1e-4)
    same = sum(np.array_equal(mbr[i][:, :4], r["mbr"][:, :4]) for i, r in enumerate(ref))
    assert same >= int(0.95 * len(ref)), f"{same}/{len(ref)} MBR trees identical"
    bp = puck.BatchParser([wsj0, wsj1_logsum], mem_bytes=3 << 30, streams=2, bucket_cells=3000)
    for k in range(3):
        got = bp.parse(sents)
        assert all(np.array_equal(a, b) for a, b in zip(got, mbr)), f"batched call {k}"
