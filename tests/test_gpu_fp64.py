"""Sentence log-likelihoods of the GPU against an FP64 restatement of the same computation (the
oracle with FP64 charts, the same rule literals, masks and scales), one case per workload
configuration: the error is printed and must stay within 1e-4 relative."""
import numpy as np
import pytest

from conftest import cidx, split_sentences, tri

pytestmark = pytest.mark.gpu


def _ll64(rec, ref, n):
    """Log-likelihood from the FP64 final chart, without the pipeline's float32 rounding."""
    root = float(rec["inside"][tri(n) + cidx(0, n, n), ref.root])
    if ref.semiring == "real":
        return np.log(root) + float(rec["in_scale"][tri(n) + cidx(0, n, n)]) + float(rec["partition_scale"])
    return root


def _check(puck, label, layers, ref_layers, words, offsets):
    from oracle.oracle import pipeline
    gpu = puck.CLParser(layers, mem_bytes=8 << 30).partitions((words, offsets)).astype(np.float64)
    recs = pipeline(ref_layers, split_sentences(words, offsets), f64_final=True)
    lens = np.diff(offsets)
    ref = np.array([_ll64(r, ref_layers[-1].ref, int(n)) for r, n in zip(recs, lens)])
    rel = np.abs(gpu - ref) / np.abs(ref)
    print(f"{label}: {len(ref)} sentences, lengths {lens.min()}-{lens.max()}, "
          f"max rel err {rel.max():.2e}, mean {rel.mean():.2e}", flush=True)
    assert rel.max() <= 1e-4


def test_config1_wsj1_real(puck, wsj1_real):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle
    words, offsets = wsj1_real.sample(seed=1, n=200, mode=0, min_len=5, max_len=20)
    _check(puck, "config 1 (wsj_1 Real, exhaustive)", [wsj1_real],
           [Oracle(RefGrammar(puck.grammar_path("wsj_1"), "real"))], words, offsets)


@pytest.fixture(scope="module")
def wsj6_refs(puck):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle
    prefix = puck.synth_wsj6_cached(seed=6)
    coarse = Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"), final_layer=False)
    return prefix, coarse, Oracle(RefGrammar(prefix, "viterbi")), Oracle(RefGrammar(prefix, "real"))


def test_config2_wsj6_viterbi(puck, wsj6_refs):
    prefix, _, fine_v, _ = wsj6_refs
    g6 = puck.Grammar.load_text(prefix)
    words, offsets = g6.sample(seed=31, n=6, mode=0, min_len=1, max_len=8)
    _check(puck, "config 2 (wsj_6 shape Viterbi, exhaustive)", [g6], [fine_v], words, offsets)


def test_config3_c2f_real(puck, wsj1, wsj6_refs):
    prefix, coarse, _, fine_r = wsj6_refs
    g6r = puck.Grammar.load_text(prefix, puck.REAL)
    words, offsets = g6r.sample(seed=33, n=6, mode=0, min_len=1, max_len=12)
    _check(puck, "config 3 (wsj_1 -> wsj_6 shape Real, masks at -9)", [wsj1, g6r], [coarse, fine_r],
           words, offsets)


def test_config4_c2f_viterbi_long(puck, wsj1, wsj6_refs):
    prefix, coarse, fine_v, _ = wsj6_refs
    g6 = puck.Grammar.load_text(prefix)
    words, offsets = g6.sample(seed=32, n=1, mode=0, min_len=41, max_len=48)
    _check(puck, "config 4 (wsj_1 -> wsj_6 shape Viterbi, long tail)", [wsj1, g6], [coarse, fine_v],
           words, offsets)
