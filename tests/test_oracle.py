"""Pins the oracle: brute force over tree shapes, closed forms and invariants (CPU only)."""
import math

import numpy as np
import pytest

from conftest import cidx, split_sentences, tri
from oracle import oracle as O


def test_chart_index_is_span_diagonal_major_bijection():
    # ChartHalf.chartIndex (ParseChart.scala:58-65) enumerates cells span by span, left to right.
    for n in range(1, 81):
        seen = []
        for w in range(1, n + 1):
            for b in range(0, n - w + 1):
                seen.append(cidx(b, b + w, n))
        assert seen == list(range(tri(n)))
        assert cidx(3 % n, 0, n) == cidx(0, 3 % n, n) if n > 3 else True


@pytest.mark.parametrize("semi", ["viterbi", "real", "logsum"])
@pytest.mark.parametrize("name", ["wsj_0", "wsj_1"])
def test_inside_root_equals_brute_force(puck, semi, name):
    from oracle.grammar_ref import RefGrammar
    g = puck.Grammar.load_text(puck.grammar_path(name))
    ref = RefGrammar(puck.grammar_path(name), semi)
    orc = O.Oracle(ref)
    words, offsets = g.sample(seed=11, n=6, mode=0, min_len=1, max_len=4)
    for s in split_sentences(words, offsets):
        n = len(s)
        chart, ps = orc.inside(s)
        root = float(chart[tri(n) + cidx(0, n, n), ref.root])
        bf = O.brute_force_root(ref, list(s), semi)
        if semi == "viterbi":
            assert math.isclose(root, bf, rel_tol=1e-6), (n, root, bf)
        elif semi == "logsum":
            # fp32 log-add in rule order vs exact logaddexp over whole trees
            assert math.isclose(root, bf, rel_tol=1e-5), (n, root, bf)
        else:
            assert math.isclose(math.log(root) + ps, math.log(bf), rel_tol=1e-6), (n, root, bf)


def test_real_inside_outside_identity(puck, wsj1, ref_wsj1_real):
    # Every derivation passes through exactly one tag at each position, so
    # sum_t in(i,t) * out(i,t) over the span-1 bottom cell equals the root inside score.
    orc = O.Oracle(ref_wsj1_real)
    words, offsets = wsj1.sample(seed=5, n=5, mode=0, min_len=3, max_len=12)
    for s in split_sentences(words, offsets):
        n = len(s)
        ins, _ = orc.inside(s, f64=True)
        out = orc.outside(ins, n, f64=True)
        z = ins[tri(n) + cidx(0, n, n), ref_wsj1_real.root]
        for i in range(n):
            c = cidx(i, i + 1, n)
            assert math.isclose(float(np.dot(ins[c], out[c])), z, rel_tol=1e-9)


def test_logsum_inside_outside_identity(puck, wsj1):
    # In log space the same identity reads: logsumexp_t (in(i,t) + out(i,t)) = in(root) at every
    # position i (FP64 oracle, log-add of RuleSemiring.scala:58-73).
    from oracle.grammar_ref import RefGrammar
    ref = RefGrammar(puck.grammar_path("wsj_1"), "logsum")
    orc = O.Oracle(ref)
    words, offsets = wsj1.sample(seed=6, n=4, mode=0, min_len=3, max_len=12)
    for s in split_sentences(words, offsets):
        n = len(s)
        ins, _ = orc.inside(s, f64=True)
        out = orc.outside(ins, n, f64=True)
        z = ins[tri(n) + cidx(0, n, n), ref.root]
        for i in range(n):
            c = cidx(i, i + 1, n)
            v = ins[c] + out[c]
            assert math.isclose(float(np.logaddexp.reduce(v[v > -1e5])), z, rel_tol=1e-9)


def test_logsum_matches_log_of_real_sum(puck, wsj1):
    # LogSum's partition is log of the Real semiring's: the two differ only by the %f rounding
    # of the rule literals (ln p vs p), far below this tolerance.
    from oracle.grammar_ref import RefGrammar
    ls = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "logsum"))
    re_ = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "real"))
    words, offsets = wsj1.sample(seed=8, n=4, mode=0, min_len=4, max_len=10)
    for s in split_sentences(words, offsets):
        n = len(s)
        a, _ = ls.inside(s, f64=True)
        b, ps = re_.inside(s, f64=True)
        r = tri(n) + cidx(0, n, n)
        assert abs(a[r, ls.ref.root] - (math.log(b[r, re_.ref.root]) + ps)) < 2e-3 * n


def test_fp32_inside_close_to_fp64(wsj1, ref_wsj1_real):
    orc = O.Oracle(ref_wsj1_real)
    words, offsets = wsj1.sample(seed=7, n=4, mode=0, min_len=10, max_len=20)
    for s in split_sentences(words, offsets):
        n = len(s)
        a, _ = orc.inside(s)
        b, _ = orc.inside(s, f64=True)
        r = tri(n) + cidx(0, n, n)
        assert abs(math.log(a[r, ref_wsj1_real.root]) - math.log(b[r, ref_wsj1_real.root])) < 1e-4


def _tree_spans(tree):
    """Walk a pre-order tree_t array, returning (node, begin, end) and checking widths."""
    out = []
    begin = 0
    for p in range(len(tree)):
        w = tree[p, 2]
        out.append((p, begin, begin + w))
        if w == 1:
            begin += 1
    return out


def test_viterbi_tree_is_consistent_and_scores_recompute(puck, wsj1, ref_wsj1):
    # The decoder's tree must cover the sentence, and each internal node's score must equal
    # (rScore + left) + right for the rule it chose, with rScore the unquantized decoder score.
    orc = O.Oracle(ref_wsj1)
    off, scores, lefts, rights = orc.dec
    nN = len(ref_wsj1.nonterms)
    words, offsets = wsj1.sample(seed=3, n=20, mode=0, min_len=2, max_len=15)
    for s in split_sentences(words, offsets):
        n = len(s)
        ins, _ = orc.inside(s)
        tree = orc.viterbi(ins, n)
        spans = _tree_spans(tree)
        assert spans[0][1:] == (0, n)
        assert sum(1 for _, b, e in spans if e - b == 1) == n
        sc = tree[:, 3].view(np.float32)
        assert tree[0, 0] == ref_wsj1.root
        for p, b, e in spans:
            if e - b == 1:
                # terminal nodes carry the tag's inside score exactly (CLViterbi.scala:538)
                assert sc[p] == ins[cidx(b, b + 1, n), tree[p, 1]]
            else:
                # internal nodes carry the best binary score for their bottom symbol, computed with
                # the decoder's unquantized rule scores: it differs from the chart cell (built with
                # 6-decimal literals) only by the quantization of the rules below it.
                assert abs(sc[p] - ins[cidx(b, e, n), tree[p, 1]]) < 1e-5 * (2 * (e - b) - 1) + 1e-4


def test_masks_closed_form(wsj0, puck):
    # bit c of a cell is set iff some refined symbol projecting to c has in + out >= root - 9
    from oracle.grammar_ref import RefGrammar
    ref = RefGrammar(puck.grammar_path("wsj_0"))
    orc = O.Oracle(ref, final_layer=False)
    words, offsets = wsj0.sample(seed=9, n=4, mode=0, min_len=4, max_len=10)
    for s in split_sentences(words, offsets):
        n = len(s)
        ins, _ = orc.inside(s)
        out = orc.outside(ins, n)
        m = orc.masks(ins, out, n)
        root = ins[n * (n + 1) - 1, ref.root]
        for cell in range(n * (n + 1)):
            proj = ref.proj_t if cell < n else ref.proj_n
            want = np.zeros(ref.mask_words, np.uint32)
            for sym in range(ref.num_syms):
                c = proj[sym] if sym < len(proj) else 0
                if np.float32(ins[cell, sym] + out[cell, sym]) >= np.float32(root + np.float32(-9.0)):
                    want[c // 32] |= np.uint32(1 << (c % 32))
            assert np.array_equal(m[cell], want)
