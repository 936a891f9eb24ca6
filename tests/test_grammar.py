"""Grammar compilation: the product's C++ compiler vs the oracle's Python restatement, pinned to
counts read from the grammar files and to Java Formatter known answers."""
import math

import numpy as np
import pytest

from oracle.grammar_ref import RefGrammar, java_f6, java_f6_string


# -- Java "%f" (SimpleGenRuleMultiply.java:151): HALF_UP on the shortest digits of the widened double
@pytest.mark.parametrize("value,expected", [
    (0.0078125, "0.007813"),       # exact x/128: HALF_UP, where C's printf rounds half-even to ...812
    (-3.5078125, "-3.507813"),
    (0.125, "0.125000"),
    (-1e-7, "-0.000000"),           # sign kept on tiny negatives
    (4.9e-7, "0.000000"),
    (6e-7, "0.000001"),
    (0.1, "0.100000"),               # float 0.1 widened: 0.10000000149011612
    (-7.123456789, "-7.123457"),
    (-12.0, "-12.000000"),
    (123456.789, "123456.789063"),   # float(123456.789) = 123456.7890625 exactly: HALF_UP
    (1.0, "1.000000"),
])
def test_java_f6_known_answers(value, expected):
    assert java_f6_string(float(np.float32(value))) == expected


def test_java_f6_rounds_exact_ties_away_from_zero():
    # k/128 (k odd) = k*15625/2 millionths: exactly half a unit in the 6th decimal. HALF_UP gives
    # (k*15625 + 1)/2 millionths; integer arithmetic, independent of the formatter.
    for k in range(1, 4001, 2):
        units = (k * 15625 + 1) // 2
        want = f"{units // 1000000}.{units % 1000000:06d}"
        assert java_f6_string(k / 128.0) == want
        assert java_f6_string(-k / 128.0) == "-" + want


@pytest.mark.parametrize("name", ["wsj_0", "wsj_1"])
@pytest.mark.parametrize("semi", ["viterbi", "real"])
def test_product_tables_match_oracle(puck, name, semi):
    ref = RefGrammar(puck.grammar_path(name), semi)
    g = puck.Grammar.load_text(puck.grammar_path(name), puck.VITERBI if semi == "viterbi" else puck.REAL)
    assert g.info.num_nonterms == len(ref.nonterms)
    assert g.info.num_terms == len(ref.terms)
    assert g.info.cell_size == ref.cell_size
    assert g.info.root == ref.root
    for f in range(4):
        idx, w, q = g.rules(f)
        b = ref.binaries[f]
        assert np.array_equal(idx[:, 0], b["p"]) and np.array_equal(idx[:, 1], b["l"])
        assert np.array_equal(idx[:, 2], b["r"])
        assert np.array_equal(w.view(np.int32), b["w"].view(np.int32))
        assert np.array_equal(q.view(np.int32), b["q"].view(np.int32))
    for fam, u in ((4, ref.unary_n), (5, ref.unary_t)):
        idx, w, q = g.rules(fam)
        assert np.array_equal(idx[:, 0], u["p"]) and np.array_equal(idx[:, 1], u["c"])
        assert np.array_equal(w.view(np.int32), u["w"].view(np.int32))
        assert np.array_equal(q.view(np.int32), u["q"].view(np.int32))
    for which in range(3):
        for f in range(4):
            nb, boc = g.blocks(which, f)
            rnb, rboc = ref.blocks(which, f)
            assert nb == rnb and np.array_equal(boc, rboc), (which, f)
    coarse_n, sub_n = g.symbols(False)
    assert [f"{ref.coarse[c]}_{s}" for c, s in zip(coarse_n, sub_n)] == ref.nonterms


def test_wsj1_counts_from_files_without_threshold():
    # Counts read from textGrammars/wsj_1.gr.* with RuleStructure's classification and no rule
    # dropped: 75 nonterminals, 69 tags, cell size 96; NN/NT/TN/TT 1318/1617/1484/2198 binaries;
    # 527 + 75 identity N-child unaries, 1224 T-child unaries; CoarseParent blocks 1/1/2/3.
    import puck_amd
    ref = RefGrammar(puck_amd.grammar_path("wsj_1"), threshold=-math.inf)
    assert (len(ref.nonterms), len(ref.terms), ref.cell_size, ref.mask_words) == (75, 69, 96, 4)
    assert [len(b) for b in ref.binaries] == [1318, 1617, 1484, 2198]
    idn = int(np.sum(ref.unary_n["p"] == ref.unary_n["c"]))
    assert (len(ref.unary_n) - idn, idn, len(ref.unary_t)) == (527, 75, 1224)
    assert [ref.blocks(0, f)[0] for f in range(4)] == [1, 1, 2, 3]


def test_threshold_drops_low_probability_rules(puck):
    g = puck.Grammar.load_text(puck.grammar_path("wsj_1"))
    # 237 binary and 1090 unary rules of wsj_1 have ln p < -12.
    assert g.info.rules_dropped == 237 + 1090
    assert sum(g.info.num_binary) == 6617 - 237


def test_lexicon_matches_oracle(puck, wsj1, ref_wsj1):
    ref_wsj1.load_lexicon()
    for word in ["the", "for", "Ministry", "smashing", "."]:
        wid = wsj1.word_id(word)
        assert wid == ref_wsj1.word_ids[word]
        tags, logp = wsj1.lexicon(wid)
        s, e = ref_wsj1.lex_off[wid], ref_wsj1.lex_off[wid + 1]
        assert np.array_equal(tags, ref_wsj1.lex_tag[s:e])
        assert np.array_equal(logp, ref_wsj1.lex_logp[s:e])
    assert wsj1.word_id("no-such-word-xyz") == -1


def test_missing_grammar_reports_io_error(puck):
    with pytest.raises(puck.PuckError, match="cannot open"):
        puck.Grammar.load_text("/nonexistent/prefix")


def test_synthetic_wsj6_has_published_shape(puck, tmp_path):
    # textGrammars/wsj_6.gr.numstates: 1084 refined states (502 nonterminal + 582 tag);
    # wsj_6.gr.num.binary sums to 1,725,570 and wsj_6.gr.num.unary to 125,619 (identities included).
    prefix = puck.synth_wsj6_cached(seed=6)
    g = puck.Grammar.load_text(prefix)
    i = g.info
    assert (i.num_nonterms, i.num_terms, i.cell_size) == (502, 582, 608)
    assert sum(i.num_binary) == 1725570
    # terminal identities are not kept (RuleStructure.scala:194), nonterminal identities are
    assert i.num_unary_n + i.num_unary_t == 125619 - 582
    assert i.rules_dropped == 0
