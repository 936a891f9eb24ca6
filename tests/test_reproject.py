"""Hierarchical reprojection (CLParser.reprojectGrammar / readHierarchy, CLParser.scala:1153-1208):
the oracle's hierarchy reader pinned by hand-derived levels and by the substate counts of the split
grammars, then the product's reprojected tables against the oracle's, bit for bit."""
import os

import numpy as np
import pytest

from oracle.grammar_ref import RefGrammar, level_of, read_hierarchy, reproject

HAND = "S\t(0 (0 (0 0 1)) (1 (1 2) (2 3 4)))\n@S\t(0 (0 (0 0)) (1 (1 1) (2 2)))\n"


def _numstates(path):
    out = {}
    for line in open(path):
        if line.strip():
            s, n = line.split()
            out[s] = int(n)
    return out


@pytest.fixture
def hand_file(tmp_path):
    p = tmp_path / "toy.hierarchy"
    p.write_text(HAND)
    return str(p)


@pytest.mark.parametrize("coarse,fine,want", [
    # Words are the level-3 substates; the root is level 0.
    (1, 3, {"S_0": ["S_0", "S_1"], "S_1": ["S_2", "S_3", "S_4"]}),
    (0, 3, {"S_0": ["S_0", "S_1", "S_2", "S_3", "S_4"]}),
    (2, 3, {"S_0": ["S_0", "S_1"], "S_1": ["S_2"], "S_2": ["S_3", "S_4"]}),
    # A fine level above the words takes the node labels at that level.
    (1, 2, {"S_0": ["S_0"], "S_1": ["S_1", "S_2"]}),
    (0, 1, {"S_0": ["S_0", "S_1"]}),
])
def test_hierarchy_levels_hand_example(hand_file, coarse, fine, want):
    got = read_hierarchy(hand_file, coarse, fine)
    assert {k: v for k, v in got.items() if k.startswith("S_")} == want


def test_level_from_name():
    assert level_of("textGrammars/wsj_6.gr") == 6
    assert level_of("/a/b12/wsj_12.gr") == 12
    assert level_of("en_sm6") == 6
    with pytest.raises(ValueError):
        level_of("grammar.gr")


def test_wsj6_hierarchy_levels_match_split_grammars(puck):
    """Level c of wsj_6's hierarchy has wsj_c's substates (c = 0, 1) and every level partitions the
    wsj_6 substates of each symbol; deeper levels refine shallower ones."""
    data = os.path.dirname(puck.grammar_path("wsj_6"))
    path = puck.grammar_path("wsj_6") + ".hierarchy"
    n6 = _numstates(puck.grammar_path("wsj_6") + ".numstates")
    levels = {c: _numstates(os.path.join(data, f"wsj_{c}.gr.numstates")) for c in (0, 1)}
    ancestor = {}
    for c in range(6):
        m = read_hierarchy(path, c, 6)
        per_sym = {}
        for coarse, fines in m.items():
            base = coarse.rsplit("_", 1)[0]
            per_sym.setdefault(base, []).append(coarse)
            for f in fines:
                ancestor[(c, f)] = coarse
        for base, n in n6.items():
            if base == "ROOT":
                continue
            subs = sorted(int(f.rsplit("_", 1)[1]) for cs in per_sym[base] for f in m[cs])
            assert subs == list(range(n)), (c, base)
            if c in levels:
                assert len(per_sym[base]) == levels[c][base], (c, base)
    for (c, f), a in ancestor.items():
        if c + 1 < 6:
            # two fine symbols with one level-(c+1) ancestor share their level-c ancestor
            mates = [g for (cc, g), b in ancestor.items() if cc == c + 1 and b == ancestor[(c + 1, f)]]
            assert all(ancestor[(c, g)] == a for g in mates)


def _compare(g, ref, puck):
    assert g.coarse_names() == ref.coarse
    assert g.info.mask_words == ref.mask_words == (len(ref.coarse) + 31) // 32
    assert np.array_equal(g.symbols(0)[0], ref.proj_n)
    assert np.array_equal(g.symbols(1)[0], ref.proj_t)
    for fam in range(4):
        idx, w, q = g.rules(fam)
        b = ref.binaries[fam]
        assert np.array_equal(idx[:, 0], b["p"])
        assert np.array_equal(w.view(np.int32), b["w"].view(np.int32)), fam
        assert np.array_equal(q.view(np.int32), b["q"].view(np.int32)), fam
    for fam, u in ((4, ref.unary_n), (5, ref.unary_t)):
        idx, w, q = g.rules(fam)
        assert np.array_equal(w.view(np.int32), u["w"].view(np.int32)), fam
        assert np.array_equal(q.view(np.int32), u["q"].view(np.int32)), fam
    for which in range(3):
        for fam in range(4):
            nb, boc = g.blocks(which, fam)
            rnb, rboc = ref.blocks(which, fam)
            assert nb == rnb and np.array_equal(boc, rboc), (which, fam)


@pytest.mark.parametrize("semi", ["viterbi", "real"])
def test_product_matches_oracle_wsj0_wsj1(puck, semi):
    rc, rf = RefGrammar(puck.grammar_path("wsj_0")), RefGrammar(puck.grammar_path("wsj_1"), semi)
    reproject(rc, rf)
    g0 = puck.Grammar.load_text(puck.grammar_path("wsj_0"))
    g1 = puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.VITERBI if semi == "viterbi" else puck.REAL)
    c, f = puck.Grammar.reproject(g0, g1)  # hierarchy and levels from the prefixes
    _compare(c, rc, puck)
    _compare(f, rf, puck)
    # Level 0 is the unsplit symbol: the projection is the base one, apart from the rules whose
    # coarse rule wsj_0's -12 threshold dropped.
    assert f.coarse_names() == [s + "_0" for s in g1.coarse_names()]
    assert np.array_equal(f.symbols(0)[0], g1.symbols(0)[0])
    changed = sum(int(np.sum(f.rules(k)[1] != g1.rules(k)[1])) for k in range(6))
    assert 0 < changed < 200
    # the inputs keep their own tables
    assert g1.coarse_names()[0] == "ROOT" and g1.info.num_coarse == 97


@pytest.fixture(scope="module")
def wsj6_pair(puck):
    prefix = puck.synth_wsj6_cached(seed=6)
    rc, rf = RefGrammar(puck.grammar_path("wsj_1")), RefGrammar(prefix)
    reproject(rc, rf, coarse_level=1, fine_level=6)
    c, f = puck.Grammar.reproject(puck.Grammar.load_text(puck.grammar_path("wsj_1")),
                                  puck.Grammar.load_text(prefix), coarse_level=1, fine_level=6)
    return rc, rf, c, f


def test_product_matches_oracle_wsj1_wsj6(puck, wsj6_pair):
    rc, rf, c, f = wsj6_pair
    _compare(c, rc, puck)
    _compare(f, rf, puck)
    assert c.info.num_coarse == 144 and c.info.mask_words == 5
    # one block per coarse-level parent with >= 300 rules, the rest lumped
    assert all(8 < f.info.num_blocks[k] <= 128 for k in range(4))


def test_reprojected_layers_round_trip_grz(puck, wsj6_pair, tmp_path):
    _, _, c, f = wsj6_pair
    path = str(tmp_path / "reproj.grz")
    puck.save_grz(path, [c, f])
    c2, f2 = puck.Grammar.load_grz(path)
    for a, b in ((c, c2), (f, f2)):
        assert a.coarse_names() == b.coarse_names()
        assert np.array_equal(a.symbols(0)[0], b.symbols(0)[0])
        for fam in range(6):
            assert np.array_equal(a.rules(fam)[1].view(np.int32), b.rules(fam)[1].view(np.int32))


def test_reprojection_errors(puck, tmp_path, hand_file):
    g0 = puck.Grammar.load_text(puck.grammar_path("wsj_0"))
    g1 = puck.Grammar.load_text(puck.grammar_path("wsj_1"))
    with pytest.raises(puck.PuckError, match="cannot open"):
        puck.Grammar.reproject(g0, g1, hierarchy=str(tmp_path / "missing.hierarchy"))
    with pytest.raises(puck.PuckError, match="coarse level"):
        puck.Grammar.reproject(g1, g0, hierarchy=puck.grammar_path("wsj_1") + ".hierarchy",
                               coarse_level=1, fine_level=1)
    # a hierarchy without most of the grammar's symbols
    with pytest.raises(puck.PuckError, match="is not in"):
        puck.Grammar.reproject(g0, g1, hierarchy=hand_file, coarse_level=0, fine_level=1)
