"""The oracle against the committed golden fixtures (CPU only)."""
import math
import os

import numpy as np
import pytest

from conftest import split_sentences
from oracle import oracle as O
from oracle.grammar_ref import RefGrammar

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.mark.parametrize("name", ["wsj_0", "wsj_1"])
def test_oracle_matches_brute_force_fixture(puck, name):
    d = load("brute_force.npz")
    w, o = d[f"{name}_words"], d[f"{name}_offsets"]
    for semi, key in (("viterbi", "viterbi"), ("real", "loglik")):
        orc = O.Oracle(RefGrammar(puck.grammar_path(name), semi))
        for i, s in enumerate(split_sentences(w, o)):
            rec = O.pipeline([orc], [s])[0]
            assert math.isclose(float(rec["partition"]), float(d[f"{name}_{key}"][i]), rel_tol=1e-5)


def test_oracle_matches_viterbi_fixture(puck):
    d = load("wsj1_viterbi.npz")
    orc = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"))
    res = O.pipeline([orc], split_sentences(d["words"], d["offsets"]))
    assert np.array_equal(np.concatenate([r["tree"] for r in res]), d["trees"])
    assert np.array_equal(np.array([r["partition"] for r in res], np.float32), d["roots"])


def test_oracle_matches_real_fixture(puck):
    d = load("wsj1_real.npz")
    orc = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "real"))
    res = O.pipeline([orc], split_sentences(d["words"], d["offsets"]))
    np.testing.assert_allclose(np.array([r["partition"] for r in res], np.float32), d["loglik"], rtol=1e-6)


def test_oracle_matches_coarse_to_fine_fixture(puck):
    d = load("wsj0_wsj1_c2f.npz")
    coarse = O.Oracle(RefGrammar(puck.grammar_path("wsj_0"), "viterbi"), final_layer=False)
    fine = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"))
    res = O.pipeline([coarse, fine], split_sentences(d["words"], d["offsets"]))
    assert np.array_equal(np.concatenate([r["coarse"][0]["masks"] for r in res]), d["masks"])
    assert np.array_equal(np.concatenate([r["tree"] for r in res]), d["trees"])
    # pruning bites: some cells of every sentence are masked out
    assert (d["masks"].sum(axis=1) == 0).sum() > 0
