"""Real sentences on the GPU: the 68 Wall Street Journal sentences of the reference's own test
treebank (tests/golden/smallbank_sentences.txt, 8-45 tokens), mapped to word ids by the text front
end. wsj_0 -> wsj_1 coarse-to-fine Viterbi trees through CLParser and BatchParser, and exhaustive
wsj_1 Viterbi trees, bit-exact against the oracle; Real coarse-to-fine log-likelihoods within 1e-4;
and the `parse` front end end to end: one bracketed tree per line whose leaves are the line's
tokens in order."""
import os
import re

import numpy as np
import pytest

from test_realtext import GOLD, smallbank, smallbank_ids

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sents(wsj1):
    ids = [np.asarray(w, np.int32) for w in smallbank_ids(wsj1)]
    words = np.concatenate(ids)
    offsets = np.concatenate([[0], np.cumsum([len(w) for w in ids])]).astype(np.int32)
    return words, offsets


@pytest.fixture(scope="module")
def layers(puck):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle
    return {"coarse": Oracle(RefGrammar(puck.grammar_path("wsj_0"), "viterbi"), final_layer=False),
            "fine_v": Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi")),
            "fine_r": Oracle(RefGrammar(puck.grammar_path("wsj_1"), "real"))}


def _same(got, want, what):
    assert len(got) == len(want), what
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), f"{what}: sentence {i}"


def test_c2f_viterbi_trees(puck, wsj0, wsj1, sents, layers):
    from oracle.oracle import batch
    ref = batch([layers["coarse"], layers["fine_v"]], *sents)
    p = puck.CLParser([wsj0, wsj1], mem_bytes=4 << 30)
    np.testing.assert_array_equal(p.masks(sents), ref["masks"])
    _same(p.parse(sents), ref["trees"], "CLParser")
    bp = puck.BatchParser([wsj0, wsj1], mem_bytes=4 << 30, streams=2, bucket_cells=4000)
    for k in range(2):  # eager, then a graph capture
        _same(bp.parse(sents), ref["trees"], f"BatchParser call {k}")


def test_exhaustive_viterbi_trees(puck, wsj1, sents, layers):
    from oracle.oracle import batch
    ref = batch([layers["fine_v"]], *sents)
    p = puck.CLParser([wsj1], mem_bytes=4 << 30, persistent=True)
    _same(p.parse(sents), ref["trees"], "exhaustive")


def test_c2f_real_partitions(puck, wsj0, wsj1_real, sents, layers):
    from oracle.oracle import batch
    ref = batch([layers["coarse"], layers["fine_r"]], *sents)
    parts = puck.CLParser([wsj0, wsj1_real], mem_bytes=4 << 30).partitions(sents)
    rel = np.abs(parts.astype(np.float64) - ref["partitions"]) / np.abs(ref["partitions"].astype(np.float64))
    print(f"[real text] Real log-likelihood max rel err {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-4


def test_parse_front_end_yields(puck, tmp_path):
    from puck_amd import cli
    grz = str(tmp_path / "g.grz")
    cli.main(["compile", "--textGrammarPrefix", puck.grammar_path("wsj_0") + ":" + puck.grammar_path("wsj_1"),
              "--grammar", grz])
    src = tmp_path / "smallbank.txt"
    src.write_text(open(os.path.join(GOLD, "smallbank_sentences.txt")).read())
    cli.main(["parse", "--grammar", grz, "--mem", "4g", str(src)])
    out = (tmp_path / "smallbank.txt.parsed").read_text().splitlines()
    toks = smallbank()
    assert len(out) == len(toks)
    for line, t in zip(out, toks):
        assert line.count("(") == line.count(")") and line != "(())"
        assert re.findall(r"\(([^\s()]+) ([^\s()]+)\)", line) and \
            [w for _, w in re.findall(r"\(([^\s()]+) ([^\s()]+)\)", line)] == t
