"""The reference's own parser tests (CLParserTest.scala:14-31) on this engine: partitions of the
first 100 train sentences of at most 15 tokens of its test treebank (tests/golden/
smallbank_train15.txt, ParserTestHarness.getTrainTrees().take(100)) under a 40 MiB allocation
budget, with one layer ("simple test") and with the same grammar as coarse and fine layer ("pruning
test"). The reference asserts only that no partition is infinite or NaN; here they also equal the
oracle's bit for bit. Its grammar is trained from the treebank by epic, which is not in the
checkout, so wsj_1 stands in."""
import os

import numpy as np
import pytest

from test_realtext import GOLD

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def train15(wsj1):
    from puck_amd import cli
    with open(os.path.join(GOLD, "smallbank_train15.txt")) as f:
        ids = [cli.word_ids(wsj1, line.split()) for line in f if line.strip()]
    assert len(ids) == 100 and all(w is not None and len(w) <= 15 for w in ids)
    words = np.concatenate([np.asarray(w, np.int32) for w in ids])
    offsets = np.concatenate([[0], np.cumsum([len(w) for w in ids])]).astype(np.int32)
    return words, offsets


def _oracle(puck, coarse_too):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle
    fine = Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"))
    if not coarse_too:
        return [fine]
    return [Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"), final_layer=False), fine]


def test_simple(puck, wsj1, train15):
    from oracle.oracle import batch
    parts = puck.CLParser([wsj1], mem_bytes=40 << 20).partitions(train15)
    assert np.isfinite(parts).all()
    np.testing.assert_array_equal(parts, batch(_oracle(puck, False), *train15)["partitions"])


def test_pruning(puck, wsj1, train15):
    from oracle.oracle import batch
    p = puck.CLParser([wsj1, wsj1], mem_bytes=40 << 20)
    parts = p.partitions(train15)
    assert np.isfinite(parts).all()
    ref = batch(_oracle(puck, True), *train15)
    np.testing.assert_array_equal(parts, ref["partitions"])
    np.testing.assert_array_equal(p.masks(train15), ref["masks"])
    trees = p.parse(train15)
    assert all(np.array_equal(t, r) for t, r in zip(trees, ref["trees"]))
