"""Real sentences: the word sequences of the reference's own test treebank (smallbank dev + test,
68 Wall Street Journal sentences of 8-45 tokens; tests/golden/make_smallbank.py) through the text
front end on the CPU. Every token maps to a lexicon entry of wsj_1 (the word itself or an UNK-*
signature), valid in wsj_0 too, so the coarse-to-fine pair reads them; the oracle parses the short
ones into trees that span their sentences."""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def smallbank():
    with open(os.path.join(GOLD, "smallbank_sentences.txt")) as f:
        return [line.split() for line in f if line.strip()]


def smallbank_ids(grammar):
    from puck_amd import cli
    return [cli.word_ids(grammar, toks) for toks in smallbank()]


def test_fixture_shape():
    sents = smallbank()
    lens = [len(s) for s in sents]
    assert len(sents) == 68 and min(lens) == 8 and max(lens) == 45


def test_every_token_maps_to_both_lexicons(wsj0, wsj1):
    sents = smallbank()
    ids = smallbank_ids(wsj1)
    assert all(w is not None for w in ids)
    assert all(max(w) < wsj0.info.num_words for w in ids)
    unknown = sum(wsj1.word_id(t) < 0 for s in sents for t in s)
    total = sum(len(s) for s in sents)
    print(f"{unknown} of {total} tokens take an UNK signature")
    assert unknown < total // 4
    # A known word keeps its own entry.
    assert ids[0][1] == wsj1.word_id("economy")


def test_oracle_parses_short_sentences(puck, wsj1):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle, batch
    ids = smallbank_ids(wsj1)
    short = [np.asarray(w, np.int32) for w in ids if len(w) <= 12]
    assert len(short) >= 3
    words = np.concatenate(short)
    offsets = np.concatenate([[0], np.cumsum([len(w) for w in short])]).astype(np.int32)
    coarse = Oracle(RefGrammar(puck.grammar_path("wsj_0"), "viterbi"), final_layer=False)
    fine = Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"))
    ref = batch([coarse, fine], words, offsets)
    for w, t in zip(short, ref["trees"]):
        n = len(w)
        assert t.shape == (2 * n - 1, 4)
        assert t[0, 2] == n  # the root node (top, bot, width, score) spans the sentence
        assert sum(int(x[2]) == 1 for x in t) == n  # one preterminal node per word
