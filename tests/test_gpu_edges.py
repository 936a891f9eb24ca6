"""Edge cases through the C ABI: empty batches, one-word sentences, sentences at the 255-word limit
(trees bit-exact against the oracle), and rejected inputs (error status and message, the handle
still usable afterwards)."""
import numpy as np
import pytest

from conftest import split_sentences

pytestmark = pytest.mark.gpu


def _oracle_trees(ref, words, offsets):
    from oracle.oracle import Oracle, pipeline
    return [r["tree"] for r in pipeline([Oracle(ref)], split_sentences(words, offsets))]


def test_empty_batch(puck, wsj0, wsj1):
    empty = (np.zeros(0, np.int32), np.zeros(1, np.int32))
    p = puck.CLParser([wsj1], mem_bytes=1 << 30)
    assert p.parse(empty) == []
    assert p.partitions(empty).shape == (0,)
    c2f = puck.CLParser([wsj0, wsj1], mem_bytes=1 << 30)
    assert c2f.masks(empty).shape[0] == 0
    assert puck.BatchParser([wsj1], mem_bytes=1 << 30).parse(empty) == []


def test_one_word_sentences(puck, wsj1, ref_wsj1):
    words, offsets = wsj1.sample(seed=91, n=40, mode=0, min_len=1, max_len=1)
    trees = puck.CLParser([wsj1], mem_bytes=1 << 30).parse((words, offsets))
    ref = _oracle_trees(ref_wsj1, words, offsets)
    assert all(t.shape[0] == 1 for t in trees)
    for i, r in enumerate(ref):
        assert np.array_equal(trees[i], r), f"sentence {i}"


def test_sentences_at_the_length_limit(puck, wsj1, ref_wsj1):
    long_w, long_o = wsj1.sample(seed=92, n=2, mode=0, min_len=230, max_len=255)
    short_w, short_o = wsj1.sample(seed=93, n=6, mode=0, min_len=1, max_len=12)
    # long and short sentences in one batch
    words = np.concatenate([long_w, short_w])
    offsets = np.concatenate([long_o, long_o[-1] + short_o[1:]]).astype(np.int32)
    print(f"lengths {np.diff(offsets).tolist()}", flush=True)
    trees = puck.CLParser([wsj1], mem_bytes=8 << 30, persistent=True).parse((words, offsets))
    ref = _oracle_trees(ref_wsj1, words, offsets)
    for i, r in enumerate(ref):
        assert np.array_equal(trees[i], r), f"sentence {i}"


def test_rejected_inputs_leave_the_parser_usable(puck, wsj1):
    p = puck.CLParser([wsj1], mem_bytes=1 << 30)
    good_w, good_o = wsj1.sample(seed=94, n=5, mode=0, min_len=3, max_len=10)
    before = p.parse((good_w, good_o))
    nw = wsj1.info.num_words
    for words, offsets, msg in [
        (np.array([0, nw], np.int32), np.array([0, 2], np.int32), "word id out of range"),
        (np.array([-1], np.int32), np.array([0, 1], np.int32), "word id out of range"),
        (np.array([0, 1], np.int32), np.array([0, 2, 2], np.int32), "sentence length"),
        (np.zeros(256, np.int32), np.array([0, 256], np.int32), "sentence length"),
    ]:
        with pytest.raises(puck.PuckError, match=msg):
            p.parse((words, offsets))
    after = p.parse((good_w, good_o))
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_batch_engine_rejects_bad_input_and_stays_usable(puck, wsj1):
    words, offsets = wsj1.sample(seed=95, n=30, mode=0, min_len=2, max_len=20)
    bp = puck.BatchParser([wsj1], mem_bytes=2 << 30, bucket_cells=256)
    before = bp.parse((words, offsets))
    bad = words.copy()
    bad[3] = wsj1.info.num_words
    with pytest.raises(puck.PuckError, match="word id out of range"):
        bp.parse((bad, offsets))
    after = bp.parse((words, offsets))
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
