"""Folded partial rows (Viterbi layers, the default): every rule slot of a target maxes its values
into one row per target with an integer atomic max on order-preserving encodings, instead of
writing one row per (operand chunk, target) slot for the reduce (PUCK_FLAG_SLOT_PARTIALS). The
maximum is exact and independent of order, so inside and outside charts, masks and trees must be
bit-identical in both forms, plain and persistent, exhaustive and under masks; trees are also
checked against the oracle."""
import numpy as np
import pytest

from conftest import split_sentences

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("persistent", [False, True])
def test_exhaustive_charts_identical(puck, wsj1, ref_wsj1, persistent):
    from oracle.oracle import Oracle
    sents = wsj1.sample(seed=81, n=40, mode=0, min_len=1, max_len=30)
    fold = puck.CLParser([wsj1], mem_bytes=2 << 30, persistent=persistent)
    slot = puck.CLParser([wsj1], mem_bytes=2 << 30, persistent=persistent, slot_partials=True)
    ca = fold.debug_charts(sents, layer=0, with_outside=True)
    cb = slot.debug_charts(sents, layer=0, with_outside=True)
    assert np.array_equal(ca["inside"].view(np.uint32), cb["inside"].view(np.uint32))
    assert np.array_equal(ca["outside"].view(np.uint32), cb["outside"].view(np.uint32))
    ta, tb = fold.parse(sents), slot.parse(sents)
    orc = Oracle(ref_wsj1)
    words, offsets = sents
    for i, s in enumerate(split_sentences(words, offsets)):
        assert np.array_equal(ta[i], tb[i]), f"sentence {i}"
        if i < 8:
            ins, _ = orc.inside(s)
            assert np.array_equal(ta[i], orc.viterbi(ins, len(s))), f"sentence {i} vs oracle"


@pytest.mark.parametrize("persistent", [False, True])
def test_coarse_to_fine_identical(puck, wsj0, wsj1, persistent):
    sents = wsj1.sample(seed=82, n=40, mode=0, min_len=1, max_len=30)
    fold = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30, persistent=persistent)
    slot = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30, persistent=persistent, slot_partials=True)
    assert np.array_equal(fold.masks(sents), slot.masks(sents))
    ca = fold.debug_charts(sents, layer=1, with_outside=True)
    cb = slot.debug_charts(sents, layer=1, with_outside=True)
    assert np.array_equal(ca["inside"].view(np.uint32), cb["inside"].view(np.uint32))
    assert np.array_equal(ca["outside"].view(np.uint32), cb["outside"].view(np.uint32))
    for i, (a, b) in enumerate(zip(fold.parse(sents), slot.parse(sents))):
        assert np.array_equal(a, b), f"sentence {i}"


def test_wsj6_shape_insideoutside_identical(puck, wsj1):
    """Many chunks per target (the wsj_6 shape), split pairs, block-major masked streams and the
    batched engine with graphs: folded and slot partials give the same trees."""
    prefix = puck.synth_wsj6_cached(seed=6)
    g6 = puck.Grammar.load_text(prefix, puck.VITERBI)
    sents = g6.sample(seed=83, n=24, mode=0, min_len=1, max_len=24)
    a = puck.CLParser([g6], mem_bytes=4 << 30, persistent=True, outside_final=True)
    b = puck.CLParser([g6], mem_bytes=4 << 30, persistent=True, outside_final=True, slot_partials=True)
    ca = a.debug_charts(sents, layer=0, with_outside=True)
    cb = b.debug_charts(sents, layer=0, with_outside=True)
    assert np.array_equal(ca["inside"].view(np.uint32), cb["inside"].view(np.uint32))
    assert np.array_equal(ca["outside"].view(np.uint32), cb["outside"].view(np.uint32))
    c2f = puck.BatchParser([wsj1, g6], mem_bytes=4 << 30, streams=2, bucket_cells=1 << 12)
    c2f_slot = puck.BatchParser([wsj1, g6], mem_bytes=4 << 30, streams=2, bucket_cells=1 << 12, slot_partials=True)
    sents2 = g6.sample(seed=84, n=48, mode=0, min_len=1, max_len=30)
    ref = c2f_slot.parse(sents2)
    for rep in range(3):  # eager, capture, replay
        for i, (x, y) in enumerate(zip(c2f.parse(sents2), ref)):
            assert np.array_equal(x, y), f"run {rep} sentence {i}"
