"""Folded partial rows (Viterbi layers, the default): every rule slot of a target maxes its values
into one row per target with an integer atomic max on order-preserving encodings, instead of
writing one row per (operand chunk, target) slot for the reduce (PUCK_FLAG_SLOT_PARTIALS). Binary
families fold when their rules spread over at least three operand chunks, as the synthetic
wsj_6-shape grammar's do (wsj_1's fit one chunk and keep slot rows). The maximum is exact and
independent of order, so inside and outside charts, masks and trees must be bit-identical in both
forms, plain and persistent, exhaustive and under masks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g6(puck):
    return puck.Grammar.load_text(puck.synth_wsj6_cached(seed=6), puck.VITERBI)


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


@pytest.mark.parametrize("persistent", [False, True])
def test_exhaustive_charts_identical(puck, g6, persistent):
    """Exhaustive insideOutside: split pairs folded in the rule kernel (even and >= 15 split counts),
    prefix-counted outside families."""
    sents = g6.sample(seed=81, n=16, mode=0, min_len=1, max_len=18)
    fold = puck.CLParser([g6], mem_bytes=4 << 30, persistent=persistent, outside_final=True)
    slot = puck.CLParser([g6], mem_bytes=4 << 30, persistent=persistent, outside_final=True, slot_partials=True)
    ca = fold.debug_charts(sents, layer=0, with_outside=True)
    cb = slot.debug_charts(sents, layer=0, with_outside=True)
    assert np.array_equal(_bits(ca["inside"]), _bits(cb["inside"]))
    assert np.array_equal(_bits(ca["outside"]), _bits(cb["outside"]))
    for i, (a, b) in enumerate(zip(fold.parse(sents), slot.parse(sents))):
        assert np.array_equal(a, b), f"sentence {i}"


@pytest.mark.parametrize("persistent", [False, True])
def test_coarse_to_fine_identical(puck, wsj1, g6, persistent):
    """wsj_1 -> wsj_6 shape under masks: planned items, block-major streams (targets carried in
    the rule groups), fills up to the planned item count."""
    sents = g6.sample(seed=82, n=48, mode=0, min_len=1, max_len=30)
    fold = puck.CLParser([wsj1, g6], mem_bytes=4 << 30, persistent=persistent)
    slot = puck.CLParser([wsj1, g6], mem_bytes=4 << 30, persistent=persistent, slot_partials=True)
    assert np.array_equal(fold.masks(sents), slot.masks(sents))
    ca = fold.debug_charts(sents, layer=1, with_outside=True)
    cb = slot.debug_charts(sents, layer=1, with_outside=True)
    assert np.array_equal(_bits(ca["inside"]), _bits(cb["inside"]))
    assert np.array_equal(_bits(ca["outside"]), _bits(cb["outside"]))
    for i, (a, b) in enumerate(zip(fold.parse(sents), slot.parse(sents))):
        assert np.array_equal(a, b), f"sentence {i}"


def test_batched_graphs_identical(puck, wsj1, g6):
    """The batched engine (streams, graph capture and replay) with folded rows gives the slot
    form's trees. (The folded default is checked against the oracle on every configuration's own
    workload in test_gpu_configs.py.)"""
    c2f = puck.BatchParser([wsj1, g6], mem_bytes=4 << 30, streams=2, bucket_cells=1 << 12)
    c2f_slot = puck.BatchParser([wsj1, g6], mem_bytes=4 << 30, streams=2, bucket_cells=1 << 12, slot_partials=True)
    sents2 = g6.sample(seed=84, n=48, mode=0, min_len=1, max_len=30)
    ref = c2f_slot.parse(sents2)
    for rep in range(3):  # eager, capture, replay
        for i, (x, y) in enumerate(zip(c2f.parse(sents2), ref)):
            assert np.array_equal(x, y), f"run {rep} sentence {i}"
