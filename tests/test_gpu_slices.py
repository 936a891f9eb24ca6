"""Workspace limits change the schedule, never the result: a parser with a small memory budget splits
the sentences into several batches and every span into several partial-buffer slices (each with its
own mask plan); its trees must equal, bit for bit, those of a parser that does neither.
Synthetic wsj_6-shape grammar (608-float cells, 16-17 blocks per family), exhaustive and
coarse-to-fine with wsj_1 masks."""
import numpy as np
import pytest


pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wsj6(puck):
    return puck.Grammar.load_text(puck.synth_wsj6_cached(seed=6), puck.VITERBI)


def _same(a, b):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"sentence {i}"


def test_small_workspace_same_trees_exhaustive(puck, wsj6):
    words, offsets = wsj6.sample(seed=11, n=60, mode=0, min_len=15, max_len=30)
    big = puck.CLParser([wsj6], mem_bytes=32 << 30).parse((words, offsets))
    small = puck.CLParser([wsj6], mem_bytes=256 << 20).parse((words, offsets))
    _same(big, small)


def test_small_workspace_same_trees_coarse_to_fine(puck, wsj1, wsj6):
    words, offsets = wsj6.sample(seed=12, n=150, mode=0, min_len=20, max_len=40)
    big_p = puck.CLParser([wsj1, wsj6], mem_bytes=32 << 30)
    big = big_p.parse((words, offsets))
    small_p = puck.CLParser([wsj1, wsj6], mem_bytes=256 << 20)
    small = small_p.parse((words, offsets))
    _same(big, small)
    # The small workspace really took several slices: more rule launches for the same spans.
    assert len(small_p.launches()[0]) > len(big_p.launches()[0])
