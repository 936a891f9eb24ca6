"""Persistent rule kernel (one work-queue launch per span for the families of a reduce group) vs
the plain form (one launch per span and family): the arithmetic per (item, rule) is the same, so
charts, masks, trees and partitions must be bit-identical; trees are also checked against the
oracle."""
import numpy as np
import pytest

from conftest import split_sentences

pytestmark = pytest.mark.gpu


def _pair(puck, layers, **kw):
    plain = puck.CLParser(layers, mem_bytes=2 << 30, **kw)
    pers = puck.CLParser(layers, mem_bytes=2 << 30, persistent=True, **kw)
    return plain, pers


def test_exhaustive_viterbi_charts_and_trees(puck, wsj1, ref_wsj1):
    from oracle.oracle import Oracle
    sents = wsj1.sample(seed=71, n=24, mode=0, min_len=2, max_len=22)
    plain, pers = _pair(puck, [wsj1])
    ca = plain.debug_charts(sents, layer=0, with_outside=True)
    cb = pers.debug_charts(sents, layer=0, with_outside=True)
    assert np.array_equal(ca["inside"], cb["inside"])
    assert np.array_equal(ca["outside"], cb["outside"])
    ta, tb = plain.parse(sents), pers.parse(sents)
    orc = Oracle(ref_wsj1)
    words, offsets = sents
    for i, s in enumerate(split_sentences(words, offsets)):
        assert np.array_equal(ta[i], tb[i]), f"sentence {i}"
        if i < 6:
            ins, _ = orc.inside(s)
            assert np.array_equal(tb[i], orc.viterbi(ins, len(s))), f"sentence {i} vs oracle"


def test_coarse_to_fine_masks_trees_partitions(puck, wsj0, wsj1, wsj1_real):
    sents = wsj1.sample(seed=72, n=40, mode=0, min_len=2, max_len=25)
    plain, pers = _pair(puck, [wsj0, wsj1])
    assert np.array_equal(plain.masks(sents), pers.masks(sents))
    for a, b in zip(plain.parse(sents), pers.parse(sents)):
        assert np.array_equal(a, b)
    plain_r, pers_r = _pair(puck, [wsj0, wsj1_real])
    assert np.array_equal(plain_r.partitions(sents), pers_r.partitions(sents))
    for a, b in zip(plain_r.parse(sents), pers_r.parse(sents)):
        assert np.array_equal(a, b)


def test_launch_shapes_give_identical_trees(puck, wsj0, wsj1):
    sents = wsj1.sample(seed=73, n=30, mode=0, min_len=2, max_len=30)
    ref = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30).parse(sents)
    p = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30)
    for persistent, target, wgs in ((True, 512, 2), (True, 64, 1), (False, 64, 1), (False, 2048, 1)):
        p.configure(persistent, piece_target=target, queue_wgs_per_cu=wgs)
        got = p.parse(sents)
        for i, (a, b) in enumerate(zip(ref, got)):
            assert np.array_equal(a, b), f"persistent={persistent} target={target} wgs={wgs} sentence {i}"


def test_reduce_workgroup_sizes_give_identical_charts(puck, wsj0, wsj1, wsj1_real):
    """The reduce kernel with 1-16 targets (waves) per workgroup folds each (cell, target) in the
    same order: inside and outside charts, masks, trees and Real partitions are bit-identical."""
    sents = wsj1.sample(seed=75, n=30, mode=0, min_len=2, max_len=28)
    ref = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30, persistent=True)
    ref_ch = ref.debug_charts(sents, layer=1, with_outside=True)
    ref_trees, ref_masks = ref.parse(sents), ref.masks(sents)
    ref_r = puck.CLParser([wsj0, wsj1_real], mem_bytes=2 << 30, persistent=True)
    ref_parts = ref_r.partitions(sents)
    for rt in (1, 2, 8, 16):
        p = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30)
        p.configure(True, reduce_targets=rt)
        ch = p.debug_charts(sents, layer=1, with_outside=True)
        assert np.array_equal(ch["inside"], ref_ch["inside"]), f"reduce_targets={rt} inside"
        assert np.array_equal(ch["outside"], ref_ch["outside"]), f"reduce_targets={rt} outside"
        assert np.array_equal(p.masks(sents), ref_masks), f"reduce_targets={rt} masks"
        for i, (a, b) in enumerate(zip(p.parse(sents), ref_trees)):
            assert np.array_equal(a, b), f"reduce_targets={rt} sentence {i}"
        pr = puck.CLParser([wsj0, wsj1_real], mem_bytes=2 << 30)
        pr.configure(True, reduce_targets=rt)
        assert np.array_equal(pr.partitions(sents), ref_parts), f"reduce_targets={rt} partitions"


def test_wsj6_shape_block_major_streams(puck, wsj1):
    """The synthetic wsj_6-shape grammar has 15-18 blocks per family: every fine binary family
    uses the block-major (skipping) layout in one shared queue launch."""
    prefix = puck.synth_wsj6_cached(seed=6)
    g6 = puck.Grammar.load_text(prefix, puck.VITERBI)
    sents = g6.sample(seed=74, n=48, mode=0, min_len=1, max_len=30)
    plain, pers = _pair(puck, [wsj1, g6])
    for i, (a, b) in enumerate(zip(plain.parse(sents), pers.parse(sents))):
        assert np.array_equal(a, b), f"sentence {i}"
