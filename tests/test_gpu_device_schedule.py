"""Device-scheduled passes (PUCK_FLAG_DEVICE_SCHEDULE): one persistent launch per pass pulls the
rule and reduce units of all spans from a device queue, a sentence's span starting once its shorter
spans are reduced (per-block counters). The kernels are the per-span launches' own, so every chart,
tree and partition must be bit-identical to theirs; trees are also checked against the oracle."""
import time

import numpy as np
import pytest

# Both GPU runs of this file's first case (wsj_1, 90 sentences) did not return within 150 s, the
# second after a change that removed the pass's whole-L2 fences; no cause has been shown, so the
# library rejects the flag and these parity tests stay skipped until one is.
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="device-scheduled passes disabled: the launch did not finish on the "
                                      "GPU and no cause has been shown")]


def _timed(label, fn):
    t = time.time()
    out = fn()
    print(f"[device schedule] {label}: {time.time() - t:.2f} s", flush=True)
    return out


def _same(a, b, label):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"{label}: sentence {i}"


@pytest.fixture(scope="module")
def g6(puck):
    return puck.Grammar.load_text(puck.synth_wsj6_cached(seed=6), puck.VITERBI)


@pytest.mark.parametrize("persistent", [False, True])
def test_wsj1_charts_trees_partitions_identical(puck, wsj1, persistent):
    sents = wsj1.sample(seed=301, n=90, mode=0, min_len=1, max_len=35)
    ref = puck.CLParser([wsj1], mem_bytes=2 << 30, persistent=persistent)
    dev = puck.CLParser([wsj1], mem_bytes=2 << 30, persistent=persistent, device_schedule=True)
    want = _timed("per-span parse", lambda: ref.parse(sents))
    _same(_timed("device-scheduled parse", lambda: dev.parse(sents)), want, "trees")
    np.testing.assert_array_equal(dev.partitions(sents), ref.partitions(sents))
    small = wsj1.sample(seed=302, n=12, mode=0, min_len=2, max_len=20)
    a = dev.debug_charts(small, layer=0, with_outside=True)
    b = ref.debug_charts(small, layer=0, with_outside=True)
    np.testing.assert_array_equal(a["inside"], b["inside"])
    np.testing.assert_array_equal(a["outside"], b["outside"])
    # the device-scheduled launches ran: one per pass
    assert len(dev.launches()[0]) <= 2


def test_wsj6_exhaustive_inside_outside_against_oracle(puck, g6):
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle, batch
    words, offsets = g6.sample(seed=303, n=24, mode=0, min_len=5, max_len=32)
    ref = batch([Oracle(RefGrammar(puck.synth_wsj6_cached(seed=6), "viterbi"))], words, offsets)
    dev = puck.CLParser([g6], mem_bytes=16 << 30, persistent=True, device_schedule=True, outside_final=True)
    for k in range(2):
        _same(_timed(f"wsj6 inside-outside call {k}", lambda: dev.parse((words, offsets))), ref["trees"], f"call {k}")
    plain = puck.CLParser([g6], mem_bytes=16 << 30, outside_final=True)
    a = dev.debug_charts((words[:offsets[3]], offsets[:4]), layer=0, with_outside=True)
    b = plain.debug_charts((words[:offsets[3]], offsets[:4]), layer=0, with_outside=True)
    np.testing.assert_array_equal(a["inside"], b["inside"])
    np.testing.assert_array_equal(a["outside"], b["outside"])


def test_coarse_to_fine_viterbi_and_mbr_identical(puck, wsj1, g6):
    sents = g6.sample(seed=304, n=80, mode=0, min_len=1, max_len=40)
    ref = puck.CLParser([wsj1, g6], mem_bytes=16 << 30, persistent=True)
    dev = puck.CLParser([wsj1, g6], mem_bytes=16 << 30, persistent=True, device_schedule=True)
    _same(_timed("c2f parse", lambda: dev.parse(sents)), ref.parse(sents), "c2f viterbi")
    np.testing.assert_array_equal(dev.masks(sents), ref.masks(sents))
    g6r = puck.Grammar.load_text(puck.synth_wsj6_cached(seed=6), puck.REAL)
    refr = puck.CLParser([wsj1, g6r], mem_bytes=16 << 30, persistent=True)
    devr = puck.CLParser([wsj1, g6r], mem_bytes=16 << 30, persistent=True, device_schedule=True)
    _same(devr.parse(sents), refr.parse(sents), "c2f mbr")
    np.testing.assert_array_equal(devr.partitions(sents), refr.partitions(sents))


def test_batched_engine_with_graphs(puck, wsj1, g6):
    words, offsets = g6.sample(seed=305, n=150, mode=2, min_len=1, max_len=60)
    ref = puck.CLParser([wsj1, g6], mem_bytes=16 << 30).parse((words, offsets))
    bp = puck.BatchParser([wsj1, g6], mem_bytes=24 << 30, streams=3, bucket_cells=20000, device_schedule=True)
    for k in range(3):
        _same(_timed(f"batched call {k}", lambda: bp.parse((words, offsets))), ref, f"call {k}")
    st = bp.stats()
    assert st["graph_replays"] == st["buckets"]
    one = puck.CLParser([g6], mem_bytes=16 << 30).parse((words[:offsets[20]], offsets[:21]))
    bp1 = puck.BatchParser([g6], mem_bytes=24 << 30, streams=2, bucket_cells=5000, device_schedule=True,
                           outside_final=True)
    for k in range(3):
        _same(bp1.parse((words[:offsets[20]], offsets[:21])), one, f"exhaustive call {k}")
