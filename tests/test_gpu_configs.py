"""GPU parity on each workload configuration's own inputs, through the C ABI, against the oracle's
batch driver (oracle/cky_batch.c, OpenMP over sentences):

  config 2  wsj_6-shape exhaustive Viterbi, lengths 15-40: trees and root scores bit-exact, through
            CLParser and through BatchParser with graph capture and replay
  config 3  wsj_1 -> wsj_6-shape coarse-to-fine (no reprojection): masks bit-exact, Real
            log-likelihoods within 1e-4 relative, MBR trees
  config 4  long tail, lengths 41-80, through BatchParser: trees bit-exact (coarse-to-fine, and
            exhaustive for two sentences of 41-44 words)
  config 5  the server on the wsj_1 -> wsj_6-shape pipeline under a 1 s Poisson stream at 1,000
            requests/s: every result equals the oracle's tree
  workspace a parser with a small memory budget (several batches, several partial-buffer slices)
            against the oracle, and the fine layer's scaling constants against ref_scaling
  ties      the hand-derived tie-order trees of tests/test_oracle_pins.py decoded on the GPU
"""
import os
import time

import numpy as np
import pytest

from conftest import split_sentences

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def w6(puck):
    """Synthetic wsj_6-shape grammar (Viterbi and Real) and the oracle layers for c2f."""
    from oracle.grammar_ref import RefGrammar
    from oracle.oracle import Oracle
    prefix = puck.synth_wsj6_cached(seed=6)
    t = time.time()
    refs = dict(
        prefix=prefix,
        g6=puck.Grammar.load_text(prefix, puck.VITERBI),
        g6r=puck.Grammar.load_text(prefix, puck.REAL),
        coarse=Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"), final_layer=False),
        fine_v=Oracle(RefGrammar(prefix, "viterbi")),
    )
    print(f"[w6] grammars and oracle tables in {time.time() - t:.1f} s", flush=True)
    return refs


def _fine_real(w6):
    if "fine_r" not in w6:
        from oracle.grammar_ref import RefGrammar
        from oracle.oracle import Oracle
        w6["fine_r"] = Oracle(RefGrammar(w6["prefix"], "real"))
    return w6["fine_r"]


def _same_trees(got, want, label):
    assert len(got) == len(want)
    bad = [i for i, (x, y) in enumerate(zip(got, want)) if not np.array_equal(x, y)]
    assert not bad, f"{label}: {len(bad)} of {len(want)} trees differ (first {bad[:5]})"


def _concat(*parts):
    words = np.concatenate([p[0] for p in parts]).astype(np.int32)
    offs = [np.asarray(parts[0][1])]
    for p in parts[1:]:
        offs.append(np.asarray(p[1][1:]) + offs[-1][-1])
    return words, np.concatenate(offs).astype(np.int32)


def test_config2_exhaustive_lengths_15_40(puck, w6):
    from oracle.oracle import batch
    g6 = w6["g6"]
    words, offsets = g6.sample(seed=201, n=32, mode=0, min_len=15, max_len=40)
    lens = np.diff(offsets)
    print(f"[config 2] 32 sentences, lengths {lens.min()}-{lens.max()} (mean {lens.mean():.1f})", flush=True)
    t = time.time()
    ref = batch([w6["fine_v"]], words, offsets)
    print(f"[config 2] oracle {time.time() - t:.1f} s on {ref['threads']} threads", flush=True)
    p = puck.CLParser([g6], mem_bytes=16 << 30)
    _same_trees(p.parse((words, offsets)), ref["trees"], "CLParser")
    np.testing.assert_array_equal(p.partitions((words, offsets)), ref["partitions"])
    del p
    bp = puck.BatchParser([g6], mem_bytes=24 << 30, streams=3, bucket_cells=6000)
    for k in range(3):  # eager, capture, replay
        _same_trees(bp.parse((words, offsets)), ref["trees"], f"BatchParser call {k}")
    st = bp.stats()
    assert st["buckets"] > 1 and st["graph_replays"] == st["buckets"]


def test_config3_c2f_masks_partitions_mbr(puck, wsj1, w6):
    from oracle.oracle import batch
    g6r = w6["g6r"]
    words, offsets = g6r.sample(seed=203, n=24, mode=0, min_len=1, max_len=25)
    t = time.time()
    ref = batch([w6["coarse"], _fine_real(w6)], words, offsets)
    print(f"[config 3] oracle {time.time() - t:.1f} s", flush=True)
    p = puck.CLParser([wsj1, g6r], mem_bytes=8 << 30)
    np.testing.assert_array_equal(p.masks((words, offsets)), ref["masks"])
    parts = p.partitions((words, offsets))
    rel = np.abs(parts.astype(np.float64) - ref["partitions"]) / np.abs(ref["partitions"].astype(np.float64))
    print(f"[config 3] Real log-likelihood max rel err {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-4
    mbr = p.parse((words, offsets))
    same = sum(np.array_equal(m[:, :4], r[:, :4]) for m, r in zip(mbr, ref["mbr"]))
    print(f"[config 3] MBR trees identical: {same}/{len(mbr)}", flush=True)
    for m, r in zip(mbr, ref["mbr"]):
        # The root node's score is the max-rule objective: equal within fp32 marginal rounding.
        np.testing.assert_allclose(m[0, 4:5].view(np.float32), r[0, 4:5].view(np.float32), rtol=1e-4)
    assert same >= len(mbr) - 1


def test_config4_long_tail_batched(puck, wsj1, w6):
    from oracle.oracle import batch
    g6 = w6["g6"]
    words, offsets = g6.sample(seed=204, n=8, mode=0, min_len=41, max_len=80)
    lens = np.diff(offsets)
    print(f"[config 4] lengths {sorted(lens.tolist())}", flush=True)
    t = time.time()
    ref = batch([w6["coarse"], w6["fine_v"]], words, offsets)
    print(f"[config 4] c2f oracle {time.time() - t:.1f} s", flush=True)
    short = g6.sample(seed=205, n=40, mode=0, min_len=1, max_len=20)
    mixed = _concat((words, offsets), short)
    ref_short = batch([w6["coarse"], w6["fine_v"]], *short)
    bp = puck.BatchParser([wsj1, g6], mem_bytes=24 << 30, streams=3, bucket_cells=20000)
    for k in range(3):
        got = bp.parse(mixed)
        _same_trees(got, ref["trees"] + ref_short["trees"], f"c2f BatchParser call {k}")
    w2, o2 = g6.sample(seed=206, n=2, mode=0, min_len=41, max_len=44)
    t = time.time()
    ref2 = batch([w6["fine_v"]], w2, o2)
    print(f"[config 4] exhaustive oracle, lengths {np.diff(o2).tolist()}: {time.time() - t:.1f} s", flush=True)
    bp1 = puck.BatchParser([g6], mem_bytes=24 << 30, streams=2, bucket_cells=2000)
    _same_trees(bp1.parse((w2, o2)), ref2["trees"], "exhaustive BatchParser")


def test_config5_server_poisson_c2f(puck, wsj1, w6):
    from oracle.oracle import batch
    g6 = w6["g6"]
    pool_w, pool_o = g6.sample(seed=207, n=200, mode=0, min_len=1, max_len=40)
    pool = split_sentences(pool_w, pool_o)
    t = time.time()
    ref = batch([w6["coarse"], w6["fine_v"]], pool_w, pool_o)["trees"]
    print(f"[config 5] oracle on the 200-sentence pool {time.time() - t:.1f} s", flush=True)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / 1000, size=1400))
    arrivals = arrivals[arrivals < 1.0]
    srv = puck.Server([wsj1, g6], mem_bytes=24 << 30, max_inflight=3)
    t0 = time.perf_counter()
    tickets = []
    for k, at in enumerate(arrivals):
        while time.perf_counter() - t0 < at:
            pass
        tickets.append(srv.submit(pool[k % len(pool)]))
    srv.flush()
    bad = []
    for k, tk in enumerate(tickets):
        s = pool[k % len(pool)]
        if not np.array_equal(srv.result(tk, len(s)), ref[k % len(pool)]):
            bad.append(k)
    st = srv.stats()
    print(f"[config 5] {len(tickets)} requests, {st['batches']} batches, p50 {st['p50_ms']:.1f} ms, "
          f"p99 {st['p99_ms']:.1f} ms", flush=True)
    assert not bad, f"{len(bad)} results differ from the oracle (first {bad[:5]})"
    assert st["batches"] > 1


def test_small_workspace_against_oracle_and_scales(puck, wsj1, w6):
    from oracle.oracle import batch, pipeline
    g6 = w6["g6"]
    words, offsets = g6.sample(seed=208, n=60, mode=0, min_len=10, max_len=30)
    ref = batch([w6["coarse"], w6["fine_v"]], words, offsets)
    small = puck.CLParser([wsj1, g6], mem_bytes=256 << 20)
    _same_trees(small.parse((words, offsets)), ref["trees"], "small workspace c2f")
    w1, o1 = g6.sample(seed=209, n=30, mode=0, min_len=15, max_len=30)
    ref1 = batch([w6["fine_v"]], w1, o1)
    small1 = puck.CLParser([g6], mem_bytes=256 << 20)
    _same_trees(small1.parse((w1, o1)), ref1["trees"], "small workspace exhaustive")
    # Scaling constants of a Real fine layer: computed on the device from the coarse charts.
    g6r = w6["g6r"]
    ws, os_ = g6r.sample(seed=210, n=4, mode=0, min_len=3, max_len=9)
    p = puck.CLParser([wsj1, g6r], mem_bytes=8 << 30)
    ch = p.debug_charts((ws, os_), layer=1, with_outside=False)
    recs = pipeline([w6["coarse"], _fine_real(w6)], split_sentences(ws, os_))
    np.testing.assert_array_equal(ch["in_scale"], np.concatenate([r["in_scale"] for r in recs]))
    np.testing.assert_array_equal(ch["out_scale"], np.concatenate([r["out_scale"] for r in recs]))
    assert np.any(ch["in_scale"] != 0) and np.any(ch["out_scale"] != 0)


def test_hand_derived_tie_order_on_gpu(puck, tmp_path):
    import test_oracle_pins as T
    for binary, cases in ((T.BINARY_A, [(3, T.TREE_A3)]),
                          (T.BINARY_B, [(3, T.TREE_B3), (2, T.TREE_B2), (1, T.TREE_B1)])):
        d = tmp_path / ("a" if binary is T.BINARY_A else "b")
        d.mkdir()
        g = puck.Grammar.load_text(T._write_grammar(d, "tie", binary))
        p = puck.CLParser([g], mem_bytes=1 << 30)
        a = g.word_id("a")
        sents = [np.full(n, a, np.int32) for n, _ in cases]
        trees = p.parse(sents)
        for (n, expect), got in zip(cases, trees):
            np.testing.assert_array_equal(got, T._tree(expect), err_msg=f"length {n}")
