"""Serving: results through the continuous-batching server equal the result of parsing each
sentence alone, whatever mix of requests the server batched it with (Viterbi: bit-exact)."""
import threading

import numpy as np
import pytest

from conftest import split_sentences

pytestmark = pytest.mark.gpu


def test_server_result_same_alone_and_in_the_mix(puck, wsj1):
    words, offsets = wsj1.sample(seed=77, n=120, mode=0, min_len=1, max_len=30)
    sents = split_sentences(words, offsets)
    alone = puck.CLParser([wsj1], mem_bytes=1 << 30)
    want = [alone.parse([s])[0] for s in sents[:12]] + alone.parse((words, offsets))[12:]
    srv = puck.Server([wsj1], mem_bytes=2 << 30, max_inflight=3, max_batch=16)
    tickets = [None] * len(sents)

    def producer(lo, hi):
        for i in range(lo, hi):
            tickets[i] = srv.submit(sents[i])

    threads = [threading.Thread(target=producer, args=(k * 40, (k + 1) * 40)) for k in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i, s in enumerate(sents):
        got = srv.result(tickets[i], len(s))
        assert np.array_equal(got, want[i]), f"request {i}"
    st = srv.stats()
    assert st["completed"] == len(sents)
    assert st["batches"] > 1 and st["p99_ms"] >= st["p50_ms"] > 0


def test_server_coarse_to_fine(puck, wsj0, wsj1):
    words, offsets = wsj1.sample(seed=78, n=30, mode=0, min_len=3, max_len=20)
    alone = puck.CLParser([wsj0, wsj1], mem_bytes=1 << 30)
    want = alone.parse((words, offsets))
    srv = puck.Server([wsj0, wsj1], mem_bytes=2 << 30, max_inflight=2, max_batch=7)
    sents = split_sentences(words, offsets)
    tickets = [srv.submit(s) for s in sents]
    srv.flush()
    for i, s in enumerate(sents):
        assert np.array_equal(srv.result(tickets[i], len(s), wait=False), want[i])


def test_server_length_classes_short_first(puck, wsj0, wsj1):
    """A slot parses its batch's sentences of at most 40 words, then the longer ones: every result is
    still the one the sentence gets alone, in Viterbi and in MBR."""
    words, offsets = wsj1.sample(seed=79, n=48, mode=2, min_len=1, max_len=70)
    lens = np.diff(offsets)
    assert (lens <= 40).any() and (lens > 40).any()
    sents = split_sentences(words, offsets)
    for layers in ([wsj0, wsj1], [wsj0, puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.REAL)]):
        want = puck.CLParser(layers, mem_bytes=2 << 30).parse((words, offsets))
        srv = puck.Server(layers, mem_bytes=4 << 30, max_inflight=1, max_batch=1000)
        tickets = [srv.submit(s) for s in sents]
        srv.flush()
        for i, s in enumerate(sents):
            np.testing.assert_array_equal(srv.result(tickets[i], len(s), wait=False), want[i], err_msg=f"request {i}")
        st = srv.stats()
        assert st["completed"] == len(sents) and st["batches"] >= 2


def test_server_rejects_bad_requests_at_submit(puck, wsj1):
    """A request with an unknown word id or a bad length is refused by submit, so it cannot fail the
    batch it would have joined; the requests around it complete normally."""
    words, offsets = wsj1.sample(seed=79, n=12, mode=0, min_len=3, max_len=15)
    want = puck.CLParser([wsj1], mem_bytes=1 << 30).parse((words, offsets))
    srv = puck.Server([wsj1], mem_bytes=2 << 30, max_inflight=2)
    sents = split_sentences(words, offsets)
    tickets = []
    for i, s in enumerate(sents):
        tickets.append(srv.submit(s))
        if i == 5:
            with pytest.raises(puck.PuckError, match="word id out of range"):
                srv.submit(np.array([0, wsj1.info.num_words, 1], np.int32))
            with pytest.raises(puck.PuckError, match="sentence length"):
                srv.submit(np.zeros(256, np.int32))
    srv.flush()
    for i, s in enumerate(sents):
        assert np.array_equal(srv.result(tickets[i], len(s)), want[i]), f"request {i}"


def test_server_ticket_errors_and_tight_memory_budget(puck, wsj1):
    """A ticket that was never issued, or was already collected, is an error (not a wait). Under a
    tight memory budget the slots admit smaller batches and every result is unchanged."""
    words, offsets = wsj1.sample(seed=80, n=40, mode=0, min_len=5, max_len=30)
    want = puck.CLParser([wsj1], mem_bytes=1 << 30).parse((words, offsets))
    srv = puck.Server([wsj1], mem_bytes=384 << 20, max_inflight=2)
    sents = split_sentences(words, offsets)
    tickets = [srv.submit(s) for s in sents]
    with pytest.raises(puck.PuckError, match="unknown or already collected"):
        srv.result(10 ** 9, 3)
    for i, s in enumerate(sents):
        assert np.array_equal(srv.result(tickets[i], len(s)), want[i]), f"request {i}"
    with pytest.raises(puck.PuckError, match="unknown or already collected"):
        srv.result(tickets[0], len(sents[0]))
    assert srv.stats()["completed"] == len(sents)
