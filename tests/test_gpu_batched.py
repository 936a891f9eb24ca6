"""Batched engine and HIP graphs vs the plain form. Length buckets on concurrent streams,
persistent rule launches and per-shape graphs change only the schedule, so every tree must be
bit-identical to CLParser.parse's; graph replays are interleaved with eager work of other shapes
and each replay is checked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.fixture(scope="module")
def mixed(wsj1):
    # Mixed lengths: short and long sentences in one call.
    w1, o1 = wsj1.sample(seed=81, n=60, mode=0, min_len=1, max_len=12)
    w2, o2 = wsj1.sample(seed=82, n=20, mode=0, min_len=20, max_len=40)
    words = np.concatenate([w1, w2]).astype(np.int32)
    offsets = np.concatenate([o1, o2[1:] + o1[-1]]).astype(np.int32)
    return words, offsets


def test_batched_viterbi_matches_plain_across_graph_replays(puck, wsj0, wsj1, mixed):
    plain = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30)
    ref = plain.parse(mixed)
    other = wsj1.sample(seed=83, n=25, mode=0, min_len=3, max_len=18)
    ref_other = plain.parse(other)
    bp = puck.BatchParser([wsj0, wsj1], mem_bytes=3 << 30, streams=3, bucket_cells=4000)
    seen = []
    for k, (sents, want) in enumerate([(mixed, ref), (other, ref_other), (mixed, ref), (other, ref_other),
                                       (mixed, ref), (mixed, ref)]):
        got = bp.parse(sents)
        assert _same(got, want), f"call {k}"
        seen.append(bp.stats())
    assert seen[0]["buckets"] > 3
    assert seen[2]["graph_captures"] > 0
    assert seen[4]["graph_replays"] == seen[4]["buckets"]
    assert seen[5]["graph_replays"] == seen[5]["buckets"]


def test_batched_mbr_matches_plain(puck, wsj0, wsj1_real, mixed):
    plain = puck.CLParser([wsj0, wsj1_real], mem_bytes=2 << 30)
    ref = plain.parse(mixed)
    bp = puck.BatchParser([wsj0, wsj1_real], mem_bytes=3 << 30, streams=2, bucket_cells=6000)
    for k in range(3):
        assert _same(bp.parse(mixed), ref), f"call {k}"


def test_parser_graphs_replay_and_new_shapes(puck, wsj0, wsj1, mixed):
    plain = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30)
    g = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30, persistent=True, graphs=True)
    ref = plain.parse(mixed)
    for k in range(4):
        assert _same(g.parse(mixed), ref), f"replay {k}"
        other = wsj1.sample(seed=90 + k, n=15, mode=0, min_len=2, max_len=20)
        assert _same(g.parse(other), plain.parse(other)), f"eager {k}"


def test_batched_exhaustive_single_layer(puck, wsj1, mixed):
    ref = puck.CLParser([wsj1], mem_bytes=2 << 30).parse(mixed)
    bp = puck.BatchParser([wsj1], mem_bytes=3 << 30, streams=4, bucket_cells=3000, persistent=False)
    for k in range(3):
        assert _same(bp.parse(mixed), ref), f"call {k}"


def test_grz_layers_parse_identically(puck, wsj0, wsj1, mixed, tmp_path):
    path = str(tmp_path / "grammar.grz")
    puck.save_grz(path, [wsj0, wsj1])
    loaded = puck.Grammar.load_grz(path)
    a = puck.CLParser([wsj0, wsj1], mem_bytes=2 << 30).parse(mixed)
    b = puck.CLParser(loaded, mem_bytes=2 << 30).parse(mixed)
    assert _same(a, b)


def test_cli_compile_and_parse(puck, tmp_path):
    from puck_amd import cli
    grz = str(tmp_path / "g.grz")
    cli.main(["compile", "--textGrammarPrefix", puck.grammar_path("wsj_0") + ":" + puck.grammar_path("wsj_1"),
              "--grammar", grz])
    src = tmp_path / "in.txt"
    long_line = " ".join(["the"] * 60)
    src.write_text("The company said it expects higher profits .\n\nInvestors were flurbing again .\n"
                   + long_line + "\nShe sold shares .\n")
    cli.main(["parse", "--grammar", grz, "--mem", "2g", str(src)])
    out = (tmp_path / "in.txt.parsed").read_text().splitlines()
    assert len(out) == 4
    assert out[2] == "(())"  # maxLength 50
    for k in (0, 1, 3):
        assert out[k].startswith("(") and out[k].count("(") == out[k].count(")") and out[k] != "(())"
    assert "company" in out[0] and "flurbing" in out[1]
