"""Autotuner search logic on the CPU, with a stand-in engine whose device time is a known
function of the configuration: the search must find the minimum along each knob, and it must
refuse a configuration that changes the parse results."""
import types

import numpy as np
import pytest

from puck_amd import autotune


def fake_puck(cost, wrong=None):
    trees = [np.arange(7, dtype=np.int32).reshape(-1, 1)]

    seen = []

    class CLParser:
        def __init__(self, layers, mem_bytes, **kw):
            seen.append(("plain", kw))

        def parse(self, s):
            return trees

        def stats(self):
            return {"device_ms": 100.0}

    class BatchParser:
        def __init__(self, layers, mem_bytes, **cfg):
            self.cfg = cfg
            seen.append(("batched", cfg))

        def parse(self, s):
            if wrong is not None and all(self.cfg.get(k) == v for k, v in wrong.items()):
                return [t + 1 for t in trees]
            return trees

        def stats(self):
            return {"device_ms": cost(self.cfg), "buckets": 3}

    return types.SimpleNamespace(CLParser=CLParser, BatchParser=BatchParser, seen=seen)


def test_coordinate_descent_finds_separable_minimum():
    want = dict(streams=2, bucket_cells=1 << 21, piece_target=1024, items_per_lane=2, persistent=True)

    def cost(c):
        return 10.0 + sum(0.0 if c[k] == v else 1.0 + abs(hash((k, c[k]))) % 5 for k, v in want.items())

    res = autotune.tune(fake_puck(cost), [None], ([0], [0, 1]))
    assert all(res["best"][k] == v for k, v in want.items())
    assert res["best_ms"] == 10.0 and res["plain_ms"] == 100.0
    assert all(r["same_trees"] for r in res["table"])


def test_configuration_that_changes_trees_is_rejected():
    with pytest.raises(RuntimeError, match="changed the parse results"):
        autotune.tune(fake_puck(lambda c: 1.0, wrong={"streams": 1}), [None], ([0], [0, 1]))


def test_save_and_load_roundtrip(tmp_path):
    res = autotune.tune(fake_puck(lambda c: 5.0 + c["streams"]), [None], ([0], [0, 1]))
    path = tmp_path / "tune.json"
    autotune.save(str(path), res, "unit")
    back = autotune.load(str(path))
    assert back["best"]["streams"] == 1 and back["workload"] == "unit"


def test_workload_options_reach_every_parser_and_are_not_tuned():
    fake = fake_puck(lambda c: 10.0 + c["reduce_targets"])
    res = autotune.tune(fake, [None], ([0], [0, 1]), fixed={"outside_final": True})
    assert all(kw.get("outside_final") is True for _, kw in fake.seen)
    assert "outside_final" not in res["best"]
    assert res["best"]["reduce_targets"] == 1
