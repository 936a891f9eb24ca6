"""C ABI: the library loads, exports every symbol include/puck.h declares, reports errors as
status codes with messages, and its host-side inputs are deterministic (CPU only)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "puck.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(puck_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(puck):
    lib = ctypes.CDLL(puck.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    # and the Python layer binds every one of them
    assert set(n for n, _, _ in puck.SIGNATURES) == set(names)


def test_error_status_and_message(puck):
    lib = puck.lib()
    h = ctypes.c_void_p()
    rc = lib.puck_grammar_load_text(b"/no/such/grammar", 0, ctypes.byref(h))
    assert rc == -2  # PUCK_ERR_IO
    assert b"cannot open" in lib.puck_last_error()
    rc = lib.puck_grammar_load_text(puck.grammar_path("wsj_0").encode(), 7, ctypes.byref(h))
    assert rc == -1 and b"semiring" in lib.puck_last_error()
    assert lib.puck_parse_viterbi(None, None, None, 0, None, 0) == -1


def test_parser_create_without_device_fails_cleanly(puck, wsj0):
    # On a host without a GPU the engine reports a device error instead of crashing.
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(puck.PuckError, match="HIP error"):
        puck.CLParser([wsj0], mem_bytes=1 << 28)


def test_rejected_configurations_fail_before_device_work(puck, wsj0):
    # Checked before any HIP call, so these run without a GPU.
    with pytest.raises(puck.PuckError, match="device-scheduled passes are disabled"):
        puck.CLParser([wsj0], mem_bytes=1 << 28, device_schedule=True)
    with pytest.raises(puck.PuckError, match="device-scheduled passes are disabled"):
        puck.BatchParser([wsj0], mem_bytes=1 << 28, device_schedule=True)
    for rt in (0, 3, 32):
        with pytest.raises(puck.PuckError, match="reduce_targets must be 1, 2, 4, 8 or 16"):
            puck.BatchParser([wsj0], mem_bytes=1 << 28, reduce_targets=rt)


def test_sampler_is_seeded_and_respects_lengths(wsj1):
    w1, o1 = wsj1.sample(seed=42, n=50, mode=0, min_len=5, max_len=20)
    w2, o2 = wsj1.sample(seed=42, n=50, mode=0, min_len=5, max_len=20)
    assert np.array_equal(w1, w2) and np.array_equal(o1, o2)
    lens = np.diff(o1)
    assert lens.min() >= 5 and lens.max() <= 20
    w3, o3 = wsj1.sample(seed=43, n=50, mode=0, min_len=5, max_len=20)
    assert not np.array_equal(w1[:20], w3[:20])
    assert w1.min() >= 0 and w1.max() < wsj1.info.num_words


def test_lognormal_and_mixed_length_modes(wsj1):
    _, o = wsj1.sample(seed=1, n=300, mode=1, min_len=1, max_len=40)
    lens = np.diff(o)
    assert lens.min() >= 1 and lens.max() <= 40
    assert 17 < lens.mean() < 27  # discretized log-normal, mean ~23 before truncation
    _, o = wsj1.sample(seed=1, n=100, mode=2, min_len=1, max_len=80)
    lens = np.diff(o)
    assert lens.max() <= 80 and (lens > 40).sum() > 10
