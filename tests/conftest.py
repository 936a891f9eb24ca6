import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def puck():
    import puck_amd
    return puck_amd


@pytest.fixture(scope="session")
def wsj1(puck):
    return puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.VITERBI)


@pytest.fixture(scope="session")
def wsj1_real(puck):
    return puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.REAL)


@pytest.fixture(scope="session")
def wsj0(puck):
    return puck.Grammar.load_text(puck.grammar_path("wsj_0"), puck.VITERBI)


@pytest.fixture(scope="session")
def ref_wsj1():
    from oracle.grammar_ref import RefGrammar
    import puck_amd
    return RefGrammar(puck_amd.grammar_path("wsj_1"), "viterbi")


@pytest.fixture(scope="session")
def ref_wsj1_real():
    from oracle.grammar_ref import RefGrammar
    import puck_amd
    return RefGrammar(puck_amd.grammar_path("wsj_1"), "real")


def split_sentences(words, offsets):
    return [words[offsets[i]:offsets[i + 1]] for i in range(len(offsets) - 1)]


def tri(n):
    return n * (n + 1) // 2


def cidx(b, e, n):
    lo, s = min(b, e), abs(e - b) - 1
    return lo + s * n - s * (s - 1) // 2


def sentence_slices(offsets):
    """Per-sentence cell ranges [c0, c1) of a batch chart."""
    lens = np.diff(offsets)
    co = np.concatenate([[0], np.cumsum(lens * (lens + 1))])
    return [(int(co[i]), int(co[i + 1])) for i in range(len(lens))]
