"""Pins the oracle's MBR decoder (CPU only): the per-cell coarse marginals against a float64
closed form over the oracle's own charts, and the max-rule tree against brute force over every
binary bracketing of short sentences (CLMBRKernels.scala:118-171, CLParser.scala:589-696)."""
import math

import numpy as np
import pytest

from conftest import cidx, split_sentences, tri
from oracle import oracle as O


@pytest.fixture(scope="module")
def real_oracle(puck):
    from oracle.grammar_ref import RefGrammar
    return O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "real"))


def _charts(orc, s):
    n = len(s)
    ins, _ = orc.inside(s)
    out = orc.outside(ins, n)
    zeros = np.zeros(n * (n + 1), np.float32)
    nodes, dec = orc.mbr(ins, out, n, zeros, zeros, False)
    return ins, out, nodes, dec


def test_mbr_cells_are_coarse_marginals(wsj1, real_oracle):
    ref = real_oracle.ref
    words, offsets = wsj1.sample(seed=31, n=5, mode=0, min_len=3, max_len=12)
    for s in split_sentences(words, offsets):
        n = len(s)
        ins, out, _, dec = _charts(real_oracle, s)
        root = float(ins[tri(n) + cidx(0, n, n), ref.root])
        for cell in range(n * (n + 1)):
            proj, nproj = (ref.proj_t, len(ref.terms)) if cell < n else (ref.proj_n, len(ref.nonterms))
            margs = np.zeros(len(ref.coarse))
            for sym in range(ref.num_syms):
                c = proj[sym] if sym < nproj else 0
                margs[c] += float(ins[cell, sym]) / root * float(out[cell, sym])
            score = dec[cell, 0:1].view(np.float32)[0]
            assert math.isclose(score, margs.max(), rel_tol=1e-4, abs_tol=1e-6), (cell, score, margs.max())
            if margs.max() > 0:
                top2 = np.sort(margs)[-2:]
                if top2[1] - top2[0] > 1e-4 * top2[1]:  # no near-tie: the argmax is determined
                    assert dec[cell, 1] == int(np.argmax(margs))


def _node_scores(dec, n):
    """Per-span node score of the max-rule DP (unscaled single layer): clamped marginals, the unary
    layer used above width 1 or when its top marginal exceeds 0.4."""
    T = tri(n)
    sc = {}
    for b in range(n):
        for e in range(b + 1, n + 1):
            top = float(dec[T + cidx(b, e, n), 0:1].view(np.float32)[0])
            bot = float(dec[cidx(b, e, n), 0:1].view(np.float32)[0])
            top, bot = (0.0 if math.isnan(top) else min(top, 1.0)), (0.0 if math.isnan(bot) else min(bot, 1.0))
            uu = (b + 1 < e) or top > 0.4
            sc[b, e] = top + bot if uu else bot
    return sc


def _best_tree(sc, b, e):
    if e - b == 1:
        return sc[b, e]
    return sc[b, e] + max(_best_tree(sc, b, s) + _best_tree(sc, s, e) for s in range(b + 1, e))


def test_mbr_tree_is_the_max_rule_tree(wsj1, real_oracle):
    words, offsets = wsj1.sample(seed=32, n=8, mode=0, min_len=2, max_len=7)
    for s in split_sentences(words, offsets):
        n = len(s)
        _, _, nodes, dec = _charts(real_oracle, s)
        sc = _node_scores(dec, n)
        best = _best_tree(sc, 0, n)
        root_score = float(nodes[0, 4:5].view(np.float32)[0])
        assert math.isclose(root_score, best, rel_tol=1e-5), (n, root_score, best)
        # the returned bracketing scores exactly what the DP says
        spans = [(int(b), int(e)) for b, e in nodes[:, 2:4]]
        assert len(spans) == 2 * n - 1 and spans[0] == (0, n)
        assert math.isclose(sum(sc[x] for x in spans), best, rel_tol=1e-5)
