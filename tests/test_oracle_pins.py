"""Pins the oracle functions that no brute-force sum pins by itself (CPU only):

* ref_scaling against the closed form of getScalingConstants (CLScalingKernels.scala:51-69) on
  hand-built charts, including the cutoff at -9000 and NaN entries;
* the Viterbi decode's tie order against trees derived by hand from the viterbi kernel
  (CLViterbi.scala:383-425,462-568: candidates NT, TN, TT (width 2), then NN over ascending splits,
  '>=' so the last best candidate wins; unary candidates likewise) on grammars where every
  derivation scores exactly 0;
* the max-plus outside pass by the identity max_t in(i,t) + out(i,t) = root at every position, and
  equality on every cell of the Viterbi tree;
* the pruning masks (CLMaskKernels.scala:172-236) against max-marginals computed by brute force:
  inside and outside evaluated separately on every binary bracketing of sentences of up to four
  words, with no shared chart.
"""
import math
import os

import numpy as np
import pytest

from conftest import cidx, split_sentences, tri
from oracle import oracle as O


# -- scaling constants ---------------------------------------------------------------------------
def test_scaling_closed_form(ref_wsj1):
    orc = O.Oracle(ref_wsj1)
    cs = ref_wsj1.cell_size
    rng = np.random.default_rng(3)
    rows = []
    expect = []

    def row(vals, want):
        r = np.full(cs, -np.inf, np.float32)
        r[:len(vals)] = vals
        rows.append(r)
        expect.append(want)

    row([], 0.0)                                   # nothing reachable: max -1e5 < -9000 -> 0
    row([-9000.0, -12000.0], -9000.0)              # exactly at the cutoff: kept
    row([-9000.5], 0.0)                            # just below: 0
    row([-150000.0], 0.0)                          # below the -100000 start value: 0
    row([3.5, -1.0, 2.0], 3.5)                     # positive maximum kept
    row([np.nan, 1.0, np.nan], 1.0)                # NaN entries never win
    row([-8999.75, -20.0], -20.0)                  # ordinary maximum
    for _ in range(40):
        v = rng.uniform(-12000, 50, size=cs).astype(np.float32)
        v[rng.random(cs) < 0.5] = -np.inf
        m = float(np.max(v))
        rows.append(v)
        expect.append(m if m >= -9000.0 else 0.0)
    got = orc.scaling(np.stack(rows))
    np.testing.assert_array_equal(got, np.array(expect, np.float32))


# -- Viterbi tie order ---------------------------------------------------------------------------
NUMSTATES = "ROOT 1\nS 2\nX 1\n"
UNARY = ("ROOT_0 ROOT_0 1.0\nROOT_0 S_0 1.0\nROOT_0 S_1 1.0\nS_0 S_0 1.0\nS_1 S_1 1.0\n"
         "S_0 X_0 1.0\nS_1 X_0 1.0\n")
# Grammar A: every family present for S_1 (two NN rules), so the last NN rule at the last split wins.
BINARY_A = ("S_0 S_0 S_0 1.0\nS_1 S_0 S_1 1.0\nS_1 S_1 S_0 1.0\nS_1 S_0 X_0 1.0\nS_1 X_0 S_0 1.0\n"
            "S_0 X_0 X_0 1.0\nS_1 X_0 X_0 1.0\n")
# Grammar B: no NN rule for S_1, so TN beats NT at width 3 and TT beats NT and TN at width 2.
BINARY_B = "S_0 S_0 S_0 1.0\nS_1 S_0 X_0 1.0\nS_1 X_0 S_0 1.0\nS_0 X_0 X_0 1.0\nS_1 X_0 X_0 1.0\n"

# Symbol ids: nonterminals ROOT_0 = 0, S_0 = 1, S_1 = 2; the one terminal X_0 = 0. Nodes are
# (top, bot, width, score) in pre-order; every score is exactly 0.0.
#
# A, "a a a": root unary scans ROOT->ROOT (-inf), ROOT->S_0 (0), ROOT->S_1 (0): S_1 (last). S_1 over
# (0,3): NT (0), TN (0), NN split 1 rules (S_0 S_1), (S_1 S_0), NN split 2 the same: the last one,
# S_1 -> S_1 S_0 at split 2. Node 1 = (0,2) top S_1: S_1 -> S_1 S_0 again at its only split; the
# leaves take X_0 through their unary-T rule.
TREE_A3 = [(0, 2, 3), (2, 2, 2), (2, 0, 1), (1, 0, 1), (1, 0, 1)]
# B, "a a a": root bottom S_1 again; NT (split 2) then TN (split 1): TN wins, left child is the
# terminal X_0 (top -1), right child (1,3) top S_0. S_0 over (1,3): TT (0) then NN split 2
# (S_0 S_0): NN wins, two leaves under S_0.
TREE_B3 = [(0, 2, 3), (-1, 0, 1), (1, 1, 2), (1, 0, 1), (1, 0, 1)]
# B, "a a": S_1 over (0,2): NT, TN, TT all 0, no NN: TT wins, both children terminals.
TREE_B2 = [(0, 2, 2), (-1, 0, 1), (-1, 0, 1)]
# B, "a": ROOT has no unary-T rule, so the scan keeps its start (sym 0) and the node takes the
# terminal score of X_0 at position 0.
TREE_B1 = [(0, 0, 1)]


def _write_grammar(d, name, binary):
    prefix = os.path.join(str(d), name)
    for ext, text in ((".numstates", NUMSTATES), (".binary", binary), (".unary", UNARY),
                      (".lexicon", "X a [1.0]\n")):
        with open(prefix + ext, "w") as f:
            f.write(text)
    return prefix


def _tree(nodes):
    return np.array([(t, b, w, 0) for t, b, w in nodes], np.int32)


@pytest.mark.parametrize("binary,sentences", [(BINARY_A, [(3, TREE_A3)]),
                                              (BINARY_B, [(3, TREE_B3), (2, TREE_B2), (1, TREE_B1)])])
def test_viterbi_tie_order_hand_derived(tmp_path, binary, sentences):
    from oracle.grammar_ref import RefGrammar
    ref = RefGrammar(_write_grammar(tmp_path, "tie", binary), "viterbi")
    assert ref.nonterms == ["ROOT_0", "S_0", "S_1"] and ref.terms == ["X_0"]
    orc = O.Oracle(ref)
    a = ref.word_ids["a"]
    for n, expect in sentences:
        ins, _ = orc.inside([a] * n)
        tree = orc.viterbi(ins, n)
        np.testing.assert_array_equal(tree, _tree(expect), err_msg=f"length {n}")


def test_viterbi_tie_order_product_grammar_agrees(tmp_path, puck):
    """The product's grammar compiler reads the same hand grammar into the same symbol ids and
    decoder order (the GPU decode is checked against these trees in test_gpu_configs.py)."""
    prefix = _write_grammar(tmp_path, "tie", BINARY_A)
    g = puck.Grammar.load_text(prefix)
    assert (g.info.num_nonterms, g.info.num_terms, g.info.root) == (3, 1, 0)
    assert g.word_id("a") == 0


# -- max-plus outside identity -------------------------------------------------------------------
def test_maxplus_inside_outside_identity(wsj1, ref_wsj1):
    """Max-plus: every derivation passes through one tag per position, so max_t in(i,t) + out(i,t)
    equals the best derivation's score at every position; every cell on the Viterbi tree attains
    it (top cell at the node's top symbol, bottom cell at its bottom symbol) and no cell exceeds it."""
    orc = O.Oracle(ref_wsj1)
    words, offsets = wsj1.sample(seed=21, n=6, mode=0, min_len=2, max_len=14)
    for s in split_sentences(words, offsets):
        n = len(s)
        ins, _ = orc.inside(s, f64=True)
        out = orc.outside(ins, n, f64=True)
        root = ins[tri(n) + cidx(0, n, n), ref_wsj1.root]
        tot = ins + out
        for i in range(n):
            assert math.isclose(np.max(tot[cidx(i, i + 1, n)]), root, rel_tol=1e-12)
        assert np.max(tot) <= root + 1e-9 * abs(root)
        ins32, _ = orc.inside(s)
        tree = orc.viterbi(ins32, n)
        begin, stack = 0, []
        # Walk the pre-order nodes with their spans.
        spans = []
        pos = [0]

        def walk(p, b):
            w = tree[p, 2]
            spans.append((p, b, b + w))
            if w == 1:
                return p + 1
            q = walk(p + 1, b)
            # the right child follows the left subtree (2 * left width - 1 nodes)
            return walk(q, b + (q - p - 1 + 1) // 2)

        walk(0, 0)
        assert len(spans) == 2 * n - 1
        for p, b, e in spans:
            top, bot = int(tree[p, 0]), int(tree[p, 1])
            if top != -1:
                assert math.isclose(tot[tri(n) + cidx(b, e, n), top], root, rel_tol=1e-12)
            assert math.isclose(tot[cidx(b, e, n), bot], root, rel_tol=1e-12)


# -- masks by brute force over bracketings -------------------------------------------------------
def _shapes(b, e):
    if e - b == 1:
        yield ("leaf", b)
        return
    for s in range(b + 1, e):
        for l in _shapes(b, s):
            for r in _shapes(s, e):
                yield ("node", l, r, b, e)


def _span(t):
    return (t[1], t[1] + 1) if t[0] == "leaf" else (t[3], t[4])


def _shape_maxmarginals(ref, words, shape, mm):
    """Inside and outside of one bracketing in max-plus (FP64), with the chart's unary layering:
    a leaf's bottom holds tags, its top the unary-T closure; a node's bottom the binary rules over
    its children (a leaf child enters T families as its tag or N families as its top), its top the
    unary-N layer. Folds in + out of every node into mm[(half, b, e)] by max."""
    nN, nT = len(ref.nonterms), len(ref.terms)
    neg = -np.inf
    q = {f: ref.binaries[f]["q"].astype(np.float64) for f in range(4)}
    qn, qt = ref.unary_n["q"].astype(np.float64), ref.unary_t["q"].astype(np.float64)
    ins = {}

    def inside(t):
        if t[0] == "leaf":
            bot = np.full(nT, neg)
            w = words[t[1]]
            for k in range(ref.lex_off[w], ref.lex_off[w + 1]):
                bot[ref.lex_tag[k]] = float(np.float32(ref.lex_logp[k]))
            top = np.full(nN, neg)
            np.maximum.at(top, ref.unary_t["p"], bot[ref.unary_t["c"]] + qt)
        else:
            lb, lt = inside(t[1])
            rb, rt = inside(t[2])
            ll, rl = t[1][0] == "leaf", t[2][0] == "leaf"
            bot = np.full(nN, neg)
            for f, L, R in ((0, lt, rt), (1, lt, rb), (2, lb, rt), (3, lb, rb)):
                if (f in (2, 3) and not ll) or (f in (1, 3) and not rl):
                    continue
                B = ref.binaries[f]
                np.maximum.at(bot, B["p"], (L[B["l"]] + R[B["r"]]) + q[f])
            top = np.full(nN, neg)
            np.maximum.at(top, ref.unary_n["p"], bot[ref.unary_n["c"]] + qn)
        ins[id(t)] = (bot, top)
        return bot, top

    inside(shape)
    root_top = ins[id(shape)][1][ref.root]

    def outside(t, out_top, out_bot_extra):
        bot, top = ins[id(t)]
        leaf = t[0] == "leaf"
        U, qu = (ref.unary_t, qt) if leaf else (ref.unary_n, qn)
        out_bot = np.full(len(bot), neg) if out_bot_extra is None else out_bot_extra.copy()
        np.maximum.at(out_bot, U["c"], out_top[U["p"]] + qu)
        b, e = _span(t)
        for half, iv, ov in (("top", top, out_top), ("bot", bot, out_bot)):
            key = (half, b, e)
            v = iv + ov
            mm[key] = v if key not in mm else np.maximum(mm[key], v)
        if leaf:
            return
        lb, lt = ins[id(t[1])]
        rb, rt = ins[id(t[2])]
        ll, rl = t[1][0] == "leaf", t[2][0] == "leaf"
        ol_top, or_top = np.full(nN, neg), np.full(nN, neg)
        ol_bot = np.full(nT, neg) if ll else None
        or_bot = np.full(nT, neg) if rl else None
        for f in range(4):
            lterm, rterm = f in (2, 3), f in (1, 3)
            if (lterm and not ll) or (rterm and not rl):
                continue
            B = ref.binaries[f]
            L, R = (lb if lterm else lt), (rb if rterm else rt)
            np.maximum.at(ol_bot if lterm else ol_top, B["l"], (out_bot[B["p"]] + R[B["r"]]) + q[f])
            np.maximum.at(or_bot if rterm else or_top, B["r"], (L[B["l"]] + out_bot[B["p"]]) + q[f])
        outside(t[1], ol_top, ol_bot)
        outside(t[2], or_top, or_bot)

    ot = np.full(nN, neg)
    ot[ref.root] = 0.0
    outside(shape, ot, None)
    return root_top


@pytest.mark.parametrize("name", ["wsj_0", "wsj_1"])
def test_masks_brute_force_over_bracketings(puck, name):
    from oracle.grammar_ref import RefGrammar
    g = puck.Grammar.load_text(puck.grammar_path(name))
    ref = RefGrammar(puck.grammar_path(name), "viterbi")
    orc = O.Oracle(ref, final_layer=False)
    words, offsets = g.sample(seed=17, n=8, mode=0, min_len=1, max_len=4)
    checked = 0
    for s in split_sentences(words, offsets):
        n = len(s)
        mm = {}
        root = max(_shape_maxmarginals(ref, list(s), shp, mm) for shp in _shapes(0, n))
        ins, _ = orc.inside(s)
        out = orc.outside(ins, n)
        masks = orc.masks(ins, out, n)
        cutoff = root - 9.0
        for (half, b, e), v in mm.items():
            cell = (tri(n) if half == "top" else 0) + cidx(b, e, n)
            proj = ref.proj_t if (half == "bot" and e - b == 1) else ref.proj_n
            for c in range(len(ref.coarse)):
                syms = np.nonzero(proj == c)[0]
                best = np.max(v[syms]) if len(syms) else -np.inf
                if abs(best - cutoff) < 1e-3:
                    continue  # fp32 chart vs FP64 brute force could fall either side
                bit = bool((masks[cell, c // 32] >> np.uint32(c % 32)) & np.uint32(1))
                assert bit == (best >= cutoff), (name, n, half, b, e, c, best, cutoff)
                checked += 1
        # cells on no bracketing of this sentence do not exist for n <= 4: every cell is covered
        assert len(mm) == n * (n + 1)
    assert checked > 1000
