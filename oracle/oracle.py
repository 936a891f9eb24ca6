"""Oracle (test infrastructure only): ctypes driver of the C restatement (cky_ref.c) plus a
brute-force enumerator over tree shapes for tiny inputs.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the
checker of the HIP path; the product never calls it. See cky_ref.c for the reference functions
each routine restates.
"""
import ctypes
import itertools
import os
import subprocess

import numpy as np

from .grammar_ref import RefGrammar

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libckyref.so")

_I32P = ctypes.POINTER(ctypes.c_int32)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_U32P = ctypes.POINTER(ctypes.c_uint32)


class _Bin(ctypes.Structure):
    _fields_ = [("n_b", ctypes.c_int32), ("p", _I32P), ("l", _I32P), ("r", _I32P), ("q", _F32P), ("w", _F32P)]


class _Un(ctypes.Structure):
    _fields_ = [("n_u", ctypes.c_int32), ("p", _I32P), ("c", _I32P), ("q", _F32P), ("w", _F32P)]


class _Gram(ctypes.Structure):
    _fields_ = [("nN", ctypes.c_int32), ("nT", ctypes.c_int32), ("cell_size", ctypes.c_int32),
                ("num_syms", ctypes.c_int32), ("root", ctypes.c_int32), ("num_coarse", ctypes.c_int32),
                ("mask_words", ctypes.c_int32), ("semiring", ctypes.c_int32), ("bin", _Bin * 4), ("un", _Un * 2),
                ("proj_n", _I32P), ("proj_t", _I32P), ("num_blocks", (ctypes.c_int32 * 4) * 3),
                ("block_of_coarse", (_I32P * 4) * 3), ("block_masks", (_U32P * 4) * 3),
                ("final_layer", ctypes.c_int32)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def tri(n):
    return n * (n + 1) // 2


def cidx(b, e, n):
    lo, s = min(b, e), abs(e - b) - 1
    return lo + s * n - s * (s - 1) // 2


class Oracle:
    """One grammar layer of the oracle."""

    def __init__(self, ref: RefGrammar, final_layer=True):
        self.ref = ref
        ref.load_lexicon()
        self.final_layer = final_layer
        self.cs = ref.cell_size
        self.zero = {"viterbi": -np.inf, "real": 0.0, "logsum": -1e6}[ref.semiring]
        self._keep = []
        g = _Gram()
        g.nN, g.nT, g.cell_size = len(ref.nonterms), len(ref.terms), ref.cell_size
        g.num_syms, g.root, g.num_coarse = ref.num_syms, ref.root, len(ref.coarse)
        g.mask_words, g.semiring = ref.mask_words, {"viterbi": 0, "real": 1, "logsum": 2}[ref.semiring]
        for f in range(4):
            b = ref.binaries[f]
            arrs = [np.ascontiguousarray(b[k]) for k in ("p", "l", "r", "q", "w")]
            self._keep += arrs
            g.bin[f] = _Bin(len(b), _p(arrs[0], _I32P), _p(arrs[1], _I32P), _p(arrs[2], _I32P),
                            _p(arrs[3], _F32P), _p(arrs[4], _F32P))
        for k, u in enumerate((ref.unary_n, ref.unary_t)):
            arrs = [np.ascontiguousarray(u[x]) for x in ("p", "c", "q", "w")]
            self._keep += arrs
            g.un[k] = _Un(len(u), _p(arrs[0], _I32P), _p(arrs[1], _I32P), _p(arrs[2], _F32P), _p(arrs[3], _F32P))
        pn = np.ascontiguousarray(ref.proj_n)
        pt = np.ascontiguousarray(ref.proj_t)
        self._keep += [pn, pt]
        g.proj_n, g.proj_t = _p(pn, _I32P), _p(pt, _I32P)
        for which in range(3):
            for f in range(4):
                nb, boc, m = ref.block_masks(which, f)
                boc = np.ascontiguousarray(boc, np.int32)
                m = np.ascontiguousarray(m, np.uint32)
                self._keep += [boc, m]
                g.num_blocks[which][f] = nb
                g.block_of_coarse[which][f] = _p(boc, _I32P)
                g.block_masks[which][f] = _p(m, _U32P)
        g.final_layer = int(final_layer)
        self.g = g
        self.dec = ref.decoder_tables()

    # -- per-sentence passes -------------------------------------------------------------------
    def new_chart(self, n, dtype=np.float32):
        return np.full((n * (n + 1), self.cs), self.zero, dtype)

    def inside(self, words, masks=None, in_scale=None, f64=False):
        """Returns (inside chart, partition scale)."""
        n = len(words)
        w = np.ascontiguousarray(words, np.int32)
        chart = self.new_chart(n, np.float64 if f64 else np.float32)
        ps = ctypes.c_double(0.0)
        r = self.ref
        isc = None if in_scale is None else np.ascontiguousarray(in_scale, np.float32)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        args_tag = (ctypes.byref(self.g), n, _p(w, _I32P), _p(r.lex_off, _I32P), _p(r.lex_tag, _I32P),
                    _p(r.lex_logp, _F64P), None if isc is None else _p(isc, _F32P))
        if f64:
            lib().ref_tags_f64(*args_tag, _p(chart, _F64P), ctypes.byref(ps))
            lib().ref_inside_f64(ctypes.byref(self.g), n, _p(chart, _F64P), None if m is None else _p(m, _U32P),
                                 None if isc is None else _p(isc, _F32P))
        else:
            lib().ref_tags(*args_tag, _p(chart, _F32P), ctypes.byref(ps))
            lib().ref_inside(ctypes.byref(self.g), n, _p(chart, _F32P), None if m is None else _p(m, _U32P),
                             None if isc is None else _p(isc, _F32P))
        return chart, ps.value

    def outside(self, inside, n, masks=None, in_scale=None, out_scale=None, f64=False):
        out = self.new_chart(n, np.float64 if f64 else np.float32)
        cells = n * (n + 1)
        isc = np.zeros(cells, np.float32) if in_scale is None else np.ascontiguousarray(in_scale, np.float32)
        osc = np.zeros(cells, np.float32) if out_scale is None else np.ascontiguousarray(out_scale, np.float32)
        m = None if masks is None else np.ascontiguousarray(masks, np.uint32)
        inside = np.ascontiguousarray(inside)
        if f64:
            lib().ref_outside_f64(ctypes.byref(self.g), n, _p(inside, _F64P), _p(out, _F64P),
                                  None if m is None else _p(m, _U32P), _p(isc, _F32P), _p(osc, _F32P))
        else:
            lib().ref_outside(ctypes.byref(self.g), n, _p(inside, _F32P), _p(out, _F32P),
                              None if m is None else _p(m, _U32P), _p(isc, _F32P), _p(osc, _F32P))
        return out

    def masks(self, inside, outside, n, thresh=-9.0):
        out = np.zeros((n * (n + 1), self.ref.mask_words), np.uint32)
        lib().ref_masks(ctypes.byref(self.g), n, _p(np.ascontiguousarray(inside, np.float32), _F32P),
                        _p(np.ascontiguousarray(outside, np.float32), _F32P), ctypes.c_float(thresh), _p(out, _U32P))
        return out

    def scaling(self, chart):
        chart = np.ascontiguousarray(chart, np.float32)
        out = np.zeros(chart.shape[0], np.float32)
        lib().ref_scaling(ctypes.byref(self.g), chart.shape[0], _p(chart, _F32P), _p(out, _F32P))
        return out

    def viterbi(self, inside, n):
        off, sc, le, ri = self.dec
        tree = np.zeros((2 * n - 1, 4), np.int32)
        lib().ref_viterbi(ctypes.byref(self.g), n, _p(np.ascontiguousarray(inside, np.float32), _F32P),
                          _p(off, _I32P), _p(sc, _F32P), _p(le, _I32P), _p(ri, _I32P), _p(tree, _I32P))
        return tree

    def mbr(self, inside, outside, n, in_scale, out_scale, has_scales):
        dec = np.zeros((n * (n + 1), 2), np.int32)
        lib().ref_mbr_cells(ctypes.byref(self.g), n, _p(np.ascontiguousarray(inside, np.float32), _F32P),
                            _p(np.ascontiguousarray(outside, np.float32), _F32P), _p(dec, _I32P))
        nodes = np.zeros((2 * n - 1, 5), np.int32)
        isc = np.ascontiguousarray(in_scale, np.float32)
        osc = np.ascontiguousarray(out_scale, np.float32)
        lib().ref_mbr_tree(n, _p(dec, _I32P), _p(isc, _F32P), _p(osc, _F32P), int(has_scales), _p(nodes, _I32P))
        return nodes, dec


def pipeline(layers, sentences, f64_final=False):
    """Coarse-to-fine fold of CLParser.parse (CLParser.scala:50-67) for each sentence.

    Returns a list of dicts with the final layer's charts, masks used, scales, partition and the
    decode (Viterbi tree or MBR nodes)."""
    out = []
    for words in sentences:
        n = len(words)
        masks, isc, osc, has_scales = None, None, None, False
        rec = {}
        for li, L in enumerate(layers):
            last = li == len(layers) - 1
            inside, ps = L.inside(words, masks, isc if has_scales else None, f64=last and f64_final)
            if not last:
                outside = L.outside(inside, n, masks)
                new_masks = L.masks(inside, outside, n)
                if layers[li + 1].ref.semiring == "real":
                    isc, osc, has_scales = L.scaling(inside), L.scaling(outside), True
                else:
                    isc = osc = None
                    has_scales = False
                rec.setdefault("coarse", []).append(dict(inside=inside, outside=outside, masks=new_masks))
                masks = new_masks
                continue
            rec["inside"] = inside
            rec["masks_in"] = masks
            rec["in_scale"] = isc if has_scales else np.zeros(n * (n + 1), np.float32)
            rec["out_scale"] = osc if has_scales else np.zeros(n * (n + 1), np.float32)
            rec["partition_scale"] = ps
            root = inside[tri(n) + cidx(0, n, n), L.ref.root]
            if L.ref.semiring == "viterbi":
                rec["partition"] = np.float32(root)
                rec["tree"] = L.viterbi(inside, n)
            elif L.ref.semiring == "logsum":
                # LogSumRuleSemiring.toLogSpace is the identity (RuleSemiring.scala:61); no scaling.
                rec["partition"] = np.float32(root)
                outside = L.outside(inside, n, masks)
                rec["outside"] = outside
                rec["mbr"], rec["mbr_cells"] = L.mbr(inside, outside, n, rec["in_scale"], rec["out_scale"], False)
            else:
                top = np.float32(rec["in_scale"][tri(n) + cidx(0, n, n)])
                rec["partition"] = np.float32(np.float32(np.log(np.float64(root))) + top) + np.float32(ps)
                outside = L.outside(inside, n, masks, rec["in_scale"], rec["out_scale"])
                rec["outside"] = outside
                rec["mbr"], rec["mbr_cells"] = L.mbr(inside, outside, n, rec["in_scale"], rec["out_scale"], has_scales)
        out.append(rec)
    return out


class _Layer(ctypes.Structure):
    _fields_ = [("g", ctypes.c_void_p), ("lex_off", _I32P), ("lex_tag", _I32P), ("lex_logp", _F64P)]


def batch(layers, words, offsets, with_outside=False, threads=0):
    """The coarse-to-fine fold of `pipeline` for a whole workload in C, parallel over sentences
    (ref_parse_batch in cky_batch.c). Returns a dict: 'partitions' float32[n]; 'trees' (Viterbi
    final layer) or 'mbr' (Real / LogSum) as per-sentence node arrays; 'masks' of the last coarse
    layer (cells x mask_words, sentences concatenated) when there is one; 'threads' used."""
    words = np.ascontiguousarray(words, np.int32)
    offsets = np.ascontiguousarray(offsets, np.int32)
    lens = np.diff(offsets).astype(np.int64)
    n = len(lens)
    tree_off = np.concatenate([[0], np.cumsum(2 * lens - 1)]).astype(np.int64)
    cell_off = np.concatenate([[0], np.cumsum(lens * (lens + 1))]).astype(np.int64)
    descs = (_Layer * len(layers))()
    for k, L in enumerate(layers):
        r = L.ref
        descs[k] = _Layer(ctypes.cast(ctypes.pointer(L.g), ctypes.c_void_p), _p(r.lex_off, _I32P),
                          _p(r.lex_tag, _I32P), _p(r.lex_logp, _F64P))
    final = layers[-1]
    viterbi = final.ref.semiring == "viterbi"
    off, sc, le, ri = final.dec
    trees = np.zeros((max(int(tree_off[-1]), 1), 4), np.int32)
    mbr = np.zeros((max(int(tree_off[-1]), 1), 5), np.int32)
    parts = np.zeros(max(n, 1), np.float32)
    mw = layers[0].ref.mask_words
    masks = np.zeros((max(int(cell_off[-1]), 1), mw), np.uint32) if len(layers) > 1 else None
    f = lib().ref_parse_batch
    f.restype = ctypes.c_int
    used = f(descs, len(layers), n, _p(words, _I32P), _p(offsets, _I32P), _p(off, _I32P), _p(sc, _F32P),
             _p(le, _I32P), _p(ri, _I32P), None if masks is None else _p(masks, _U32P), _p(cell_off, ctypes.POINTER(ctypes.c_int64)),
             _p(trees, _I32P) if viterbi else None, None if viterbi else _p(mbr, _I32P),
             _p(tree_off, ctypes.POINTER(ctypes.c_int64)), _p(parts, _F32P), int(with_outside), int(threads))
    nodes = trees if viterbi else mbr
    out = {"partitions": parts[:n], "threads": int(used),
           "trees" if viterbi else "mbr": [nodes[tree_off[i]:tree_off[i + 1]].copy() for i in range(n)]}
    if masks is not None:
        out["masks"] = masks[:int(cell_off[-1])]
    return out


# -- brute force over tree shapes (tiny inputs) ------------------------------------------------
def brute_force_root(ref: RefGrammar, words, semiring="viterbi"):
    """Best (or summed) score of ROOT over all binary bracketings of the sentence, computed per tree
    shape with the chart's unary layering: every node has a bottom symbol (tag at width 1, binary
    rule above) and a top symbol reached by exactly one unary rule (identity included). A child of
    width 1 enters a binary rule either as a tag (terminal families) or as its top symbol."""
    ref.load_lexicon()
    nN, nT = len(ref.nonterms), len(ref.terms)
    neg = -np.inf
    logspace = semiring in ("viterbi", "logsum")
    add = {"viterbi": np.maximum, "logsum": np.logaddexp, "real": lambda a, b: a + b}[semiring]
    zero = neg if logspace else 0.0
    q = lambda arr: arr["q"].astype(np.float64)

    def tags(pos):
        v = np.full(nT, zero)
        w = words[pos]
        for e in range(ref.lex_off[w], ref.lex_off[w + 1]):
            lp = float(np.float32(ref.lex_logp[e]))
            v[ref.lex_tag[e]] = lp if logspace else float(np.float32(np.exp(lp)))
        return v

    def unary(vec, rules):
        out = np.full(nN, zero)
        for p, c, qq in zip(rules["p"], rules["c"], q(rules)):
            val = vec[c] + qq if logspace else vec[c] * qq
            out[p] = add(out[p], val)
        return out

    def shapes(b, e):
        if e - b == 1:
            yield ("leaf", b)
            return
        for s in range(b + 1, e):
            for l in shapes(b, s):
                for r in shapes(s, e):
                    yield ("node", l, r)

    def evaluate(shape):
        """Returns (bot, top) vectors of a shape (bot over T for leaves, over N otherwise)."""
        if shape[0] == "leaf":
            bot = tags(shape[1])
            return bot, unary(bot, ref.unary_t), True
        lb, lt, lterm = evaluate(shape[1])
        rb, rt, rterm = evaluate(shape[2])
        bot = np.full(nN, zero)
        fams = [(0, lt, rt)]
        if lterm:
            fams.append((2, lb, rt))
        if rterm:
            fams.append((1, lt, rb))
        if lterm and rterm:
            fams.append((3, lb, rb))
        for f, L, R in fams:
            rules = ref.binaries[f]
            for p, l, r, qq in zip(rules["p"], rules["l"], rules["r"], q(rules)):
                val = (L[l] + R[r]) + qq if logspace else L[l] * R[r] * qq
                bot[p] = add(bot[p], val)
        return bot, unary(bot, ref.unary_n), False

    best = zero
    n = len(words)
    for shp in shapes(0, n):
        _, top, _ = evaluate(shp)
        best = add(best, top[ref.root])
    return best
