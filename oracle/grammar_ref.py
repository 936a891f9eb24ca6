"""Oracle (test infrastructure only): CPU restatement of the reference's grammar compilation.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, and only as
the checker. It restates, independently of the product's C++ grammar compiler:
  - RuleStructure symbol classification and rule families  (RuleStructure.scala:147-198)
  - rule scores, fromLogSpace                               (CLParser.scala:1266-1274, RuleSemiring.scala:31,87)
  - the "%f" score literals of the rule kernels             (SimpleGenRuleMultiply.java:151,304;
                                                             java.util.Formatter %f on a float argument)
  - CoarseParent rule blocks                                (CoarseParentSymbolGrammarClusterer.java:38-121,
                                                             JavaFriendlyGenRuleMultiply.scala:248-256,
                                                             CLParser.scala:783-785)
  - Viterbi decoder rule arrays                             (CLViterbi.scala:287-309,610-628)
  - hierarchical reprojection of a coarse/fine pair         (CLParser.scala:1153-1208)
Symbol ids follow (coarse id in .numstates order, substate); rule order is file order; rules with
ln p < -12 are dropped (parseBerkeleyText(prefix, -12, ...), CLParser.scala:1026). These are the
documented choices for behaviour that lives in epic 0.2, which is not in the reference checkout.
"""
import math
from decimal import Decimal, ROUND_HALF_UP

import numpy as np

THRESHOLD = -12.0


def java_f6(x):
    """String.format("%f", (float) x) -> fp32 literal value, restated with Python's shortest-repr
    digits of the widened double and decimal HALF_UP rounding (Formatter / FormattedFloatingDecimal)."""
    f = float(np.float32(x))
    s = java_f6_string(f)
    return float(np.float32(float(s)))


def java_f6_string(f):
    if math.isnan(f):
        return "NaN"
    if math.isinf(f):
        return "-Infinity" if f < 0 else "Infinity"
    neg = math.copysign(1.0, f) < 0
    digits = Decimal(repr(abs(f)))  # shortest round-trip digits of the double
    q = digits.quantize(Decimal("0.000001"), rounding=ROUND_HALF_UP)
    body = format(q, "f")
    if "." not in body:
        body += ".000000"
    return ("-" if neg else "") + body


def _split(sym):
    base, sub = sym.rsplit("_", 1)
    return base, int(sub)


class RefGrammar:
    """Rule tables of one text grammar under one semiring ('viterbi', 'real' or 'logsum'; LogSum
    keeps rule scores in log space like Viterbi, RuleSemiring.scala:45-63)."""

    def __init__(self, prefix, semiring="viterbi", threshold=THRESHOLD):
        self.prefix = prefix
        self.semiring = semiring
        self.coarse = []
        self.coarse_id = {}
        for line in open(prefix + ".numstates"):
            parts = line.split()
            if not parts:
                continue
            self.coarse_id[parts[0]] = len(self.coarse)
            self.coarse.append(parts[0])
        self.mask_words = (len(self.coarse) + 31) // 32
        binary, unary = [], []
        for line in open(prefix + ".binary"):
            p = line.split()
            if p:
                binary.append((p[0], p[1], p[2], float(p[3])))
        for line in open(prefix + ".unary"):
            p = line.split()
            if p:
                unary.append((p[0], p[1], float(p[2])))
        keep = lambda pr: pr > 0 and math.log(pr) >= threshold
        binary = [r for r in binary if keep(r[3])]
        unary = [r for r in unary if r[0] == r[1] or keep(r[2])]
        self.dropped = None
        nonterms, rhs = set(), set()
        for p, l, r, _ in binary:
            nonterms.add(p)
            rhs.update((l, r))
        for p, c, _ in unary:
            if p != c:
                nonterms.add(p)
                rhs.add(c)
        key = lambda s: (self.coarse_id[_split(s)[0]], _split(s)[1])
        self.nonterms = sorted(nonterms, key=key)
        self.terms = sorted(rhs - nonterms, key=key)
        self.nid = {s: i for i, s in enumerate(self.nonterms)}
        self.tid = {s: i for i, s in enumerate(self.terms)}
        self.proj_n = np.array([self.coarse_id[_split(s)[0]] for s in self.nonterms], np.int32)
        self.proj_t = np.array([self.coarse_id[_split(s)[0]] for s in self.terms], np.int32)
        self.root = self.nid["ROOT_0"]
        self.num_syms = max(len(self.nonterms), len(self.terms))
        self.cell_size = (self.num_syms + 31) // 32 * 32

        def score(prob):
            s = np.float32(math.log(prob))
            if semiring == "real":
                return np.float32(math.exp(float(s)))
            return s

        fam = {0: [], 1: [], 2: [], 3: []}
        for p, l, r, pr in binary:
            lt, rt = l in self.tid, r in self.tid
            f = (3 if rt else 2) if lt else (1 if rt else 0)
            w = score(pr)
            fam[f].append((self.nid[p], self.tid[l] if lt else self.nid[l], self.tid[r] if rt else self.nid[r],
                           float(w), java_f6(w)))
        self.binaries = [np.array(fam[f], dtype=[("p", "i4"), ("l", "i4"), ("r", "i4"), ("w", "f4"), ("q", "f4")])
                         for f in range(4)]
        un, ut = [], []
        for p, c, pr in unary:
            if p == c and p not in self.nid:
                continue
            w = score(pr)
            if c in self.nid:
                un.append((self.nid[p], self.nid[c], float(w), java_f6(w)))
            else:
                ut.append((self.nid[p], self.tid[c], float(w), java_f6(w)))
        dt = [("p", "i4"), ("c", "i4"), ("w", "f4"), ("q", "f4")]
        self.unary_n = np.array(un, dtype=dt)
        self.unary_t = np.array(ut, dtype=dt)
        self._lexicon_loaded = False

    # -- CoarseParent blocks -------------------------------------------------------------------
    def blocks(self, which, family):
        """(num_blocks, block_of_coarse) for which=0 inside (parent), 1 outside-left (left child),
        2 outside-right (right child)."""
        rules = self.binaries[family]
        lt, rt = family in (2, 3), family in (1, 3)
        if which == 0:
            keys = self.proj_n[rules["p"]]
        elif which == 1:
            keys = (self.proj_t if lt else self.proj_n)[rules["l"]]
        else:
            keys = (self.proj_t if rt else self.proj_n)[rules["r"]]
        counts = np.bincount(keys, minlength=len(self.coarse)) if len(keys) else np.zeros(len(self.coarse), int)
        present = [c for c in range(len(self.coarse)) if counts[c] > 0]
        present.sort(key=lambda c: -counts[c])  # stable: ties keep ascending coarse id
        boc = np.full(len(self.coarse), -1, np.int32)
        nb = 0
        i = 0
        while i < len(present) and counts[present[i]] >= 300:
            boc[present[i]] = nb
            nb += 1
            i += 1
        if i < len(present):
            for c in present[i:]:
                boc[c] = nb
            nb += 1
        return nb, boc

    def block_masks(self, which, family):
        nb, boc = self.blocks(which, family)
        m = np.zeros((max(nb, 1), self.mask_words), np.uint32)
        for c, b in enumerate(boc):
            if b >= 0:
                m[b, c // 32] |= np.uint32(1 << (c % 32))
        return nb, boc, m

    # -- Viterbi decoder arrays ----------------------------------------------------------------
    def decoder_tables(self):
        nN = len(self.nonterms)
        offsets = [0]
        scores, lefts, rights = [], [], []
        groups = [(b["p"], b["w"], b["l"], b["r"]) for b in self.binaries]
        groups += [(self.unary_n["p"], self.unary_n["w"], self.unary_n["c"], None),
                   (self.unary_t["p"], self.unary_t["w"], self.unary_t["c"], None)]
        for parents, w, l, r in groups:
            order = np.argsort(parents, kind="stable")
            start = len(scores)
            sp = parents[order]
            scores.extend(w[order].tolist())
            lefts.extend(l[order].tolist())
            if r is not None:
                rights.extend(r[order].tolist())
            for sym in range(nN):
                offsets.append(start + int(np.searchsorted(sp, sym, side="right")))
        return (np.array(offsets, np.int32), np.array(scores, np.float32), np.array(lefts, np.int32),
                np.array(rights, np.int32))

    # -- lexicon -------------------------------------------------------------------------------
    def load_lexicon(self):
        if self._lexicon_loaded:
            return
        self.word_ids = {}
        entries = []
        for line in open(self.prefix + ".lexicon"):
            if not line.strip():
                continue
            tag, rest = line.split(None, 1)
            lb = rest.rfind("[")
            word = rest[:lb].strip()
            probs = [float(x) for x in rest[lb + 1:rest.rfind("]")].split(",") if x.strip()]
            if word not in self.word_ids:
                self.word_ids[word] = len(entries)
                entries.append({})
            e = entries[self.word_ids[word]]
            for i, pr in enumerate(probs):
                if pr > 0:
                    t = self.tid.get(f"{tag}_{i}")
                    if t is not None:
                        e[t] = math.log(pr)
        off = [0]
        tags, logp = [], []
        for e in entries:
            for t in sorted(e):
                tags.append(t)
                logp.append(e[t])
            off.append(len(tags))
        self.lex_off = np.array(off, np.int32)
        self.lex_tag = np.array(tags, np.int32)
        self.lex_logp = np.array(logp, np.float64)
        self._lexicon_loaded = True


# -- hierarchical reprojection -------------------------------------------------------------------
def _hier_parse(text):
    """One .hierarchy tree as nested [label, children, words] lists (PennTreeReader.parseHard as
    readHierarchy uses it, CLParser.scala:1187-1188): brackets holding only words are a leaf."""
    toks = text.replace("(", " ( ").replace(")", " ) ").split()
    pos = 0

    def node():
        nonlocal pos
        assert toks[pos] == "(", text
        label = toks[pos + 1]
        pos += 2
        kids, words = [], []
        while toks[pos] != ")":
            if toks[pos] == "(":
                kids.append(node())
            else:
                words.append(toks[pos])
                pos += 1
        pos += 1
        if kids and words:
            raise ValueError("hierarchy node mixes words and subtrees: " + text)
        return [label, kids, words]

    root = node()
    if pos != len(toks):
        raise ValueError("trailing text after hierarchy tree: " + text)
    return root


def _left_height(t):
    return 0 if not t[1] else 1 + _left_height(t[1][0])


def _preorder(t):
    yield t
    for k in t[1]:
        yield from _preorder(k)


def _leaves(t):
    """Words under t, left to right (the node's span of the line's words)."""
    if not t[1]:
        return list(t[2])
    out = []
    for k in t[1]:
        out.extend(_leaves(k))
    return out


def read_hierarchy(path, coarse_level, fine_level):
    """readHierarchy (CLParser.scala:1182-1208): {coarse-level symbol: [fine symbols]}."""
    sym_map = {}
    for line in open(path):
        if not line.strip():
            continue
        sym, rest = line.split(None, 1)
        tree = _hier_parse(rest)
        total = _left_height(tree)
        for sub in _preorder(tree):
            if _left_height(sub) != total - coarse_level:
                continue
            if fine_level - 1 == total:
                fines = [f"{sym}_{q}" for q in _leaves(sub)]
            else:
                fines = [f"{sym}_{d[0]}" for d in _preorder(sub) if _left_height(d) == total - fine_level]
            sym_map[f"{sym}_{sub[0]}"] = fines
    return sym_map


def level_of(name):
    """The last number in a grammar name (readHierarchy's regex on the grammar names)."""
    import re
    m = re.match(r"^.*(?:^|[^0-9])([0-9]+)([^0-9]*)$", name)
    if not m:
        raise ValueError("no split level in " + name)
    return int(m.group(1))


def reproject(coarse, fine, hierarchy=None, coarse_level=None, fine_level=None):
    """reprojectGrammar (CLParser.scala:1153-1180) on two RefGrammars, in place: the coarse
    grammar's refined symbols become the coarse symbols of both; the coarse layer projects each
    symbol to itself, the fine layer to its ancestor at the coarse level; fine rules whose
    projection is not a coarse rule are scored fromLogSpace(-12) (CLParser.scala:1266-1274)."""
    hierarchy = hierarchy or fine.prefix + ".hierarchy"
    cl = level_of(coarse.prefix) if coarse_level is None else coarse_level
    fl = level_of(fine.prefix) if fine_level is None else fine_level
    up = {}
    for c, fines in read_hierarchy(hierarchy, cl, fl).items():
        for f in fines:
            if up.get(f, c) != c:
                raise ValueError(f"{f} is under two coarse symbols")
            up[f] = c
    up["ROOT_0"] = "ROOT_0"
    key = lambda s: (coarse.coarse_id[_split(s)[0]], _split(s)[1])
    names = sorted(coarse.nonterms + coarse.terms, key=key)
    cid = {s: i for i, s in enumerate(names)}
    for g in (coarse, fine):
        g.coarse = list(names)
        g.coarse_id = None  # base-symbol ids no longer index the coarse symbols
        g.mask_words = (len(names) + 31) // 32
    coarse.proj_n = np.array([cid[s] for s in coarse.nonterms], np.int32)
    coarse.proj_t = np.array([cid[s] for s in coarse.terms], np.int32)
    fine.proj_n = np.array([cid[up[s]] for s in fine.nonterms], np.int32)
    fine.proj_t = np.array([cid[up[s]] for s in fine.terms], np.int32)

    def proj(g, fam, which, idx):
        term = (fam in (2, 3)) if which == "l" else (fam in (1, 3))
        return (g.proj_t if term else g.proj_n)[idx]

    have_b = set()
    for fam in range(4):
        b = coarse.binaries[fam]
        for p, l, r in zip(coarse.proj_n[b["p"]], proj(coarse, fam, "l", b["l"]), proj(coarse, fam, "r", b["r"])):
            have_b.add((int(p), int(l), int(r)))
    have_u = {(int(p), int(c)) for p, c in zip(coarse.proj_n[coarse.unary_n["p"]], coarse.proj_n[coarse.unary_n["c"]])}
    have_u |= {(int(p), int(c)) for p, c in zip(coarse.proj_n[coarse.unary_t["p"]], coarse.proj_t[coarse.unary_t["c"]])}
    w12 = np.float32(math.exp(-12.0)) if fine.semiring == "real" else np.float32(-12.0)
    q12 = java_f6(w12)
    for fam in range(4):
        b = fine.binaries[fam]
        keys = zip(fine.proj_n[b["p"]], proj(fine, fam, "l", b["l"]), proj(fine, fam, "r", b["r"]))
        miss = np.array([(int(p), int(l), int(r)) not in have_b for p, l, r in keys], bool)
        b["w"][miss] = w12
        b["q"][miss] = q12
    for arr, cp in ((fine.unary_n, fine.proj_n), (fine.unary_t, fine.proj_t)):
        miss = np.array([(int(p), int(c)) not in have_u for p, c in zip(fine.proj_n[arr["p"]], cp[arr["c"]])], bool)
        arr["w"][miss] = w12
        arr["q"][miss] = q12
    return coarse, fine
