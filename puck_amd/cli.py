"""Command-line front ends of the parse path: `compile` (CompileGrammar) and `parse` (RunParser).

  python -m puck_amd.cli compile --textGrammarPrefix a.gr:b.gr --grammar grammar.grz [--viterbi true]
      CompileGrammar.main (CompileGrammar.scala:19-74): every prefix is a layer, coarse first;
      coarse layers use the Viterbi semiring, the last one Viterbi (--viterbi true), LogSum
      (--viterbi false --logsum true) or Real (--viterbi false), the last two with max-rule
      decoding.
  python -m puck_amd.cli parse --grammar grammar.grz [--maxLength 50] [--mem 16g] [files...]
      RunParser.main (RunParser.scala:31-138) with its "newline" sentence segmenter
      (NewLineSentenceSegmenter.scala:18-39) and "whitespace" tokenizer: one sentence per input
      line, bracketed trees out, "(())" for sentences of maxLength words or more and for failures
      (RunParser.scala:86-96). Reads stdin and writes stdout without files, else writes
      <file>.parsed (--outputSuffix).

Unknown words are mapped to the lexicon's UNK-* signatures by word shape (capitalisation, digits,
dashes, suffix classes), the scheme the Berkeley lexicons' signature entries are named by; the
reference delegates this to its lexicon library, which is not in the checkout, so the mapping is
this module's own (signatures absent from the lexicon back off to shorter ones, then to "UNK").
Trees are rendered with coarse labels; nodes whose label starts with '@' (binarization
intermediates) are spliced into their parent and identity unaries are printed once.
"""
import argparse
import os
import sys

import numpy as np

SUFFIXES = ("ed", "ing", "ion", "er", "est", "ly", "ity", "y", "al")


def signature(word, pos, known_lower):
    """UNK-* signature of an unknown word by its shape."""
    sig = "UNK"
    caps = sum(c.isupper() for c in word)
    digit = any(c.isdigit() for c in word)
    dash = "-" in word
    lower = any(c.islower() for c in word)
    if word[:1].isupper():
        if pos == 0 and caps == 1:
            sig += "-INITC"
            if known_lower:
                sig += "-KNOWNLC"
        else:
            sig += "-CAPS"
    elif not word[:1].isalpha() and caps > 0:
        sig += "-CAPS"
    elif lower:
        sig += "-LC"
    if digit:
        sig += "-NUM"
    if dash:
        sig += "-DASH"
    lw = word.lower()
    if lw.endswith("s") and len(lw) >= 3 and lw[-2] not in "siu":
        sig += "-s"
    elif len(lw) >= 5 and not dash and not (digit and caps > 0):
        for suf in SUFFIXES:
            if lw.endswith(suf):
                sig += "-" + suf
                break
    return sig


def word_ids(grammar, tokens):
    """Lexicon ids of a tokenized sentence; None if a word has no usable signature."""
    out = []
    for i, w in enumerate(tokens):
        wid = grammar.word_id(w)
        if wid < 0:
            sig = signature(w, i, grammar.word_id(w.lower()) >= 0)
            parts = sig.split("-")
            while parts:
                wid = grammar.word_id("-".join(parts))
                if wid >= 0:
                    break
                parts.pop()
        if wid < 0:
            return None
        out.append(wid)
    return out


def render(nodes, tokens, nt_names, t_names):
    """Bracketed tree of pre-order Viterbi nodes (top, bot, width, score) over `tokens`."""
    pos = [0]

    def node(i):
        top, bot, width = int(nodes[i][0]), int(nodes[i][1]), int(nodes[i][2])
        if width == 1:
            label = t_names[bot]
            inner = [f"({label} {tokens[pos[0]]})"]
            pos[0] += 1
            size = 1
        else:
            left, lsize = node(i + 1)
            right, rsize = node(i + 1 + lsize)
            label = nt_names[bot]
            kids = left + right
            inner = kids if label.startswith("@") else [f"({label} {' '.join(kids)})"]
            size = 1 + lsize + rsize
        if top >= 0:
            tl = nt_names[top]
            if not (tl == label and not label.startswith("@")):
                inner = inner if tl.startswith("@") else [f"({tl} {' '.join(inner)})"]
        return inner, size

    out, _ = node(0)
    return " ".join(out)


def parse_mem(s):
    s = s.strip().lower()
    mult = {"k": 1 << 10, "m": 1 << 20, "g": 1 << 30}.get(s[-1:], 1)
    return int(float(s.rstrip("kmg")) * mult)


def cmd_compile(args):
    import puck_amd as puck
    prefixes = args.textGrammarPrefix.split(":")
    # CompileGrammar.scala:46-52: Viterbi if --viterbi, else LogSum if --logsum, else Real.
    if args.viterbi.lower() == "true":
        fine = puck.VITERBI
    elif args.logsum.lower() == "true":
        fine = puck.LOGSUM
    else:
        fine = puck.REAL
    layers = [puck.Grammar.load_text(p, puck.VITERBI) for p in prefixes[:-1]]
    layers.append(puck.Grammar.load_text(prefixes[-1], fine))
    if args.reproject.lower() == "true" and len(layers) > 1:
        # CompileGrammar.scala:57-59: the first and last grammars, reprojected, replace the list.
        layers = list(puck.Grammar.reproject(layers[0], layers[-1]))
    puck.save_grz(args.grammar, layers)
    print(f"wrote {args.grammar}: {len(layers)} layer(s)", file=sys.stderr)


def cmd_parse(args):
    import puck_amd as puck
    layers = puck.Grammar.load_grz(args.grammar)
    if layers[-1].semiring != puck.VITERBI:
        raise SystemExit("parse renders Viterbi trees: compile the grammar with --viterbi true")
    fine = layers[-1]
    coarse = fine.coarse_names()
    nt_c, _ = fine.symbols(False)
    t_c, _ = fine.symbols(True)
    nt_names = [coarse[c] for c in nt_c]
    t_names = [coarse[c] for c in t_c]
    parser = puck.BatchParser(layers, device=args.device, mem_bytes=parse_mem(args.mem), graphs=False)
    streams = [(sys.stdin, sys.stdout)] if not args.files else [
        (open(f), open(f + args.outputSuffix, "w")) for f in args.files]
    for src, sink in streams:
        sents = [line.split() for line in src if line.strip()]
        ids, keep = [], []
        for k, toks in enumerate(sents):
            w = word_ids(fine, toks) if len(toks) < args.maxLength else None
            # Every layer reads the same word ids: they must be in each layer's lexicon.
            if w is not None and all(max(w) < layer.info.num_words for layer in layers):
                ids.append(np.asarray(w, np.int32))
                keep.append(k)
        out = ["(())"] * len(sents)
        if ids:
            trees = parser.parse(ids)
            for k, t in zip(keep, trees):
                out[k] = render(t, sents[k], nt_names, t_names)
        for line in out:
            print(line, file=sink)
        if sink is not sys.stdout:
            sink.close()


def main(argv=None):
    ap = argparse.ArgumentParser(prog="puck_amd.cli")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("compile", help="text grammars -> grammar.grz (CompileGrammar)")
    c.add_argument("--textGrammarPrefix", required=True)
    c.add_argument("--grammar", default="grammar.grz")
    c.add_argument("--viterbi", default="true")
    c.add_argument("--logsum", default="false")
    c.add_argument("--reproject", default="false")
    c.add_argument("--device", default="0")
    p = sub.add_parser("parse", help="parse text with a grammar.grz (RunParser)")
    p.add_argument("--grammar", default="grammar.grz")
    p.add_argument("--maxLength", type=int, default=50)
    p.add_argument("--mem", default="16g")
    p.add_argument("--device", type=int, default=0)
    p.add_argument("--outputSuffix", default=".parsed")
    p.add_argument("files", nargs="*")
    args = ap.parse_args(argv)
    if args.cmd == "compile":
        cmd_compile(args)
    else:
        cmd_parse(args)


if __name__ == "__main__":
    main()
