"""Writes the seeded golden fixtures under tests/golden/ (run from the repository root:
`python tests/golden/make_golden.py`).

Fixtures (npz, small):
  brute_force.npz     wsj_0 / wsj_1 sentences of length 1-4 and their root scores from the
                      brute-force enumeration over tree shapes (Viterbi max and Real sum, FP64)
  wsj1_viterbi.npz    12 wsj_1 sentences (length 3-9): oracle root inside scores and Viterbi trees
  wsj1_real.npz       the same sentences: oracle log-likelihoods under the Real semiring
  wsj0_wsj1_c2f.npz   6 sentences: coarse wsj_0 masks (threshold -9) and the masked fine wsj_1
                      Viterbi trees
The reference pins no parse results (its tests only check partitions for NaN), so these vectors
come from the oracle and the brute-force enumerator; they pin both against regressions.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import puck_amd as puck  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oracle.grammar_ref import RefGrammar  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sample(name, seed, n, lo, hi):
    g = puck.Grammar.load_text(puck.grammar_path(name))
    return g.sample(seed=seed, n=n, mode=0, min_len=lo, max_len=hi)


def main():
    rec = {}
    for name in ("wsj_0", "wsj_1"):
        w, o = sample(name, 21, 6, 1, 4)
        rv = RefGrammar(puck.grammar_path(name), "viterbi")
        rr = RefGrammar(puck.grammar_path(name), "real")
        vit, real = [], []
        for i in range(len(o) - 1):
            s = list(w[o[i]:o[i + 1]])
            vit.append(O.brute_force_root(rv, s, "viterbi"))
            real.append(np.log(O.brute_force_root(rr, s, "real")))
        rec[f"{name}_words"], rec[f"{name}_offsets"] = w, o
        rec[f"{name}_viterbi"], rec[f"{name}_loglik"] = np.array(vit), np.array(real)
    np.savez(os.path.join(OUT, "brute_force.npz"), **rec)

    w, o = sample("wsj_1", 31, 12, 3, 9)
    orc = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"))
    res = O.pipeline([orc], [w[o[i]:o[i + 1]] for i in range(len(o) - 1)])
    trees = np.concatenate([r["tree"] for r in res])
    np.savez(os.path.join(OUT, "wsj1_viterbi.npz"), words=w, offsets=o,
             roots=np.array([r["partition"] for r in res], np.float32), trees=trees)
    orr = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "real"))
    res = O.pipeline([orr], [w[o[i]:o[i + 1]] for i in range(len(o) - 1)])
    np.savez(os.path.join(OUT, "wsj1_real.npz"), words=w, offsets=o,
             loglik=np.array([r["partition"] for r in res], np.float32))

    # Coarse-to-fine needs one vocabulary: sample from wsj_1 (wsj_0 shares the lexicon words).
    w, o = sample("wsj_1", 41, 6, 4, 10)
    coarse = O.Oracle(RefGrammar(puck.grammar_path("wsj_0"), "viterbi"), final_layer=False)
    fine = O.Oracle(RefGrammar(puck.grammar_path("wsj_1"), "viterbi"))
    res = O.pipeline([coarse, fine], [w[o[i]:o[i + 1]] for i in range(len(o) - 1)])
    np.savez(os.path.join(OUT, "wsj0_wsj1_c2f.npz"), words=w, offsets=o,
             masks=np.concatenate([r["coarse"][0]["masks"] for r in res]),
             trees=np.concatenate([r["tree"] for r in res]))
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
