"""Autotuner of the batched engine's launch shape and chart packing.

The reference fixes its device work shapes when it generates kernels: rule segments of a coarse
parent block split into 8 sub-segments per kernel (CoarseParentSymbolGrammarClusterer.java:132-150),
work groups of 96 work-items (JavaFriendlyGenRuleMultiply.scala:62-63) and batches closed by
workspace size (WorkSpace.scala:20-41). The knobs of this engine that play those roles are:

  piece_target    rule-block partition: how finely each wave's rule stream is cut into work units
                  (1-16 pieces per stream while tiles x chunks x 2 x pieces stays within the
                  target, so the rules one workgroup runs per unit shrink as the target grows);
                  the role of the reference's 8 sub-segments per coarse parent block
  items_per_lane  items per lane of the rule kernels: 64 or 128 (cell, split) items per
                  1,024-thread workgroup
  reduce_targets  workgroup size of the reduce kernel: 1-16 chart targets (waves) per workgroup,
                  64-1,024 work-items; the role of the reference's fixed 96-item work groups
  persistent      one work-queue launch per span for all rule families, or one launch per family
  bucket_cells    chart cells per length bucket (chart packing)
  streams         buckets in flight on separate HIP streams

tune() measures device time of the batched engine on a sample of the caller's workload, one knob
at a time from the defaults (coordinate descent, two rounds), checks that every configuration
returns the same trees as the plain parser, and returns the best configuration with the table of
measurements. save()/load() keep the result as JSON.
"""
import json
import time

import numpy as np

DEFAULTS = dict(streams=3, bucket_cells=1 << 22, persistent=True, graphs=True, piece_target=512, items_per_lane=2,
                reduce_targets=4)
SPACE = dict(
    streams=[1, 2, 3, 4],
    bucket_cells=[1 << 18, 1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 23],
    piece_target=[64, 128, 256, 512, 1024, 2048],
    items_per_lane=[1, 2],
    reduce_targets=[1, 2, 4, 8, 16],
    persistent=[False, True],
)


def _measure(puck, layers, sents, cfg, mem_bytes, reps, fixed):
    bp = puck.BatchParser(layers, mem_bytes=mem_bytes, **cfg, **fixed)
    trees = None
    # First call runs eagerly, the second captures the bucket graphs, later calls replay them.
    for _ in range(2):
        trees = bp.parse(sents)
    best = float("inf")
    for _ in range(reps):
        trees = bp.parse(sents)
        best = min(best, bp.stats()["device_ms"])
    st = bp.stats()
    del bp
    return best, trees, st


def tune(puck, layers, sentences, mem_bytes=96 << 30, rounds=2, reps=2, log=None, fixed=None):
    """Returns {"best": cfg, "best_ms": t, "plain_ms": t0, "table": [...]} for the workload sample
    `sentences` ((words, offsets) or a list of word-id arrays). `fixed`: options of the workload
    itself, passed to every parser and not tuned (outside_final)."""
    fixed = dict(fixed or {})
    plain = puck.CLParser(layers, mem_bytes=mem_bytes, **fixed)
    ref = plain.parse(sentences)
    plain.parse(sentences)
    plain_ms = plain.stats()["device_ms"]
    del plain
    table = []
    seen = {}

    def run(cfg):
        key = tuple(sorted(cfg.items()))
        if key in seen:
            return seen[key]
        t0 = time.time()
        ms, trees, st = _measure(puck, layers, sentences, cfg, mem_bytes, reps, fixed)
        same = len(trees) == len(ref) and all(np.array_equal(a, b) for a, b in zip(trees, ref))
        row = dict(cfg, device_ms=round(ms, 3), buckets=int(st["buckets"]), same_trees=bool(same),
                   wall_s=round(time.time() - t0, 2))
        table.append(row)
        if log:
            log(row)
        if not same:
            raise RuntimeError(f"configuration {cfg} changed the parse results")
        seen[key] = ms
        return ms

    best = dict(DEFAULTS)
    best_ms = run(best)
    for _ in range(rounds):
        improved = False
        for knob, values in SPACE.items():
            for v in values:
                if best[knob] == v:
                    continue
                cand = dict(best, **{knob: v})
                ms = run(cand)
                if ms < best_ms * 0.99:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    return {"best": best, "best_ms": best_ms, "plain_ms": plain_ms, "table": table}


def save(path, result, workload):
    with open(path, "w") as f:
        json.dump(dict(result, workload=workload), f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)


def main(argv=None):
    """python -m puck_amd.autotune --out FILE [--workload mixed|c2f|inside_outside]: tunes the
    batched engine on a seeded sample of a workload and writes the result as JSON.

      mixed           the mixed-length coarse-to-fine workload (wsj_1 -> synthetic wsj_6 shape,
                      Viterbi, 70% log-normal 1-40 + 30% uniform 41-80; default 4,000 sentences)
      c2f             coarse-to-fine Viterbi (wsj_1 -> synthetic wsj_6 shape) on PCFG samples of
                      lengths 1-40 (default 5,000 sentences)
      inside_outside  exhaustive insideOutside + Viterbi decode on the synthetic wsj_6 shape,
                      log-normal lengths 1-40 (default 1,000 sentences)
    """
    import argparse
    import sys
    import puck_amd as puck
    ap = argparse.ArgumentParser(prog="puck_amd.autotune")
    ap.add_argument("--out", required=True)
    ap.add_argument("--workload", choices=("mixed", "c2f", "inside_outside"), default="mixed")
    ap.add_argument("--sentences", type=int, default=0,
                    help="sample size (default 4000 mixed, 5000 c2f, 1000 inside_outside)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--rounds", type=int, default=2)
    args = ap.parse_args(argv)
    g6 = puck.Grammar.load_text(puck.synth_wsj6_cached(seed=6), puck.VITERBI)
    log = lambda row: print(f"[autotune] {row}", file=sys.stderr, flush=True)  # noqa: E731
    if args.workload in ("mixed", "c2f"):
        mixed = args.workload == "mixed"
        n = args.sentences or (4000 if mixed else 5000)
        g1 = puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.VITERBI)
        sents = g6.sample(seed=args.seed, n=n, mode=2 if mixed else 0, min_len=1, max_len=80 if mixed else 40)
        res = tune(puck, [g1, g6], sents, rounds=args.rounds, log=log)
        workload = (f"wsj_1 -> synthetic wsj_6-shape c2f Viterbi, {n} sentences (seed {args.seed}), "
                    + ("70% log-normal 1-40 + 30% uniform 41-80" if mixed else "PCFG samples of lengths 1-40"))
    else:
        n = args.sentences or 1000
        sents = g6.sample(seed=args.seed, n=n, mode=1, min_len=1, max_len=40)
        res = tune(puck, [g6], sents, rounds=args.rounds, log=log, fixed=dict(outside_final=True))
        workload = (f"synthetic wsj_6-shape exhaustive insideOutside + Viterbi decode, {n} sentences "
                    f"(seed {args.seed}), log-normal lengths 1-40")
    save(args.out, res, workload)
    print(json.dumps({"best": res["best"], "best_ms": res["best_ms"], "plain_ms": res["plain_ms"]}))


if __name__ == "__main__":
    main()
