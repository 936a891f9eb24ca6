"""Benchmark of the CKY parse path on one MI355X.

Workload (the headline configuration): exhaustive Viterbi parsing with the synthetic wsj_6-shape
grammar (1,084 refined states = 502 nonterminals + 582 tags, 1,725,570 binary and 125,619 unary
rules; the real wsj_6 rule files are not in the reference checkout) of 2,000 seeded PCFG-sampled
sentences, lengths 1-40 from a discretized log-normal (mean ~23). One step = one parse of the
2,000 sentences: tag scores, inside pass (one launch per span and rule family), Viterbi decode.

Prints one JSON line. `value` is sentences/s from device time (HIP events from the first kernel
after the word upload to the tree download); the PCIe-inclusive host-API rate is reported beside
it as `e2e_value`.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sentences/sec (wsj_6, len≤40) batched+served vs plain GPU; p50/p99 latency"
VALU_PEAK = 256 * 64 * 2.4e9  # fp32 lane-ops/s without packed math


def hip_sync():
    try:
        ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()
    except OSError:
        pass


def cpu_baseline(puck, prefix, words, offsets, sample, threads):
    """The oracle (CPU restatement, OpenMP over sentences) on the first `sample` sentences."""
    from oracle.grammar_ref import RefGrammar
    from oracle import oracle as O
    t0 = time.time()
    ref = RefGrammar(prefix, "viterbi")
    orc = O.Oracle(ref)
    load_s = time.time() - t0
    lens = np.diff(offsets)[:sample]
    sub_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    sub_words = np.ascontiguousarray(words[:sub_off[-1]], np.int32)
    tree_off = np.concatenate([[0], np.cumsum(2 * lens - 1)]).astype(np.int64)
    trees = np.zeros((int(tree_off[-1]), 4), np.int32)
    off, sc, le, ri = orc.dec
    lib = O.lib()
    lib.ref_parse_viterbi_batch.restype = ctypes.c_int
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    t0 = time.time()
    used = lib.ref_parse_viterbi_batch(
        ctypes.byref(orc.g), ref.cell_size, len(lens), P(sub_words, ctypes.c_int32), P(sub_off, ctypes.c_int32),
        P(ref.lex_off, ctypes.c_int32), P(ref.lex_tag, ctypes.c_int32), P(ref.lex_logp, ctypes.c_double),
        P(off, ctypes.c_int32), P(sc, ctypes.c_float), P(le, ctypes.c_int32), P(ri, ctypes.c_int32),
        P(trees, ctypes.c_int32), P(tree_off, ctypes.c_int64), threads)
    dt = time.time() - t0
    return dict(value=len(lens) / dt, unit="sentences/s", cores=int(used), kind="port",
                sample=f"first {len(lens)} of the same 2000 sentences (oracle C restatement, "
                       f"OpenMP over sentences, {dt:.1f} s; grammar load {load_s:.1f} s excluded)"), trees, tree_off


PROFILE = os.path.join(ROOT, "profiles", "k_rules_wsj6_exhaustive_viterbi.md")


def profiled_traffic():
    """HBM bytes per k_rules launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of
    this workload (copy-calibrated), or None when the summary is absent."""
    try:
        with open(PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def timed_parses(parser, words, offsets, steps):
    """Warm-up parse, then `steps` parses; returns (device ms per parse, [(ms, apps, family)] of the
    last parse)."""
    parser.parse((words, offsets))
    dev = 0.0
    for _ in range(steps):
        parser.parse((words, offsets))
        dev += parser.stats()["device_ms"]
    return dev / steps, parser.launches()


def fine_roofline(launches):
    """fp32 VALU roofline of the fine layer's inside binary rule launches (pruned rule-apps)."""
    ms, apps, fam = launches
    sel = (fam >= 16) & ((fam % 16) < 4)
    t = float(ms[sel].sum()) / 1e3
    ach = 3.0 * float(apps[sel].sum()) / t if t > 0 else 0.0
    return {"bound": "valu", "achieved": round(ach / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 2),
            "unit": "TFLOP/s", "frac": round(ach / VALU_PEAK, 4), "kernel": "k_rules (fine inside binary, masked)",
            "launches": int(sel.sum()), "rule_apps": float(apps[sel].sum())}


def serve(puck, layers, pool, rate, seconds, seed):
    """Open-loop Poisson arrivals at `rate` requests/s for `seconds`; returns throughput and latency."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * seconds * 1.2) + 16))
    arrivals = arrivals[arrivals < seconds]
    srv = puck.Server(layers, mem_bytes=96 << 30, max_inflight=3, max_batch=20000)
    srv.stats(reset=True)
    t0 = time.perf_counter()
    for k, at in enumerate(arrivals):
        while True:
            dt = at - (time.perf_counter() - t0)
            if dt <= 0:
                break
            if dt > 0.002:
                time.sleep(dt - 0.001)
        srv.submit(pool[k % len(pool)])
    srv.flush()
    wall = time.perf_counter() - t0
    st = srv.stats()
    del srv
    return {"rate": rate, "seconds": seconds, "requests": int(len(arrivals)),
            "throughput": round(st["completed"] / wall, 1), "p50_ms": round(st["p50_ms"], 2),
            "p99_ms": round(st["p99_ms"], 2), "max_ms": round(st["max_ms"], 2), "batches": int(st["batches"]),
            "mean_batch": round(st["mean_batch"], 1)}


def side_workloads(puck, g6v, prefix, args):
    """Coarse-to-fine, mixed-length and serving workloads, measured beside the headline."""
    out = {}
    g1 = puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.VITERBI)
    words, offsets = g6v.sample(seed=3, n=args.c2f_sentences, mode=0, min_len=1, max_len=40)
    lens = np.diff(offsets)
    p = puck.CLParser([g1, g6v], device=0, mem_bytes=64 << 30)
    ms, launches = timed_parses(p, words, offsets, 2)
    out["c2f_viterbi"] = {
        "workload": f"wsj_1 -> synthetic wsj_6-shape Viterbi, masks at -9, {len(lens)} PCFG sentences 1-40",
        "value": round(len(lens) / (ms / 1e3), 1), "unit": "sentences/s", "ms_per_parse": round(ms, 2),
        "mean_length": round(float(lens.mean()), 2), "roofline": fine_roofline(launches)}
    print(f"[bench] c2f viterbi {out['c2f_viterbi']['value']} sentences/s", file=sys.stderr, flush=True)
    g6r = puck.Grammar.load_text(prefix, puck.REAL)
    pr = puck.CLParser([g1, g6r], device=0, mem_bytes=64 << 30)
    ms, launches = timed_parses(pr, words, offsets, 2)
    out["c2f_mbr"] = {
        "workload": f"wsj_1 -> synthetic wsj_6-shape Real inside/outside + MBR, {len(lens)} PCFG sentences 1-40",
        "value": round(len(lens) / (ms / 1e3), 1), "unit": "sentences/s", "ms_per_parse": round(ms, 2),
        "roofline": fine_roofline(launches)}
    print(f"[bench] c2f mbr {out['c2f_mbr']['value']} sentences/s", file=sys.stderr, flush=True)
    del pr, g6r
    mw, mo = g6v.sample(seed=4, n=args.mixed_sentences, mode=2, min_len=1, max_len=80)
    ml = np.diff(mo)
    ms, launches = timed_parses(p, mw, mo, 1)
    out["mixed_lengths"] = {
        "workload": f"c2f Viterbi, {len(ml)} sentences, 70% log-normal 1-40 + 30% uniform 41-80",
        "value": round(len(ml) / (ms / 1e3), 1), "unit": "sentences/s", "ms_per_parse": round(ms, 2),
        "mean_length": round(float(ml.mean()), 2), "roofline": fine_roofline(launches)}
    print(f"[bench] mixed lengths {out['mixed_lengths']['value']} sentences/s", file=sys.stderr, flush=True)
    del p
    pool_w, pool_o = g6v.sample(seed=5, n=20000, mode=0, min_len=1, max_len=40)
    pool = [np.ascontiguousarray(pool_w[pool_o[i]:pool_o[i + 1]]) for i in range(len(pool_o) - 1)]
    out["serving"] = []
    for rate in (200, 1000, 5000):
        r = serve(puck, [g1, g6v], pool, rate, args.serve_seconds, seed=rate)
        out["serving"].append(r)
        print(f"[bench] serving {r}", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sentences", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--ipl", type=int, default=2, help="items per lane of the rule kernels (1 or 2)")
    ap.add_argument("--cpu-sample", type=int, default=64, help="sentences timed on the CPU oracle (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--check", type=int, default=8, help="sentences whose trees are checked against the oracle")
    ap.add_argument("--side", type=int, default=1, help="also run the c2f, mixed-length and serving workloads")
    ap.add_argument("--c2f-sentences", type=int, default=5000)
    ap.add_argument("--mixed-sentences", type=int, default=2000)
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="seconds of Poisson arrivals per rate")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many MiB three times (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch
        n = args.copy_probe << 18  # float32 elements
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 4 * n, "launches": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for {args.gpus}); "
              "sentences are independent and would shard by sentence across GPUs.", file=sys.stderr)
        sys.exit(2)

    import puck_amd as puck
    t0 = time.time()
    prefix = puck.synth_wsj6_cached(seed=6)
    g = puck.Grammar.load_text(prefix, puck.VITERBI)
    words, offsets = g.sample(seed=args.seed, n=args.sentences, mode=1, min_len=1, max_len=40)
    setup_s = time.time() - t0
    lens = np.diff(offsets)
    parser = puck.CLParser([g], device=0, mem_bytes=64 << 30, items_per_lane=args.ipl)
    print(f"[bench] setup {setup_s:.1f}s; {len(lens)} sentences, mean length {lens.mean():.1f}, "
          f"max {lens.max()}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        parser.parse((words, offsets))
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    hip_sync()
    dev_ms, wall, kern_ms, kern_apps, launches = 0.0, 0.0, 0.0, 0.0, 0
    trees = None
    for i in range(args.steps):
        t = time.time()
        trees = parser.parse((words, offsets))
        hip_sync()
        wall += time.time() - t
        st = parser.stats()
        dev_ms += st["device_ms"]
        ms, apps, fam = parser.launches()
        binary = fam < 4  # inside binary families: the dominant kernel
        kern_ms += float(ms[binary].sum())
        kern_apps += float(apps[binary].sum())
        launches += int(binary.sum())
        print(f"[bench] step {i}: device {st['device_ms']:.1f} ms, wall {time.time() - t:.2f} s",
              file=sys.stderr, flush=True)
    n = len(lens) * args.steps
    default_workload = (args.sentences, args.seed, args.ipl) == (2000, 2, 2)
    value = n / (dev_ms / 1e3)
    achieved = 3.0 * kern_apps / (kern_ms / 1e3)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "sentences/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dev_ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded PCFG samples of a synthetic grammar with the wsj_6 shape",
        "config": {"workload": "exhaustive Viterbi parse (tags + inside + decode), synthetic wsj_6-shape grammar, "
                               f"{len(lens)} sentences, lengths 1-40 log-normal",
                   "sentences": int(len(lens)), "mean_length": round(float(lens.mean()), 2),
                   "mode": "plain (one launch per span and rule family)", "items_per_lane": args.ipl},
        "e2e_value": round(n / wall, 2),
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 2),
                     "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK, 4),
                     "traffic": profiled_traffic() if default_workload else None,
                     "kernel": "k_rules (inside binary families)", "launches": launches,
                     "avg_launch_ms": round(kern_ms / max(launches, 1), 4),
                     "ops_per_launch": round(3.0 * kern_apps / max(launches, 1), 1)},
    }
    # Parity spot check against the oracle, and the CPU baseline on a bounded sample.
    if args.cpu_sample > 0:
        base, ref_trees, tree_off = cpu_baseline(puck, prefix, words, offsets, args.cpu_sample, args.cpu_threads)
        line["cpu_baseline"] = base
        k = min(args.check, args.cpu_sample)
        ok = all(np.array_equal(trees[i], ref_trees[tree_off[i]:tree_off[i + 1]]) for i in range(k))
        line["trees_match_oracle"] = f"{k}/{k}" if ok else "MISMATCH"
    if args.side:
        line.update(side_workloads(puck, g, prefix, args))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
