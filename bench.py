"""Benchmark of the CKY parse path on one MI355X.

Workload (the headline, config 2): exhaustive parsing with the synthetic wsj_6-shape grammar
(1,084 refined states = 502 nonterminals + 582 tags, 1,725,570 binary and 125,619 unary rules; the
real wsj_6 rule files are not in the reference checkout) of 2,000 seeded PCFG-sampled sentences,
lengths 1-40 from a discretized log-normal (mean ~23). One step = CLParser.insideOutside on the
2,000 sentences as config 2 states it: tag scores, FP32 inside pass, outside pass, then the
Viterbi decode (max-product in log space, whose per-cell scaling exponent is identically 0).

Two forms run on the same sentences and must return identical trees:
  plain    one launch per span and rule family, one batch (the reference's schedule)
  batched  the batched engine: length buckets on concurrent HIP streams, one persistent
           work-queue launch per span for all rule families, per-shape HIP graphs
`value` is the batched engine's sentences/s from device time (HIP events from its first kernel to
its last; inputs resident); `plain` and `batched_vs_plain` stand beside it. The roofline is that of
the persistent rule kernel (k_rules_q, inside and outside binary families) over the same work, from
per-launch HIP events of timed persistent parses on the engine stream; `hbm` states the algorithmic
chart and tree bytes of a step against 8 TB/s. Side workloads: config 1 (wsj_1), coarse-to-fine
Viterbi and MBR (5,000 sentences), mixed lengths to 80 (20,000 sentences), each plain vs batched,
and Poisson serving with p50/p99 latency. CPU baselines: the oracle on the same inputs. Prints one
JSON line.
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


def _oracle_layers(puck, specs):
    from oracle.grammar_ref import RefGrammar
    from oracle import oracle as O
    out = []
    for k, (prefix, semi) in enumerate(specs):
        out.append(O.Oracle(RefGrammar(prefix, semi), final_layer=k == len(specs) - 1))
    return out


def cpu_time(layers, words, offsets, sample, threads, with_outside, label):
    """The oracle's batch driver (C restatement, OpenMP over sentences) on the first `sample`
    sentences; returns the cpu_baseline record and the oracle's results."""
    from oracle import oracle as O
    lens = np.diff(offsets)[:sample]
    sub_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    sub_words = np.ascontiguousarray(words[:sub_off[-1]], np.int32)
    t0 = time.time()
    res = O.batch(layers, sub_words, sub_off, with_outside=with_outside, threads=threads)
    dt = time.time() - t0
    rec = dict(value=len(lens) / dt, unit="sentences/s", cores=res["threads"], kind="port",
               sample=f"{label}: {len(lens)} of {len(offsets) - 1} sentences (oracle C restatement, "
                      f"OpenMP over sentences, {dt:.1f} s)")
    return rec, res


PROFILE = os.path.join(ROOT, "profiles", "k_rules_q_wsj6_inside_outside_viterbi.md")
HBM_PEAK = 8.0e12  # B/s


def profiled_traffic():
    """HBM bytes per k_rules_q launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of
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


def timed_batched(bp, words, offsets, steps):
    """Eager run, graph capture run, then `steps` replays; device ms per parse and the trees."""
    trees = None
    for _ in range(2):
        trees = bp.parse((words, offsets))
    dev = 0.0
    for _ in range(steps):
        trees = bp.parse((words, offsets))
        dev += bp.stats()["device_ms"]
    return dev / steps, trees


def same_trees(a, b):
    return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))


def plain_vs_batched(puck, layers, words, offsets, steps, cfg, label, outside_final=False,
                     kernel_labels=("k_rules (fine inside binary, masked)",
                                    "k_rules_q (fine inside binary, masked, persistent)")):
    """Plain and batched forms on one workload: rates, ratio, tree identity, and the fine layer's
    rule-kernel roofline in both forms (plain k_rules launches; persistent k_rules_q launches of the
    kind the batched engine runs, timed by per-launch HIP events on a parser's stream)."""
    lens = np.diff(offsets)
    p = puck.CLParser(layers, device=0, mem_bytes=64 << 30, outside_final=outside_final)
    ms_p, launches = timed_parses(p, words, offsets, steps)
    ref = p.parse((words, offsets))
    del p
    q = puck.CLParser(layers, device=0, mem_bytes=64 << 30, items_per_lane=cfg["items_per_lane"], persistent=True,
                      outside_final=outside_final)
    q.configure(True, piece_target=cfg["piece_target"], reduce_targets=cfg["reduce_targets"])
    _, launches_q = timed_parses(q, words, offsets, 1)
    del q
    bp = puck.BatchParser(layers, mem_bytes=96 << 30, **cfg)
    ms_b, trees = timed_batched(bp, words, offsets, steps)
    st = bp.stats()
    del bp
    out = {"workload": label, "sentences": int(len(lens)), "mean_length": round(float(lens.mean()), 2),
           "value": round(len(lens) / (ms_b / 1e3), 1), "unit": "sentences/s", "ms_per_parse": round(ms_b, 2),
           "plain": round(len(lens) / (ms_p / 1e3), 1), "plain_ms_per_parse": round(ms_p, 2),
           "batched_vs_plain": round(ms_p / ms_b, 3), "trees_identical": same_trees(trees, ref),
           "buckets": int(st["buckets"]), "graph_replays": int(st["graph_replays"]), "batched": cfg,
           "roofline_plain": fine_roofline(launches, kernel_labels[0]),
           "roofline_batched": fine_roofline(launches_q, kernel_labels[1])}
    hb = pipeline_bytes(puck, layers, lens, outside_final)
    out["hbm"] = {"algorithmic_bytes_per_parse": hb, "achieved": round(hb / (ms_b / 1e3) / 1e9, 1),
                  "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(hb / (ms_b / 1e3) / HBM_PEAK, 5),
                  "note": "every layer's chart passes written and read once, masks written and read, 16 B "
                          "per tree node; the path is VALU-bound, so this fraction is small"}
    print(f"[bench] {label}: batched {out['value']} vs plain {out['plain']} sentences/s", file=sys.stderr, flush=True)
    return out


def fine_roofline(launches, kernel):
    """fp32 VALU roofline of the last layer's inside binary rule launches (pruned rule-apps)."""
    ms, apps, fam = launches
    last = fam.max() // 16 if len(fam) else 0
    sel = (fam // 16 == last) & ((fam % 16) < 4)
    t = float(ms[sel].sum()) / 1e3
    ach = 3.0 * float(apps[sel].sum()) / t if t > 0 else 0.0
    return {"bound": "valu", "achieved": round(ach / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 2),
            "unit": "TFLOP/s", "frac": round(ach / VALU_PEAK, 4), "kernel": kernel,
            "launches": int(sel.sum()), "rule_apps": float(apps[sel].sum()),
            "avg_launch_ms": round(float(ms[sel].mean()), 4) if sel.any() else None}


def pipeline_bytes(puck, layers, lens, outside_final):
    """Algorithmic HBM bytes of a coarse-to-fine parse: every layer's chart passes (inside, and
    outside where it runs: coarse layers, a Real final layer, insideOutside) each write and read
    their chart once, every coarse layer writes its masks and the next reads them, and the decode
    writes 16 bytes per tree node."""
    lens = np.asarray(lens, np.int64)
    cells = int(np.sum(lens * (lens + 1)))
    total = 16 * int(np.sum(2 * lens - 1))
    for k, g in enumerate(layers):
        last = k == len(layers) - 1
        passes = 2 if (not last or outside_final or g.semiring != puck.VITERBI) else 1
        total += passes * 2 * 4 * g.info.cell_size * cells
        if not last:
            total += 2 * 4 * g.info.mask_words * cells
    return total


def chart_bytes(lens, cell_size, passes):
    """Algorithmic HBM bytes of a parse: each chart pass writes its chart (4 cellSize n(n+1) bytes)
    once and reads it once; the decode writes 16 bytes per tree node."""
    lens = np.asarray(lens, np.int64)
    return int(np.sum(passes * 2 * 4 * cell_size * lens * (lens + 1) + 16 * (2 * lens - 1)))


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


SIDE = ("config1", "c2f_viterbi", "c2f_mbr", "c2f_reproject", "mixed_lengths", "serving", "serving_short")


def side_workloads(puck, g6v, prefix, args, which=SIDE):
    """Config 1, coarse-to-fine, mixed-length and serving workloads, measured beside the headline,
    with the oracle's CPU time on config 1 (in full) and config 3 (a stated sample). Each batched
    workload runs its own autotuned configuration."""
    out = {}
    cfg_io, cfg = batched_config(args, TUNED_IO), batched_config(args, TUNED_C2F)
    cfg_mixed = batched_config(args, TUNED)
    g1 = puck.Grammar.load_text(puck.grammar_path("wsj_1"), puck.VITERBI)
    words, offsets = g6v.sample(seed=3, n=args.c2f_sentences, mode=0, min_len=1, max_len=40)
    if "config1" in which:
        # Config 1: wsj_1, 200 PCFG sentences of 5-20 words, exhaustive insideOutside + Viterbi.
        w1, o1 = g1.sample(seed=1, n=200, mode=0, min_len=5, max_len=20)
        c1 = dict(cfg_io, outside_final=True)
        rec = plain_vs_batched(puck, [g1], w1, o1, 3, c1, "config 1: wsj_1 exhaustive insideOutside + Viterbi, "
                               "200 PCFG sentences 5-20", outside_final=True,
                               kernel_labels=("k_rules (inside binary)", "k_rules_q (inside binary, persistent)"))
        if args.cpu_sample > 0:
            base, res = cpu_time(_oracle_layers(puck, [(puck.grammar_path("wsj_1"), "viterbi")]), w1, o1, 200,
                                 args.cpu_threads, True, "config 1 in full")
            rec["cpu_baseline"] = base
            bp = puck.BatchParser([g1], mem_bytes=8 << 30, **c1)
            rec["trees_match_oracle"] = f"{200}/{200}" if same_trees(bp.parse((w1, o1)), res["trees"]) else "MISMATCH"
            del bp
        out["config1"] = rec
    if "c2f_viterbi" in which:
        out["c2f_viterbi"] = plain_vs_batched(
            puck, [g1, g6v], words, offsets, 2, cfg,
            f"config 3 inputs: wsj_1 -> synthetic wsj_6-shape Viterbi, masks at -9, {args.c2f_sentences} PCFG "
            "sentences 1-40")
    if "c2f_mbr" in which:
        g6r = puck.Grammar.load_text(prefix, puck.REAL)
        out["c2f_mbr"] = plain_vs_batched(
            puck, [g1, g6r], words, offsets, 2, cfg,
            f"config 3: wsj_1 -> synthetic wsj_6-shape Real inside/outside + MBR, {args.c2f_sentences} PCFG "
            "sentences 1-40")
        if args.cpu_sample > 0:
            t = time.time()
            layers = _oracle_layers(puck, [(puck.grammar_path("wsj_1"), "viterbi"), (prefix, "real")])
            load_s = time.time() - t
            n3 = len(offsets) - 1 if args.cpu_full else min(args.c2f_cpu_sample, len(offsets) - 1)
            base, res = cpu_time(layers, words, offsets, n3, args.cpu_threads, False,
                                 "config 3" + (" in full" if n3 == len(offsets) - 1 else " (seeded prefix)"))
            base["sample"] += f"; grammar load {load_s:.1f} s excluded"
            out["c2f_mbr"]["cpu_baseline"] = base
            p = puck.CLParser([g1, g6r], device=0, mem_bytes=64 << 30)
            sub = (words[:offsets[n3]], offsets[:n3 + 1])
            gm = p.masks(sub)
            out["c2f_mbr"]["masks_match_oracle"] = bool(np.array_equal(gm, res["masks"]))
            parts = p.partitions(sub).astype(np.float64)
            rel = np.abs(parts - res["partitions"]) / np.abs(res["partitions"].astype(np.float64))
            out["c2f_mbr"]["partition_max_rel_err_vs_oracle"] = float(rel.max())
            mb = p.parse(sub)
            same = sum(np.array_equal(a[:, :4], b[:, :4]) for a, b in zip(mb, res["mbr"]))
            out["c2f_mbr"]["mbr_trees_match_oracle"] = f"{same}/{n3}"
            del p, layers
        del g6r
    if "c2f_reproject" in which:
        # --reproject: masks keep one bit per wsj_1 substate; fine symbols project through the
        # wsj_6 split hierarchy.
        rc, rf = puck.Grammar.reproject(g1, g6v, coarse_level=1, fine_level=6)
        out["c2f_reproject"] = plain_vs_batched(
            puck, [rc, rf], words, offsets, 2, cfg,
            f"wsj_1 -> synthetic wsj_6-shape Viterbi, reprojected through the wsj_6 hierarchy "
            f"({rc.info.num_coarse} coarse substates), masks at -9, {args.c2f_sentences} PCFG sentences 1-40")
        del rc, rf
    if "mixed_lengths" in which:
        mw, mo = g6v.sample(seed=4, n=args.mixed_sentences, mode=2, min_len=1, max_len=80)
        out["mixed_lengths"] = plain_vs_batched(
            puck, [g1, g6v], mw, mo, 1, cfg_mixed,
            f"config 4: c2f Viterbi, {args.mixed_sentences} sentences, 70% log-normal 1-40 + 30% uniform 41-80")
    if "serving" in which:
        # Config 5: request lengths as in config 4 (70% log-normal 1-40, 30% uniform 41-80).
        pool_w, pool_o = g6v.sample(seed=5, n=20000, mode=2, min_len=1, max_len=80)
        pool = [np.ascontiguousarray(pool_w[pool_o[i]:pool_o[i + 1]]) for i in range(len(pool_o) - 1)]
        out["serving"] = []
        for rate in (200, 1000, 5000):
            r = serve(puck, [g1, g6v], pool, rate, args.serve_seconds, seed=rate)
            r["lengths"] = "config 4: 70% log-normal 1-40 + 30% uniform 41-80"
            out["serving"].append(r)
            print(f"[bench] serving {r}", file=sys.stderr, flush=True)
    if "serving_short" in which:
        # The same server on requests of the coarse-to-fine workload's lengths (PCFG samples 1-40).
        pool_w, pool_o = g6v.sample(seed=5, n=20000, mode=0, min_len=1, max_len=40)
        pool = [np.ascontiguousarray(pool_w[pool_o[i]:pool_o[i + 1]]) for i in range(len(pool_o) - 1)]
        out["serving_short"] = []
        for rate in (200, 1000, 5000):
            r = serve(puck, [g1, g6v], pool, rate, args.serve_short_seconds, seed=rate)
            r["lengths"] = "PCFG samples 1-40"
            out["serving_short"].append(r)
            print(f"[bench] serving_short {r}", file=sys.stderr, flush=True)
    return out


# Autotuned configurations (python -m puck_amd.autotune), one per workload: the exhaustive
# insideOutside workload's serves the headline and config 1, the coarse-to-fine workload's the
# config-3 workloads, the mixed-length workload's config 4.
TUNED = os.path.join(ROOT, "puck_amd", "tuning", "batched_wsj6.json")
TUNED_IO = os.path.join(ROOT, "puck_amd", "tuning", "batched_wsj6_inside_outside.json")
TUNED_C2F = os.path.join(ROOT, "puck_amd", "tuning", "batched_wsj6_c2f.json")


def batched_config(args, path=TUNED):
    """The autotuned configuration for a workload when one is committed (else the mixed workload's,
    else the defaults)."""
    from puck_amd import autotune
    cfg = dict(autotune.DEFAULTS)
    if not args.no_tuned:
        for p in dict.fromkeys((path, TUNED)):
            if os.path.exists(p):
                cfg.update(autotune.load(p)["best"])
                break
    return cfg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sentences", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--ipl", type=int, default=2, help="items per lane of the rule kernels (1 or 2)")
    ap.add_argument("--cpu-sample", type=int, default=100,
                    help="headline sentences timed on the CPU oracle (0: skip every CPU baseline)")
    ap.add_argument("--c2f-cpu-sample", type=int, default=300, help="config 3 sentences timed on the CPU oracle")
    ap.add_argument("--cpu-full", action="store_true", help="time config 3 on the CPU in full (5,000 sentences)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--side", type=int, default=1, help="also run the c2f, mixed-length and serving workloads")
    ap.add_argument("--only", default="", help="comma list of side workloads to run alone, without the headline "
                                               "(" + ",".join(SIDE) + "); for profiling")
    ap.add_argument("--c2f-sentences", type=int, default=5000)
    ap.add_argument("--mixed-sentences", type=int, default=20000)
    ap.add_argument("--no-tuned", action="store_true", help="use the default batched configuration")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="seconds of Poisson arrivals per rate")
    ap.add_argument("--serve-short-seconds", type=float, default=10.0,
                    help="seconds per rate of the serving run on requests of length 1-40")
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
        print(f"bench.py measures one GPU only: the engine has no multi-GPU path (asked for {args.gpus}); "
              "sentences are independent and would shard by sentence across GPUs.", file=sys.stderr)
        sys.exit(2)

    import puck_amd as puck
    t0 = time.time()
    prefix = puck.synth_wsj6_cached(seed=6)
    g = puck.Grammar.load_text(prefix, puck.VITERBI)
    if args.only:
        which = tuple(w for w in args.only.split(",") if w)
        bad = [w for w in which if w not in SIDE]
        if bad:
            sys.exit(f"unknown workloads {bad}; choose from {SIDE}")
        print(json.dumps(side_workloads(puck, g, prefix, args, which)))
        return
    words, offsets = g.sample(seed=args.seed, n=args.sentences, mode=1, min_len=1, max_len=40)
    setup_s = time.time() - t0
    lens = np.diff(offsets)
    cfg = batched_config(args, TUNED_IO)
    print(f"[bench] setup {setup_s:.1f}s; {len(lens)} sentences, mean length {lens.mean():.1f}, "
          f"max {lens.max()}; batched config {cfg}", file=sys.stderr, flush=True)

    cfg = dict(cfg, outside_final=True)
    # Plain form (one launch per span and family) on the same sentences.
    plain = puck.CLParser([g], device=0, mem_bytes=64 << 30, items_per_lane=args.ipl, outside_final=True)
    ms_plain, _ = timed_parses(plain, words, offsets, args.steps)
    ref_trees = plain.parse((words, offsets))
    del plain
    print(f"[bench] plain {ms_plain:.1f} ms per parse", file=sys.stderr, flush=True)

    # Roofline of the persistent rule kernel: per-launch HIP events on the engine stream over timed
    # persistent parses of the same batch the batched engine runs (inside and outside binary
    # families: family ids 0-11 of layer 0).
    pers = puck.CLParser([g], device=0, mem_bytes=64 << 30, items_per_lane=cfg["items_per_lane"], persistent=True,
                         outside_final=True)
    pers.configure(True, piece_target=cfg["piece_target"], reduce_targets=cfg["reduce_targets"])
    pers.parse((words, offsets))
    kern_ms, kern_apps, launches = 0.0, 0.0, 0
    for i in range(args.steps):
        pers.parse((words, offsets))
        ms, apps, fam = pers.launches()
        binary = fam < 12
        kern_ms += float(ms[binary].sum())
        kern_apps += float(apps[binary].sum())
        launches += int(binary.sum())
    del pers

    # Batched engine: the timed steps.
    bp = puck.BatchParser([g], mem_bytes=96 << 30, **cfg)
    for i in range(max(args.warmup, 2)):
        bp.parse((words, offsets))
        print(f"[bench] batched warmup {i} done", file=sys.stderr, flush=True)
    hip_sync()
    dev_ms, wall = 0.0, 0.0
    trees = None
    for i in range(args.steps):
        t = time.time()
        trees = bp.parse((words, offsets))
        hip_sync()
        wall += time.time() - t
        st = bp.stats()
        dev_ms += st["device_ms"]
        print(f"[bench] step {i}: device {st['device_ms']:.1f} ms, wall {time.time() - t:.2f} s, "
              f"buckets {st['buckets']}, graph replays {st['graph_replays']}", file=sys.stderr, flush=True)
    del bp
    n = len(lens) * args.steps
    default_workload = (args.sentences, args.seed) == (2000, 2)
    value = n / (dev_ms / 1e3)
    achieved = 3.0 * kern_apps / (kern_ms / 1e3)
    step_s = dev_ms / args.steps / 1e3
    hbm_bytes = chart_bytes(lens, g.info.cell_size, passes=2)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "sentences/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 2),
        "ms_per_step": round(dev_ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded PCFG samples of a synthetic grammar with the wsj_6 shape",
        "config": {"workload": "config 2: exhaustive FP32 insideOutside (tags + inside + outside) then Viterbi decode, "
                               f"synthetic wsj_6-shape grammar, {len(lens)} sentences, lengths 1-40 log-normal",
                   "sentences": int(len(lens)), "mean_length": round(float(lens.mean()), 2),
                   "mode": "batched (length buckets, concurrent streams, persistent rule launches, HIP graphs)",
                   "batched": cfg},
        "plain": {"value": round(len(lens) / (ms_plain / 1e3), 2), "ms_per_step": round(ms_plain, 3),
                  "mode": "plain (one launch per span and rule family)"},
        "batched_vs_plain": round(ms_plain / (dev_ms / args.steps), 3),
        "trees_identical_to_plain": same_trees(trees, ref_trees),
        "e2e_value": round(n / wall, 2),
        "roofline": {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 2),
                     "unit": "TFLOP/s", "frac": round(achieved / VALU_PEAK, 4),
                     "traffic": profiled_traffic() if default_workload else None,
                     "kernel": "k_rules_q (inside and outside binary families, one persistent launch per span "
                               "and pass)",
                     "launches": launches, "avg_launch_ms": round(kern_ms / max(launches, 1), 4),
                     "ops_per_launch": round(3.0 * kern_apps / max(launches, 1), 1)},
        "hbm": {"algorithmic_bytes_per_step": hbm_bytes, "achieved": round(hbm_bytes / step_s / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(hbm_bytes / step_s / HBM_PEAK, 5),
                "note": "inside and outside charts each written and read once (4 x 608 x n(n+1) bytes) plus "
                        "16 B per tree node; the path is VALU-bound, so this fraction is small"},
    }
    # CPU baseline on a bounded sample (the oracle's config-2 step: inside, outside, decode), which
    # also checks the batched trees of those sentences.
    if args.cpu_sample > 0:
        t = time.time()
        layers = _oracle_layers(puck, [(prefix, "viterbi")])
        load_s = time.time() - t
        base, res = cpu_time(layers, words, offsets, args.cpu_sample, args.cpu_threads, True,
                             "config 2 (seeded prefix)")
        base["sample"] += f"; grammar load {load_s:.1f} s excluded"
        line["cpu_baseline"] = base
        k = len(res["trees"])
        ok = same_trees(trees[:k], res["trees"])
        line["trees_match_oracle"] = f"{k}/{k}" if ok else "MISMATCH"
        del layers, res
    if args.side:
        line.update(side_workloads(puck, g, prefix, args))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
