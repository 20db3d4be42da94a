#!/usr/bin/env python3
"""Engine autotuner for a resident batch.

Tries every engine x block size x tile-sweep x control candidate on the same
resident batch, replays each candidate (hipGraph replay when the candidate
uses graph control, a persistent cooperative launch otherwise) and keeps the
fastest.  Each candidate must reproduce the first candidate's target distances
bit for bit, otherwise the tuner stops with an error: every schedule of a
monotone fp32 relaxation reaches the same fixed point, so a difference is a
defect, never a tuning trade-off.

Per-class choice: the queries are split into small / medium / large classes by
box side, the engines are timed on each class alone, and the mixed assignment
is kept only when it beats the best uniform engine on the whole batch.

Output: tuning/tuning_table.json (one entry per workload name, read by
bench.py) and a TSV record of every trial (tuning/trials_<workload>.tsv), so a
decision can be audited and replayed.  The reference keeps the same kind of
evidence as shootout manifests (src/engine_shootout.cpp:787-1567, TSV at
:2189).
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import bfnew_amd as bfn  # noqa: E402

WORKLOADS = {
    "c3_batched_128x128x512_q256": dict(tiles=(128, 128, 512), seed=3, queries=256, query_seed=33,
                                        side=(4, 64), unbounded=0.01),
    "c2_plain_64x64x256_q64": dict(tiles=(64, 64, 256), seed=2, queries=64, query_seed=2,
                                   side=(4, 32), unbounded=0.0),
}


def log(*a):
    print("[autotune]", *a, file=sys.stderr, flush=True)


def query_class(q):
    if q.unbounded or q.box_side > 32:
        return "large"
    return "medium" if q.box_side >= 16 else "small"


def target_bits(ws):
    return np.array([p["distance"] for p in ws.paths()], np.float32).view(np.uint32)


def time_candidate(dev, dg, batch, cand, reps):
    ws = bfn.Workspace(dev)
    try:
        opts = bfn.run_options(engine=cand["engine"], control_mode=cand["control"],
                               block_size=cand["block"], rounds_per_chunk=cand.get("chunk", 8))
        tune = bfn.tuning(tile_sweeps=cand["sweeps"], items_per_claim=cand["items_per_claim"],
                          use_graph=cand.get("use_graph", 0),
                          graph_rounds=cand.get("graph_rounds", 0),
                          boundary_mode=cand.get("boundary_mode", 0),
                          min_waves=cand.get("min_waves", 0))
        ws.submit(dg, batch, opts, tune)
        st, _, _ = ws.status()
        if not st.converged:
            return None, None, None
        bits = target_bits(ws)
        ws.rerun()
        times = []
        for _ in range(reps):
            t = time.perf_counter()
            ws.rerun()
            times.append((time.perf_counter() - t) * 1e3)
        rounds = ws.counters().rounds_executed
        return statistics.median(times), bits, rounds
    finally:
        ws.close()


def candidates(engines, blocks, quick):
    """Persistent candidates; each is followed by its hipGraph twin, whose
    captured round count is set from the persistent run."""
    out = []
    for engine, block in itertools.product(engines, blocks):
        jacobi = engine == bfn.JACOBI_PULL
        waves = (0, 6, 8) if (not jacobi and block in (256, 512)) else (0,)
        for sw in ((0, 1) if jacobi else (0,)):
            for mode in ((0,) if jacobi else (0, 1)):
                for mw in waves:
                    for ipc in ((2,) if quick else (1, 2)):
                        out.append(dict(engine=engine, block=block, sweeps=sw, items_per_claim=ipc,
                                        boundary_mode=mode, min_waves=mw,
                                        control=bfn.PERSISTENT))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3_batched_128x128x512_q256", choices=sorted(WORKLOADS))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--blocks", default="128,256,512")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning"))
    args = ap.parse_args()
    cfg = WORKLOADS[args.workload]

    dev = bfn.Device(0)
    hg = bfn.HostGraph.synthetic(*cfg["tiles"], cfg["seed"], host_graph=False)
    dg = bfn.DeviceGraph(dev, hg.layout())
    qs = hg.synthetic_queries(cfg["queries"], seed=cfg["query_seed"], side_min=cfg["side"][0],
                              side_max=cfg["side"][1], unbounded_fraction=cfg["unbounded"])
    tiles = [hg.query_tiles(q) for q in qs]
    batch = bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], tiles)

    blocks = [int(b) for b in args.blocks.split(",")]
    engines = (bfn.JACOBI_PULL, bfn.DENSE_CHAOTIC_PUSH, bfn.FRONTIER_PUSH)
    trials = []
    ref_bits = None

    def trial(cand):
        """Times one candidate, checks its distances, records it; returns its
        round count (None when it did not converge)."""
        nonlocal ref_bits
        ms, bits, rounds = time_candidate(dev, dg, batch, cand, args.reps)
        if ms is None:
            log("did not converge:", cand)
            return None
        if ref_bits is None:
            ref_bits = bits
        elif not np.array_equal(bits, ref_bits):
            raise SystemExit(f"candidate {cand} changed target distances")
        trials.append(dict(cand, ms=ms, rounds=rounds, scope="all"))
        log(f"{bfn.ENGINE_NAMES[cand['engine']]:>20} block {cand['block']:4d} sweeps "
            f"{cand['sweeps']} boundary {cand['boundary_mode']} waves {cand['min_waves']} "
            f"ipc {cand['items_per_claim']} control "
            f"{bfn.CONTROL_NAMES[cand['control']]}{'+graph' if cand.get('use_graph') else ''}: "
            f"{ms:.3f} ms, {rounds} rounds")
        return rounds

    def with_graph_twin(cand):
        rounds = trial(cand)
        if rounds is not None and not args.quick:
            # Push schedules differ by a round or two between runs: capture a
            # small margin so that replays rarely need the host-polled tail.
            margin = 0 if cand["engine"] == bfn.JACOBI_PULL else 2
            trial(dict(cand, control=bfn.CHUNKED, use_graph=1, graph_rounds=rounds + margin))

    for cand in candidates(engines, blocks, args.quick):
        with_graph_twin(cand)
    best = min(trials, key=lambda t: t["ms"])

    # Per-class engines at the best uniform shape.
    classes = {}
    for cls in ("small", "medium", "large"):
        idx = [i for i, q in enumerate(qs) if query_class(q) == cls]
        if not idx:
            continue
        sub = bfn.Batch([[qs[i].source] for i in idx], [[qs[i].target] for i in idx],
                        [tiles[i] for i in idx])
        timing = {}
        for engine in engines:
            cand = dict(best, engine=engine, sweeps=best["sweeps"] if engine == 0 else 0,
                        boundary_mode=0 if engine == 0 else best.get("boundary_mode", 0),
                        min_waves=0 if engine == 0 else best.get("min_waves", 0))
            ms, _, rounds = time_candidate(dev, dg, sub, cand, args.reps)
            if ms is not None:
                timing[engine] = ms
                trials.append(dict(cand, ms=ms, rounds=rounds, scope=cls))
        classes[cls] = min(timing, key=timing.get)
        log(f"class {cls} ({len(idx)} queries):",
            {bfn.ENGINE_NAMES[e]: round(v, 3) for e, v in timing.items()})
    entry = {k: best[k] for k in ("engine", "block", "sweeps", "control", "items_per_claim",
                                  "boundary_mode", "min_waves")}
    entry["use_graph"] = best.get("use_graph", 0)
    entry["graph_rounds"] = best.get("graph_rounds", 0)
    entry["ms"] = best["ms"]
    if len(set(classes.values())) > 1:
        batch.engines = [classes[query_class(q)] for q in qs]
        mixed = dict(best, engine=best["engine"])
        ms, bits, rounds = time_candidate(dev, dg, batch, mixed, args.reps)
        batch.engines = None
        if ms is not None:
            if not np.array_equal(bits, ref_bits):
                raise SystemExit("per-class engines changed target distances")
            trials.append(dict(mixed, ms=ms, rounds=rounds, scope="mixed"))
            log(f"per-class mix: {ms:.3f} ms vs uniform {best['ms']:.3f} ms")
            if ms < best["ms"]:
                entry["classes"] = {k: int(v) for k, v in classes.items()}
                entry["ms"] = ms
    os.makedirs(args.out, exist_ok=True)
    table_path = os.path.join(args.out, "tuning_table.json")
    table = {}
    if os.path.exists(table_path):
        with open(table_path) as f:
            table = json.load(f)
    table[args.workload] = entry
    with open(table_path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    with open(os.path.join(args.out, f"trials_{args.workload}.tsv"), "w") as f:
        cols = ["scope", "engine", "block", "sweeps", "boundary_mode", "min_waves", "items_per_claim",
                "control", "use_graph", "graph_rounds", "rounds", "ms"]
        f.write("\t".join(cols) + "\n")
        for t in trials:
            f.write("\t".join(str(t.get(c, 0)) for c in cols) + "\n")
    log("chosen:", json.dumps(entry))
    print(json.dumps({"workload": args.workload, "chosen": entry}))


if __name__ == "__main__":
    main()
