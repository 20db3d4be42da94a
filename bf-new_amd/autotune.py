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
decision can be audited and replayed.

--shootout runs the per-query campaign of the reference's engine shootout on
the workload instead: every query alone under each configuration of a
catalog (the three engines, per-round host poll), one warmup and --reps timed
repetitions, and writes the reference's manifest, catalog and samples TSV
files (shootout.py) under tuning/shootout_<workload>/, with the
configuration of smallest median wall time.
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
import shootout  # noqa: E402

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
                          min_waves=cand.get("min_waves", 0),
                          claim_order=cand.get("claim_order", 0),
                          push_window=cand.get("push_window", 0.0))
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


def candidates(engines, blocks, quick, boundary_modes=(0, 1)):
    """Persistent candidates; each is followed by its hipGraph twin, whose
    captured round count is set from the persistent run."""
    out = []
    for engine, block in itertools.product(engines, blocks):
        jacobi = engine == bfn.JACOBI_PULL
        waves = (0, 6, 7, 8) if (not jacobi and block in (256, 512)) else (0,)
        for sw in ((0, 1) if jacobi else (0,)):
            for mode in ((0,) if jacobi else boundary_modes):
                for mw in waves:
                    for ipc in ((2,) if quick else (1, 2, 3)):
                        out.append(dict(engine=engine, block=block, sweeps=sw, items_per_claim=ipc,
                                        boundary_mode=mode, min_waves=mw,
                                        control=bfn.PERSISTENT))
    return out


def _fnv_words(values):
    """Two 64-bit FNV-1a lanes (offset bases 14695981039346656037 and its
    byte-reversal) over the little-endian u64 encoding of `values`."""
    a, b = 14695981039346656037, 0x25233184D3AB5CCB
    for v in values:
        for k in range(8):
            byte = (int(v) >> (8 * k)) & 0xFF
            a = ((a ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
            b = ((b ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return a, b


def shootout_campaign(args, cfg, dev, hg, dg, qs, tiles):
    """Every query alone under each catalog configuration: a point-Jacobi
    pilot for the expected round count, then one warmup and args.reps timed
    reruns per (configuration, query); returns the manifest, catalog and
    samples."""
    arrays = hg.layout_arrays()
    owner = arrays["owner_tiles"]
    tile_vertices = np.diff(arrays["tile_vertex_offsets"].astype(np.int64))
    src = np.repeat(np.arange(arrays["vertex_count"], dtype=np.int64),
                    np.diff(arrays["csr_row_offsets"].astype(np.int64)))
    src_tile, dst_tile = owner[src], owner[arrays["csr_destinations"]]
    graph_words = hg.fingerprint()
    # Query ids are the queries' positions in the workload.
    query_words = _fnv_words([v for i, (q, t) in enumerate(zip(qs, tiles)) for v in (i, q.source, q.target, len(t), *t)])
    fp = shootout.Fingerprint(tuple(int(w) for w in graph_words), query_words, len(qs))
    workload = shootout.Workload(shootout.WORKLOAD_SYNTHETIC, 1 + sum(cfg["tiles"]) * 1000 + cfg["queries"],
                                 args.workload)
    configs = [shootout.Tuning(i, engine, shootout.PER_ROUND, 1, 256, maximum_rounds=1 << 20)
               for i, engine in enumerate((bfn.JACOBI_PULL, bfn.DENSE_CHAOTIC_PUSH, bfn.FRONTIER_PUSH))]
    features = []
    for i, (q, t) in enumerate(zip(qs, tiles)):
        sel = np.zeros(arrays["tile_count"], bool)
        sel[np.asarray(t, np.int64)] = True
        one = bfn.Batch([[q.source]], [[q.target]], [t])
        ws = bfn.Workspace(dev)
        ws.submit(dg, one, bfn.run_options(engine=bfn.JACOBI_PULL, control_mode=bfn.PER_ROUND,
                                           rounds_per_chunk=1), bfn.tuning(tile_sweeps=1))
        _, per, _ = ws.status()
        ws.close()
        features.append(shootout.QueryFeatures(
            i, int(tile_vertices[sel].sum()), int(np.count_nonzero(sel[src_tile] & sel[dst_tile])),
            1, 1, max(1, int(per[0].rounds_completed))))
    bins = shootout.quartile_bins(features)
    manifest = shootout.Manifest(fp, workload, cfg["query_seed"], 0, 1, args.reps, len(qs),
                                 list(zip(features, bins)))
    catalog = shootout.Catalog(fp, workload, configs)
    samples, ordinal = [], 0
    for c in configs:
        for i, (q, t) in enumerate(zip(qs, tiles)):
            one = bfn.Batch([[q.source]], [[q.target]], [t])
            ws = bfn.Workspace(dev)
            ws.submit(dg, one, bfn.run_options(engine=c.engine, control_mode=c.control_mode,
                                               rounds_per_chunk=c.rounds_per_chunk, block_size=c.block_size,
                                               maximum_rounds=c.maximum_rounds),
                      bfn.tuning(tile_sweeps=1))
            for rep in range(args.reps):
                t0 = time.perf_counter()
                ws.rerun()
                wall = (time.perf_counter() - t0) * 1e3
                st, per, work = ws.status()
                timeline = ws.timing()[1]
                s = shootout.Sample(fp, workload, shootout.RUN_TIMING, ordinal, rep, i,
                                    c.configuration_id, c.engine, c.control_mode, correctness_passed=True)
                s.status.update(final_distance_slot=st.final_distance_slot, converged=st.converged,
                                rounds_completed=st.rounds_completed, reached_target_mask=st.reached_target_mask,
                                bounding_box_miss_mask=st.bounding_box_miss_mask,
                                valid_lane_mask=st.valid_lane_mask, active_lane_mask=st.active_lane_mask,
                                converged_lane_mask=st.converged_lane_mask, stop_reason=st.stop_reason,
                                error_bits=st.error_bits)
                for k in shootout.WORK_FIELDS:
                    s.work[k] = int(getattr(work, k))
                s.timing["timeline"] = shootout.evidence(timeline)
                s.timing["wall"] = shootout.evidence(wall)
                samples.append(s)
                ordinal += 1
            ws.close()
    return manifest, catalog, samples


def write_shootout(args, cfg, dev, hg, dg, qs, tiles):
    manifest, catalog, samples = shootout_campaign(args, cfg, dev, hg, dg, qs, tiles)
    out = os.path.join(args.out, f"shootout_{args.workload}")
    os.makedirs(out, exist_ok=True)
    texts = {"manifest.v2.tsv": shootout.serialize_manifest(manifest),
             "catalog.v2.tsv": shootout.serialize_catalog(catalog),
             "samples.v2.tsv": shootout.serialize_samples(manifest, catalog.tunings, samples,
                                                          shootout.RUN_TIMING)}
    for name, text in texts.items():
        with open(os.path.join(out, name), "w", newline="\n") as f:
            f.write(text)
    # Replay check: the files read back to the same records.
    assert shootout.deserialize_manifest(texts["manifest.v2.tsv"]) == manifest
    assert shootout.samples_equal(shootout.deserialize_samples(manifest, catalog.tunings, texts["samples.v2.tsv"],
                                                               shootout.RUN_TIMING), samples)
    medians = {}
    for c in catalog.tunings:
        walls = [s.timing["wall"][1] for s in samples if s.configuration_id == c.configuration_id]
        medians[c.configuration_id] = statistics.median(walls)
    best = min(medians, key=medians.get)
    decision = {"workload": args.workload, "selected_configuration": best,
                "engine": bfn.ENGINE_NAMES[catalog.tunings[best].engine],
                "median_wall_ms": {bfn.ENGINE_NAMES[c.engine]: round(medians[c.configuration_id], 4)
                                   for c in catalog.tunings},
                "samples": len(samples), "directory": os.path.relpath(out, HERE)}
    log("shootout:", json.dumps(decision))
    print(json.dumps(decision))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3_batched_128x128x512_q256", choices=sorted(WORKLOADS))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--blocks", default="128,256,512")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--engines", default="0,1,2", help="engines to try (0 Jacobi pull, 1 dense push, 2 frontier push)")
    ap.add_argument("--windows", default="2,4,8", help="push windows to try besides none")
    ap.add_argument("--boundary-modes", default="0,1", help="push boundary modes to try")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning"))
    ap.add_argument("--shootout", action="store_true",
                    help="run the per-query shootout campaign and write its TSV records")
    args = ap.parse_args()
    cfg = WORKLOADS[args.workload]

    dev = bfn.Device(0)
    hg = bfn.HostGraph.synthetic(*cfg["tiles"], cfg["seed"], host_graph=False)
    dg = bfn.DeviceGraph(dev, hg.layout())
    qs = hg.synthetic_queries(cfg["queries"], seed=cfg["query_seed"], side_min=cfg["side"][0],
                              side_max=cfg["side"][1], unbounded_fraction=cfg["unbounded"])
    tiles = [hg.query_tiles(q) for q in qs]
    batch = bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], tiles)
    if args.shootout:
        write_shootout(args, cfg, dev, hg, dg, qs, tiles)
        return

    blocks = [int(b) for b in args.blocks.split(",")]
    engines = tuple(int(e) for e in args.engines.split(","))
    windows = [float(w) for w in args.windows.split(",") if float(w) > 0]
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
            f"{bfn.CONTROL_NAMES[cand['control']]}{'+graph' if cand.get('use_graph') else ''}"
            f"{' slot-order' if cand.get('claim_order') else ''}"
            f"{' window ' + str(cand['push_window']) if cand.get('push_window') else ''}: "
            f"{ms:.3f} ms, {rounds} rounds")
        return rounds

    def with_graph_twin(cand):
        rounds = trial(cand)
        if rounds is not None and not args.quick:
            # Push schedules differ by a round or two between runs: capture a
            # small margin so that replays rarely need the host-polled tail.
            margin = 0 if cand["engine"] == bfn.JACOBI_PULL else 2
            twin = dict(cand, control=bfn.CHUNKED, use_graph=1, graph_rounds=rounds + margin)
            trial(twin)
            # and in slot order (queue bits compacted between round launches)
            trial(dict(twin, claim_order=1))
        if rounds is not None and cand["engine"] != bfn.JACOBI_PULL:
            # Push windows: a persistent run gives the round count to capture.
            for w in windows:
                wr = trial(dict(cand, push_window=w))
                if wr is not None:
                    trial(dict(cand, push_window=w, control=bfn.CHUNKED, use_graph=1,
                               graph_rounds=wr + 2))

    modes = tuple(int(m) for m in args.boundary_modes.split(","))
    for cand in candidates(engines, blocks, args.quick, modes):
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
    entry["claim_order"] = best.get("claim_order", 0)
    entry["push_window"] = best.get("push_window", 0.0)
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
                "control", "use_graph", "graph_rounds", "claim_order", "push_window", "rounds", "ms"]
        f.write("\t".join(cols) + "\n")
        for t in trials:
            f.write("\t".join(str(t.get(c, 0)) for c in cols) + "\n")
    log("chosen:", json.dumps(entry))
    print(json.dumps({"workload": args.workload, "chosen": entry}))


if __name__ == "__main__":
    main()
