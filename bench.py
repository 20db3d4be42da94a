#!/usr/bin/env python3
"""Batched SSSP edge relaxations/s on one MI355X.

Workload (the headline configuration): a 128 x 128 tile grid with 512 nodes
per tile (8,388,608 nodes, 83,886,080 edges, fp32 node-entry costs uniform in
[1, 4]), 256 route queries whose bounding boxes have a side of 4-64 tiles
(log-uniform) with 1% unbounded, all solved in one batched launch.  A step is
one pass of the engine over the resident batch: label initialisation,
relaxation to the fixed point, per-query finalisation (graph, batch metadata
and queries already resident in HBM).

value  = work-normalised edge relaxations per second: per step, the sum over
         queries of the admitted edges whose source is reachable (exactly what
         Dijkstra relaxes), times steps, over the wall time of the timed steps.
Also reported: the plain one-query-per-launch mode on the 64 x 64 x 256 graph
(64 queries) beside the batched mode on the same queries, the roofline of the
relaxation kernel, and the reference CPU path timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bf-new_amd"))

import numpy as np  # noqa: E402

import bfnew_amd as bfn  # noqa: E402

HBM_PEAK_GBS = 8000.0

C3 = dict(name="c3_batched_128x128x512_q256", tiles=(128, 128, 512), seed=3, queries=256,
          query_seed=33, side=(4, 64), unbounded=0.01)
C2 = dict(name="c2_plain_64x64x256_q64", tiles=(64, 64, 256), seed=2, queries=64, query_seed=2,
          side=(4, 32), unbounded=0.0)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_tuning(workload):
    """Per-workload engine choices written by bf-new_amd/autotune.py."""
    path = os.path.join(ROOT, "bf-new_amd", "tuning", "tuning_table.json")
    if os.path.exists(path):
        with open(path) as f:
            table = json.load(f)
        if workload in table:
            return table[workload]
    return None


def query_class(q):
    if q.unbounded or q.box_side > 32:
        return "large"
    return "medium" if q.box_side >= 16 else "small"


def make_batch(hg, cfg):
    qs = hg.synthetic_queries(cfg["queries"], seed=cfg["query_seed"], side_min=cfg["side"][0],
                              side_max=cfg["side"][1], unbounded_fraction=cfg["unbounded"])
    tiles = [hg.query_tiles(q) for q in qs]
    return qs, bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], tiles)


def apply_tuning(batch, qs, tune_entry):
    t = dict(engine=bfn.JACOBI_PULL, block=256, sweeps=0, control=bfn.PERSISTENT,
             classes=None, items_per_claim=1, use_graph=0, graph_rounds=0)
    if tune_entry:
        t.update(tune_entry)
    if t.get("classes"):
        batch.engines = [int(t["classes"][query_class(q)]) for q in qs]
    return t


def timed_steps(ws, steps, warmup):
    for _ in range(warmup):
        ws.rerun()
    bfn.hip_lib().bfn_device_synchronize(ws.device.h)
    relax_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ws.rerun()
        relax_ms.append(ws.timing()[0])
    bfn.hip_lib().bfn_device_synchronize(ws.device.h)
    return time.perf_counter() - t0, relax_ms


def cpu_baseline(hg_view, qs, tiles, budget_s):
    """The reference CPU path (dijkstra_oracle built from the reference's own
    sources, oracle/_ref) on a bounded sample of the same queries, one core.
    Falls back to the oracle's C restatement when _ref is not built."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    V = hg_view["vertex_count"]
    kind = "reference" if oracle.ref_available() else "port"
    relax_total, seconds, n = 0, 0.0, 0
    # Spread over box sizes; boxes above 1024 tiles are left out to bound the
    # untimed induced-graph construction of the reference builder.
    order = [i for i in sorted(range(len(qs)), key=lambda i: len(tiles[i]))
             if len(tiles[i]) <= 1024]
    picked = order[::max(1, len(order) // 64)]
    wall = time.perf_counter()
    for i in picked:
        if time.perf_counter() - wall > 4 * budget_s:
            break
        sel = np.zeros(hg_view["tile_count"], np.uint8)
        sel[tiles[i]] = 1
        if kind == "reference":
            ns, relax = oracle.ref_csr_sssp(V, hg_view["row_offsets"], hg_view["destinations"],
                                            hg_view["weights"], hg_view["owner_tiles"], sel,
                                            [qs[i].source])
            seconds += ns * 1e-9
        else:
            t = time.perf_counter()
            _, relax = oracle.dijkstra(V, hg_view["row_offsets"], hg_view["destinations"],
                                       hg_view["weights"], hg_view["owner_tiles"], sel,
                                       [qs[i].source])
            seconds += time.perf_counter() - t
        relax_total += relax
        n += 1
        if seconds > budget_s:
            break
    return dict(value=relax_total / seconds, unit="edge relaxations/s", cores=1, kind=kind,
                sample=f"{n} queries of the headline workload spread over box sizes "
                       f"({relax_total} relaxations in {seconds:.2f} s of Dijkstra time; "
                       "induced-subgraph construction excluded)")


def plain_vs_batched(dev, steps):
    """Config 2: 64 queries one per launch (plain) vs all 64 in one launch."""
    hg = bfn.HostGraph.synthetic(*C2["tiles"], C2["seed"])
    dg = bfn.DeviceGraph(dev, hg.layout())
    qs, batch = make_batch(hg, C2)
    ws = bfn.Workspace(dev)
    ws.submit(dg, batch, bfn.run_options(engine=bfn.JACOBI_PULL, control_mode=bfn.PERSISTENT),
              bfn.tuning(tile_sweeps=0, count_relaxations=1))
    _, per, _ = ws.status()
    relax = [p.relaxations for p in per]
    elapsed, _ = timed_steps(ws, steps, 1)
    batched = sum(relax) * steps / elapsed
    out = {}
    for eng in (bfn.JACOBI_PULL, bfn.DENSE_CHAOTIC_PUSH, bfn.FRONTIER_PUSH):
        wss = []
        for i in range(len(qs)):
            one = bfn.Batch([[qs[i].source]], [[qs[i].target]], [batch.tiles[i]])
            w = bfn.Workspace(dev)
            w.submit(dg, one, bfn.run_options(engine=eng, control_mode=bfn.PER_ROUND),
                     bfn.tuning(tile_sweeps=1))
            wss.append(w)
        lat = []
        t0 = time.perf_counter()
        for w in wss:
            t = time.perf_counter()
            w.rerun()
            lat.append(time.perf_counter() - t)
        elapsed_plain = time.perf_counter() - t0
        out[bfn.ENGINE_NAMES[eng]] = dict(
            value=sum(relax) / elapsed_plain, p50_ms=float(np.percentile(lat, 50) * 1e3),
            p99_ms=float(np.percentile(lat, 99) * 1e3))
        for w in wss:
            w.close()
    best_plain = max(v["value"] for v in out.values())
    return dict(workload=C2["name"], batched_value=batched, plain=out,
                batched_over_best_plain=batched / best_plain)


def copy_calibration():
    """Three copies of 1 GiB between two device buffers: a kernel of known
    traffic (1 GiB read, 1 GiB written per dispatch) against which the HBM
    FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 3 x 1073741824 bytes copied")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-budget-s", type=float, default=10.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--engine", type=int, default=None, help="override the tuned engine")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(queries shard across GPUs as independent batches, see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    dev = bfn.Device(0)
    info = dev.info()
    log("device", info.arch.decode(), info.compute_units, "CUs")
    t = time.perf_counter()
    hg = bfn.HostGraph.synthetic(*C3["tiles"], C3["seed"], host_graph=False)
    log(f"graph generated in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    dg = bfn.DeviceGraph(dev, hg.layout())
    log(f"graph resident in {time.perf_counter() - t:.1f}s", dg.info().device_bytes >> 20, "MiB")
    qs, batch = make_batch(hg, C3)
    tune = apply_tuning(batch, qs, load_tuning(C3["name"]))
    if args.engine is not None:
        tune.update(engine=args.engine, classes=None)
        batch.engines = None
    opts = bfn.run_options(engine=tune["engine"], control_mode=tune["control"],
                           block_size=tune["block"])
    ws = bfn.Workspace(dev)
    # Validation run (untimed): work-normalised relaxation count and counters.
    ws.submit(dg, batch, bfn.run_options(engine=tune["engine"], control_mode=tune["control"],
                                         block_size=tune["block"], instrumentation=1),
              bfn.tuning(tile_sweeps=tune["sweeps"], count_relaxations=1,
                         items_per_claim=tune["items_per_claim"], use_graph=tune["use_graph"],
                         graph_rounds=tune["graph_rounds"]))
    st, per, _ = ws.status()
    if not st.converged:
        raise SystemExit("batch did not converge")
    relax_per_step = sum(p.relaxations for p in per)
    counters = ws.counters()
    # Timed steps with instrumentation off.
    ws.submit(dg, batch, opts, bfn.tuning(tile_sweeps=tune["sweeps"],
                                          items_per_claim=tune["items_per_claim"],
                                          use_graph=tune["use_graph"],
                                          graph_rounds=tune["graph_rounds"]))
    elapsed, relax_ms = timed_steps(ws, args.steps, args.warmup)
    value = relax_per_step * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    kernel_ms = float(np.mean(relax_ms))
    achieved = counters.algorithmic_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"{value / 1e9:.2f} G relax/s, {ms_per_step:.2f} ms/step, relax kernel {kernel_ms:.3f} ms, "
        f"{counters.tile_visits} tile visits, {counters.rounds_executed} rounds")

    result = {
        "metric": "batched SSSP edge relaxations/s + p99 query latency, 1 MI355X, vs plain mode",
        "value": value,
        "unit": "edge relaxations/s (work-normalised)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": C3["name"], "graph": "128x128 tiles x 512 nodes, 10 edges/node",
                   "queries": C3["queries"], "box_side": list(C3["side"]),
                   "unbounded_fraction": C3["unbounded"], "engine": bfn.ENGINE_NAMES[tune["engine"]],
                   "per_class_engines": tune.get("classes"), "block": tune["block"],
                   "tile_sweeps": tune["sweeps"],
                   "control": bfn.CONTROL_NAMES[tune["control"]] +
                   (f"+hipgraph({tune['graph_rounds']} rounds)" if tune["use_graph"] else ""),
                   "relaxations_per_step": relax_per_step},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "relax (tile visits)", "kernel_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": counters.algorithmic_bytes,
                     "edges_relaxed_per_launch": counters.edges_relaxed},
    }
    if not args.no_plain:
        pm = plain_vs_batched(dev, max(3, args.steps // 2))
        result["plain_mode"] = pm
        log("plain vs batched (C2):", json.dumps(pm))
    if not args.no_cpu_baseline:
        v = hg.layout_arrays()
        view = dict(vertex_count=v["vertex_count"], tile_count=v["tile_count"],
                    row_offsets=v["csr_row_offsets"].astype(np.uint64),
                    destinations=v["csr_destinations"], weights=v["csr_weights"],
                    owner_tiles=v["owner_tiles"])
        cb = cpu_baseline(view, qs, batch.tiles, args.cpu_budget_s)
        result["cpu_baseline"] = cb
        result["speedup_vs_cpu"] = value / cb["value"]
        log("cpu baseline:", json.dumps(cb))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
