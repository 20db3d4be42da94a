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


def profiled_traffic(workload, tune):
    """HBM bytes per step of the relaxation kernels, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes over this workload and configuration and
    recorded in profiles/relax_traffic.json (None when the recorded
    configuration is not the one being run)."""
    path = os.path.join(ROOT, "profiles", "relax_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        entry = json.load(f).get(workload)
    if not entry:
        return None, None
    cfg = entry["config"]
    keys = ("engine", "block", "sweeps", "boundary_mode", "use_graph", "min_waves", "claim_order",
            "push_window", "items_per_claim")
    if any(float(cfg.get(k, 0)) != float(tune.get(k, 0)) for k in keys):
        return None, None
    return entry["hbm_bytes_per_step"], entry["source"]


def query_class(q):
    if q.unbounded or q.box_side > 32:
        return "large"
    return "medium" if q.box_side >= 16 else "small"


def make_batch(hg, cfg):
    qs = hg.synthetic_queries(cfg["queries"], seed=cfg["query_seed"], side_min=cfg["side"][0],
                              side_max=cfg["side"][1], unbounded_fraction=cfg["unbounded"])
    tiles = [hg.query_tiles(q) for q in qs]
    return qs, bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], tiles)


def scheduling(tune):
    """The claim order and push window of a tuning entry (absent: off)."""
    return dict(claim_order=tune.get("claim_order", 0), push_window=tune.get("push_window", 0.0))


def tuned_defaults(tune_entry):
    t = dict(engine=bfn.JACOBI_PULL, block=256, sweeps=0, control=bfn.PERSISTENT,
             classes=None, items_per_claim=1, use_graph=0, graph_rounds=0, boundary_mode=0,
             min_waves=0, claim_order=0, push_window=0.0)
    if tune_entry:
        t.update(tune_entry)
    return t


def apply_tuning(batch, qs, tune_entry):
    t = tuned_defaults(tune_entry)
    if t.get("classes"):
        batch.engines = [int(t["classes"][query_class(q)]) for q in qs]
    return t


def timed_steps(ws, steps, warmup, paths=False):
    """Times `steps` reruns of the resident batch (with paths: each followed
    by compact-path reconstruction of every target and the copy of the
    summaries, offsets and path arenas to host arrays).  Returns the wall
    time, the relaxation time of each step (HIP events on the workspace
    stream), the round launches of the last step and the wall time of each
    paths call."""
    for _ in range(warmup):
        ws.rerun()
        if paths:
            ws.path_arrays()
    bfn.hip_lib().bfn_device_synchronize(ws.device.h)
    relax_ms, paths_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ws.rerun()
        relax_ms.append(ws.timing()[0])
        if paths:
            t = time.perf_counter()
            ws.path_arrays()
            paths_ms.append((time.perf_counter() - t) * 1e3)
    bfn.hip_lib().bfn_device_synchronize(ws.device.h)
    return time.perf_counter() - t0, relax_ms, ws.timing()[2], paths_ms


def persistent_control(dev, dg, batch, tune, target_bits, steps):
    """The same batch and tuning under one persistent cooperative launch per
    step (grid barrier between rounds) instead of the tuned control, with
    the entry's persistent overrides (its wave bound): wall and
    relaxation-kernel ms per step, and whether every target distance has
    the bits of the tuned run."""
    t = dict(tune, **{k: v for k, v in (tune.get("persistent") or {}).items() if k == "min_waves"})
    ws = bfn.Workspace(dev)
    try:
        ws.submit(dg, batch, bfn.run_options(engine=t["engine"], control_mode=bfn.PERSISTENT,
                                             block_size=t["block"]),
                  bfn.tuning(tile_sweeps=t["sweeps"], items_per_claim=t["items_per_claim"],
                             boundary_mode=t["boundary_mode"], min_waves=t["min_waves"],
                             push_window=t.get("push_window", 0.0)))
        st, _, _ = ws.status()
        bits = np.array([p["distance"] for p in ws.paths()], np.float32).view(np.uint32)
        elapsed, relax_ms, _, _ = timed_steps(ws, steps, 1)
        return {"control": "persistent_cooperative", "min_waves": t["min_waves"], "steps": steps,
                "ms_per_step": elapsed * 1e3 / steps, "kernel_ms_per_step": float(np.mean(relax_ms)),
                "rounds": int(ws.counters().rounds_executed), "converged": bool(st.converged),
                "same_target_bits": bool(np.array_equal(bits, target_bits))}
    finally:
        ws.close()


def cpu_baseline(hg_view, qs, tiles, budget_s):
    """The reference CPU path (dijkstra_oracle built from the reference's own
    sources, oracle/_ref) on a bounded sample of the same queries: one core,
    and all cores with one query per worker thread.  Falls back to the
    oracle's C restatement (one core) when _ref is not built."""
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
    result = dict(value=relax_total / seconds, unit="edge relaxations/s", cores=1, kind=kind,
                  sample=f"{n} queries of the headline workload spread over box sizes "
                         f"({relax_total} relaxations in {seconds:.2f} s of Dijkstra time; "
                         "induced-subgraph construction excluded)")
    if kind == "reference":
        # The same queries, one per worker thread (waves of `threads`).
        nproc = len(os.sched_getaffinity(0))
        threads = max(1, min(16, nproc))
        done = picked[:n]
        ns, relax = oracle.ref_csr_sssp_parallel(
            V, hg_view["row_offsets"], hg_view["destinations"], hg_view["weights"],
            hg_view["owner_tiles"], hg_view["tile_count"], [tiles[i] for i in done],
            [qs[i].source for i in done], threads)
        result["all_cores"] = dict(value=relax / (ns * 1e-9), threads=threads, nproc=nproc,
                                   hardware_concurrency=os.cpu_count(),
                                   sample=f"the same {len(done)} queries, one per thread")
    return result


def plain_vs_batched(dev, steps):
    """Config 2: 64 queries one per launch (plain) vs all 64 in one launch.
    Plain mode gives each query its own workspace and runs it as one
    persistent cooperative launch with the same engine knobs as the batched
    run (tile closure, direct boundary pushes): the launch count, not the
    engine, is what differs."""
    hg = bfn.HostGraph.synthetic(*C2["tiles"], C2["seed"])
    dg = bfn.DeviceGraph(dev, hg.layout())
    qs, batch = make_batch(hg, C2)
    # The batched run takes the autotuner's choice for this workload (block
    # Jacobi in one persistent launch when there is none).
    tune_entry = load_tuning(C2["name"])
    tune = tuned_defaults(tune_entry or dict(sweeps=0))
    ws = bfn.Workspace(dev)
    ws.submit(dg, batch, bfn.run_options(engine=tune["engine"], control_mode=tune["control"],
                                         block_size=tune["block"]),
              bfn.tuning(tile_sweeps=tune["sweeps"], count_relaxations=1,
                         items_per_claim=tune["items_per_claim"], use_graph=tune["use_graph"],
                         graph_rounds=tune["graph_rounds"], boundary_mode=tune["boundary_mode"],
                         min_waves=tune["min_waves"], **scheduling(tune)))
    _, per, _ = ws.status()
    relax = [p.relaxations for p in per]
    elapsed, _, _, _ = timed_steps(ws, steps, 1)
    batched = sum(relax) * steps / elapsed
    out = {}
    for eng in (bfn.JACOBI_PULL, bfn.DENSE_CHAOTIC_PUSH, bfn.FRONTIER_PUSH):
        wss = []
        for i in range(len(qs)):
            one = bfn.Batch([[qs[i].source]], [[qs[i].target]], [batch.tiles[i]])
            w = bfn.Workspace(dev)
            w.submit(dg, one, bfn.run_options(engine=eng, control_mode=bfn.PERSISTENT),
                     bfn.tuning(tile_sweeps=0, boundary_mode=1))
            wss.append(w)
        lat = []
        t0 = time.perf_counter()
        for w in wss:
            t = time.perf_counter()
            w.rerun()
            lat.append(time.perf_counter() - t)
        elapsed_plain = time.perf_counter() - t0
        out[bfn.ENGINE_NAMES[eng]] = dict(
            value=sum(relax) / elapsed_plain, p50_ms=percentile(lat, 50) * 1e3,
            p99_ms=percentile(lat, 99) * 1e3, control="persistent_cooperative, one launch per query")
        for w in wss:
            w.close()
    ws.close()
    dg.close()
    best_plain = max(v["value"] for v in out.values())
    return dict(workload=C2["name"], batched_value=batched,
                batched_engine=bfn.ENGINE_NAMES[tune["engine"]],
                batched_control=bfn.CONTROL_NAMES[tune["control"]] + ("+hipgraph" if tune["use_graph"] else ""),
                plain=out, batched_over_best_plain=batched / best_plain)


SERVING = dict(name="c4_serving_64x64x256", tiles=(64, 64, 256), seed=2, query_seed=44,
               rates=(2000, 10000, 50000), max_requests=2000)


def percentile(values, p):
    """Nearest-rank percentile (the reference's engine_shootout.cpp:1538-1567 rule)."""
    v = sorted(values)
    if not v:
        return None
    k = max(1, int(np.ceil(p / 100.0 * len(v))))
    return v[k - 1]


def serving(dev):
    """Poisson arrivals into the continuous-batching server on the 64x64x256
    graph, box mix 70/25/5; latency = response time minus scheduled arrival
    (host monotonic clock), nearest-rank p50/p99."""
    hg = bfn.HostGraph.synthetic(*SERVING["tiles"], SERVING["seed"], host_graph=False)
    dg = bfn.DeviceGraph(dev, hg.layout())
    out = {"workload": SERVING["name"], "mix": "70% side 4-8, 25% 16-32, 5% 64 or unbounded"}
    rng = np.random.default_rng(7)
    for rate in SERVING["rates"]:
        n = min(SERVING["max_requests"], rate)
        qs = hg.synthetic_queries(n, seed=SERVING["query_seed"] + rate, mix=1)
        tiles = [hg.query_tiles(q) for q in qs]
        gaps = rng.exponential(1.0 / rate, n)
        offsets_ns = np.cumsum(gaps) * 1e9
        server = bfn.Server(dev, dg, budget_bytes=8 << 30, max_active_queries=1024,
                            rounds_per_chunk=2, engine=bfn.FRONTIER_PUSH, tile_sweeps=0)
        got = {}
        try:
            t0 = time.monotonic_ns()
            i = 0
            while i < n:
                now = time.monotonic_ns() - t0
                while i < n and offsets_ns[i] <= now:
                    server.submit(i, qs[i].source, qs[i].target, tiles[i])
                    i += 1
                for r in server.poll(capacity=1024, timeout_ms=0):
                    got[r["request_id"]] = r
                if i < n:  # sleep towards the next arrival, at most 0.2 ms
                    wait_ns = offsets_ns[i] - (time.monotonic_ns() - t0)
                    if wait_ns > 0:
                        time.sleep(min(wait_ns, 2e5) / 1e9)
            deadline = time.time() + 120
            while len(got) < n and time.time() < deadline:
                for r in server.poll(capacity=1024, timeout_ms=50):
                    got[r["request_id"]] = r
            t_end = time.monotonic_ns()
        finally:
            server.close()
        lat = [(got[k]["complete_ns"] - (t0 + offsets_ns[k])) / 1e6 for k in got]
        out[str(rate)] = {"requests": n, "completed": len(got),
                          "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                          "throughput_qps": len(got) / ((t_end - t0) / 1e9)}
        log(f"serving {rate}/s:", json.dumps(out[str(rate)]))
    dg.close()
    return out


C5 = dict(name="c5_negotiated_256x256x512_q1024", tiles=(256, 256, 512), seed=5, queries=1024,
          query_seed=55, side=(4, 64), unbounded=0.0, rounds=4)


def negotiated(dev, tune):
    """Negotiated-cost rounds on the 256x256x512 graph: 1024 queries routed
    four times; after each round every edge u->v is re-costed on device to
    base[v] * (1 + 0.5 * occupancy[v]) from the round's compact paths."""
    t = time.perf_counter()
    hg = bfn.HostGraph.synthetic(*C5["tiles"], C5["seed"], host_graph=False)
    log(f"negotiated: graph generated in {time.perf_counter() - t:.1f}s")
    dg = bfn.DeviceGraph(dev, hg.layout())
    base = hg.view()["node_costs"]
    dg.update_node_costs(base)
    qs, batch = make_batch(hg, C5)
    ws = bfn.Workspace(dev)
    opts = bfn.run_options(engine=tune["engine"], control_mode=tune["control"],
                           block_size=tune["block"])
    rounds = []
    for r in range(C5["rounds"]):
        ws.submit(dg, batch, opts, bfn.tuning(tile_sweeps=tune["sweeps"], count_relaxations=1,
                                              items_per_claim=tune["items_per_claim"],
                                              use_graph=tune["use_graph"],
                                              graph_rounds=tune["graph_rounds"],
                                              boundary_mode=tune["boundary_mode"],
                                              min_waves=tune["min_waves"], **scheduling(tune)))
        st, per, _ = ws.status()
        relax_ms = ws.timing()[0]
        relaxations = sum(p.relaxations for p in per)
        t = time.perf_counter()
        paths = ws.paths()
        paths_ms = (time.perf_counter() - t) * 1e3
        t = time.perf_counter()
        dg.update_node_costs_from_workspace(ws)
        update_ms = (time.perf_counter() - t) * 1e3
        rounds.append({"converged": int(st.converged), "relaxations": relaxations,
                       "relax_ms": relax_ms, "relax_per_s": relaxations / (relax_ms * 1e-3),
                       "paths_ms": paths_ms, "cost_update_ms": update_ms,
                       "complete_paths": sum(1 for p in paths if p["reconstruction"] == 0)})
        log(f"negotiated round {r}:", json.dumps(rounds[-1]))
    ws.close()
    dg.close()
    return {"workload": C5["name"], "engine": bfn.ENGINE_NAMES[tune["engine"]],
            "control": bfn.CONTROL_NAMES[tune["control"]] +
            (f"+hipgraph({tune['graph_rounds']} rounds)" if tune["use_graph"] else ""),
            "tuning": "the headline workload's autotuned entry", "rounds": rounds}


def copy_calibration():
    """Kernels of known traffic against which the HBM FETCH_SIZE / WRITE_SIZE
    counters are scaled: three copies of 1 GiB between two device buffers
    (1 GiB read and 1 GiB written per dispatch), then three sums of 1 GiB
    (read only) and three fills of 1 GiB (write only), so that each counter's
    scale is checked on a kernel that moves traffic in one direction."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    s = [a.sum() for _ in range(3)]
    torch.cuda.synchronize()
    for _ in range(3):
        b.fill_(2.0)
    torch.cuda.synchronize()
    del s
    log("copy calibration: 3 x 1073741824 bytes copied, summed and filled")


EXPANSION_DRIVER = os.path.join(ROOT, "bf-new_amd", "lib", "bfnew_expansion_bench")


def expansion(tune):
    """run_batched_expansion (compact paths) on the 64x64x256 graph, 256
    queries starting from their two terminal tiles, one record per expansion
    schedule, for both device runners: one run per planned lane batch and one
    run per planning pass (bf-new_amd/csrc/bench/expansion_bench.cpp)."""
    import subprocess
    out = {}
    for runner in ("batch", "pass"):
        cmd = [EXPANSION_DRIVER, "--tiles", "64", "64", "256", "--queries", "256", "--runner", runner,
               "--engine", str(tune["engine"]), "--control", str(bfn.CHUNKED),
               "--sweeps", str(tune["sweeps"]), "--boundary", str(tune["boundary_mode"]),
               "--graph", str(tune["use_graph"]), "--graph-rounds", "8",
               "--claim", str(tune["items_per_claim"]), "--min-waves", str(tune["min_waves"])]
        log("expansion:", " ".join(cmd[1:]))
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        sys.stderr.write(p.stderr)
        if p.returncode != 0:
            raise SystemExit(f"expansion driver failed ({p.returncode})")
        out[runner] = json.loads(p.stdout.strip().splitlines()[-1])
        log("expansion:", json.dumps(out[runner]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-budget-s", type=float, default=10.0)
    ap.add_argument("--negotiated", action="store_true",
                    help="also run the negotiated-cost rounds on the 256x256x512 graph")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the Poisson-arrival serving measurement")
    ap.add_argument("--expansion", action="store_true",
                    help="also time the expansion orchestrator over the device executor")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--engine", type=int, default=None, help="override the tuned engine")
    ap.add_argument("--claim-order", type=int, default=None, help="override the tuned claim order")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(queries shard across GPUs as independent batches).",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    # The expansion driver is its own process; it runs before this one opens
    # the device.
    expansion_result = expansion(tuned_defaults(load_tuning(C3["name"]))) if args.expansion else None
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
    if args.claim_order is not None:
        tune.update(claim_order=args.claim_order)
    opts = bfn.run_options(engine=tune["engine"], control_mode=tune["control"],
                           block_size=tune["block"])
    ws = bfn.Workspace(dev)
    # Validation run (untimed): work-normalised relaxation count and counters.
    ws.submit(dg, batch, bfn.run_options(engine=tune["engine"], control_mode=tune["control"],
                                         block_size=tune["block"], instrumentation=1),
              bfn.tuning(tile_sweeps=tune["sweeps"], count_relaxations=1,
                         items_per_claim=tune["items_per_claim"], use_graph=tune["use_graph"],
                         graph_rounds=tune["graph_rounds"], boundary_mode=tune["boundary_mode"],
                         min_waves=tune["min_waves"], **scheduling(tune)))
    st, per, _ = ws.status()
    if not st.converged:
        raise SystemExit("batch did not converge")
    relax_per_step = sum(p.relaxations for p in per)
    counters = ws.counters()
    # Timed steps with instrumentation off.
    ws.submit(dg, batch, opts, bfn.tuning(tile_sweeps=tune["sweeps"],
                                          items_per_claim=tune["items_per_claim"],
                                          use_graph=tune["use_graph"],
                                          graph_rounds=tune["graph_rounds"],
                                          boundary_mode=tune["boundary_mode"],
                                          min_waves=tune["min_waves"], **scheduling(tune)))
    elapsed, relax_ms, launches, paths_ms = timed_steps(ws, args.steps, args.warmup, paths=True)
    # one round kernel + one advance kernel per host-polled or replayed round
    round_launches = 1 if tune["control"] == bfn.PERSISTENT and not tune["use_graph"] else launches // 2
    value = relax_per_step * args.steps / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    kernel_ms = float(np.mean(relax_ms))
    achieved = counters.algorithmic_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"{value / 1e9:.2f} G relax/s, {ms_per_step:.2f} ms/step, relax kernel {kernel_ms:.3f} ms, "
        f"compact paths {float(np.mean(paths_ms)):.3f} ms, "
        f"{counters.tile_visits} tile visits, {counters.rounds_executed} rounds")
    target_bits = np.array([p["distance"] for p in ws.paths()], np.float32).view(np.uint32)
    persistent = persistent_control(dev, dg, batch, tune, target_bits, max(3, args.steps // 2))
    log("persistent cooperative control:", json.dumps(persistent))
    traffic, traffic_source = profiled_traffic(C3["name"], tune)
    cycles = [int(c) for c in counters.phase_cycles]
    visit_profile = {
        "tile_visits": counters.tile_visits,
        "closure_sweeps_per_visit": counters.closure_sweeps / max(1, counters.tile_visits),
        "phase_share": {n: c / max(1, sum(cycles)) for n, c in zip(bfn.PHASE_NAMES, cycles)},
        "cycles_per_visit": sum(cycles) / max(1, counters.tile_visits),
        "atomic_attempts": counters.atomic_attempts,
        "atomic_successes": counters.atomic_successes,
        # work per work-normalised relaxation: edge records the visits load,
        # and candidate labels they compute (internal edges from frontier
        # sources per closure sweep, plus boundary candidates)
        "records_per_relaxation": counters.edges_relaxed / max(1, relax_per_step),
        "candidates_per_relaxation": counters.candidate_evaluations / max(1, relax_per_step),
    }
    log("visit profile (instrumented run):", json.dumps(visit_profile))

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
                   "tile_sweeps": tune["sweeps"], "boundary_mode": tune["boundary_mode"],
                   "min_waves": tune["min_waves"], "claim_order": tune["claim_order"],
                   "push_window": tune["push_window"],
                   "control": bfn.CONTROL_NAMES[tune["control"]] +
                   (f"+hipgraph({tune['graph_rounds']} rounds)" if tune["use_graph"] else ""),
                   "relaxations_per_step": relax_per_step,
                   "step": "label initialisation + relaxation to the fixed point + per-query "
                           "finalisation + compact paths of every target (summaries, offsets, "
                           "vertex / label / edge-id arenas) copied to host arrays",
                   "compact_paths_ms_per_step": float(np.mean(paths_ms))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": "relax_round_kernel" if round_launches > 1 else
                     "relax_persistent_kernel",
                     "per": "batch step: every round launch of one relaxation (achieved and "
                            "traffic are bytes per step over kernel time per step)",
                     "kernel_ms_per_step": kernel_ms, "round_launches_per_step": round_launches,
                     "avg_launch_us": kernel_ms * 1e3 / round_launches,
                     "algorithmic_bytes_per_step": counters.algorithmic_bytes,
                     "edges_relaxed_per_step": counters.edges_relaxed},
        "visit_profile": visit_profile,
        "persistent_control": persistent,
    }
    if not args.no_plain:
        pm = plain_vs_batched(dev, max(3, args.steps // 2))
        result["plain_mode"] = pm
        log("plain vs batched (C2):", json.dumps(pm))
    if not args.no_serving:
        sv = serving(dev)
        result["serving"] = sv
        result["p99_query_latency_ms"] = {r: sv[r]["p99_ms"] for r in map(str, SERVING["rates"])}
    if args.negotiated:
        result["negotiated"] = negotiated(dev, tune)
    if expansion_result is not None:
        result["expansion"] = expansion_result
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
    # Every handle is released before the interpreter exits: a workspace left
    # to process teardown still holds its stream and hipGraph executable.
    ws.close()
    dg.close()
    dev.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
