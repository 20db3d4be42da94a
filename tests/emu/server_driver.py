"""Runs the serving loop (bfn_server_*) of the engine library named by
argv[1] -- the CPU emulation of the kernels -- for the engine argv[2], on a
small synthetic graph with more requests than active places, so requests are
admitted while others are in flight and the rounds advance across chunks.
Every response is checked against the oracle; prints one JSON line."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
sys.path.insert(0, os.path.join(ROOT, "bf-new_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, TESTS)

import numpy as np  # noqa: E402

import bfnew_amd as bfn  # noqa: E402

bfn.HIP_LIB_PATH = sys.argv[1]

from helpers import oracle_query  # noqa: E402


def main():
    engine = int(sys.argv[2])
    dev = bfn.Device(0)
    hg = bfn.HostGraph.synthetic(8, 8, 32, 4, host_graph=True)
    view = hg.view()
    dg = bfn.DeviceGraph(dev, hg.layout())
    qs = hg.synthetic_queries(6, seed=21, side_min=2, side_max=5)
    tiles = [hg.query_tiles(q) for q in qs]
    server = bfn.Server(dev, dg, budget_bytes=64 << 20, max_active_queries=3,
                        rounds_per_chunk=2, engine=engine, tile_sweeps=0)
    got = {}
    try:
        for i, q in enumerate(qs):
            server.submit(i, q.source, q.target, tiles[i])
        end = time.time() + 60
        while len(got) < len(qs) and time.time() < end:
            for r in server.poll_paths(capacity=64, timeout_ms=100):
                got[r["request_id"]] = r
    finally:
        server.close()
    failures = []
    for i, q in enumerate(qs):
        if i not in got:
            failures.append(f"request {i}: no response")
            continue
        ref, _ = oracle_query(view, [q.source], tiles[i])
        r = got[i]
        expect = np.float32(ref[q.target])
        if np.float32(r["distance"]).view(np.uint32) != expect.view(np.uint32):
            failures.append(f"request {i}: distance {r['distance']} != {expect}")
        elif np.isfinite(expect) and (len(r["vertices"]) != r["path_length"] + 1 or
                                      r["vertices"][0] != q.source or
                                      r["vertices"][-1] != q.target):
            failures.append(f"request {i}: path does not run source to target")
    dg.close()
    dev.close()
    print(json.dumps(dict(engine=engine, responses=len(got), failures=failures)), flush=True)


if __name__ == "__main__":
    main()
