"""Serving loop (bfn_server_*): requests admitted into a running batch.

The reference has no serving loop; the bar is the engine's own: a request's
result must not depend on what else is in flight.  Each response is compared
with the same query run alone through bfn_batch_submit and with the oracle.
"""
import time

import numpy as np
import pytest

import bfnew_amd as bfn
from helpers import oracle_query

pytestmark = pytest.mark.gpu

SERVING_MIX = 1


def _graph(device):
    hg = bfn.HostGraph.synthetic(32, 32, 128, 4, host_graph=True)
    return hg, bfn.DeviceGraph(device, hg.layout())


def _drain(server, want, deadline_s=60.0):
    got = {}
    end = time.time() + deadline_s
    while len(got) < want and time.time() < end:
        for r in server.poll(capacity=256, timeout_ms=200):
            got[r["request_id"]] = r
    return got


@pytest.mark.parametrize("engine", [bfn.FRONTIER_PUSH, bfn.JACOBI_PULL])
def test_server_results_equal_alone_and_oracle(device, engine):
    hg, dg = _graph(device)
    view = hg.view()
    qs = hg.synthetic_queries(40, seed=21, side_min=4, side_max=32, mix=SERVING_MIX)
    tiles = [hg.query_tiles(q) for q in qs]
    print(f"server test engine={engine} queries={len(qs)}", flush=True)
    server = bfn.Server(device, dg, budget_bytes=256 << 20, max_active_queries=16,
                        rounds_per_chunk=2, engine=engine, tile_sweeps=0)
    try:
        # Staggered arrivals: requests join while others are half way.
        for i, q in enumerate(qs):
            server.submit(i, q.source, q.target, tiles[i])
            if i % 5 == 4:
                time.sleep(0.002)
        got = _drain(server, len(qs))
        stats = server.stats()
    finally:
        server.close()
    assert len(got) == len(qs), f"{len(got)} of {len(qs)} responses"
    assert stats["completed"] == len(qs)
    ws = bfn.Workspace(device)
    for i, q in enumerate(qs):
        ref, _ = oracle_query(view, [q.source], tiles[i])
        r = got[i]
        expect = np.float32(ref[q.target])
        assert np.float32(r["distance"]).view(np.uint32) == expect.view(np.uint32), i
        assert r["reached"] == int(np.isfinite(expect))
        assert r["complete_ns"] >= r["submit_ns"]
        if i < 6:  # the same request alone through the batch entry point
            alone = bfn.Batch([[q.source]], [[q.target]], [tiles[i]])
            ws.submit(dg, alone, bfn.run_options(engine=engine), bfn.tuning(tile_sweeps=0))
            p = ws.paths()[0]
            assert np.float32(p["distance"]).view(np.uint32) == expect.view(np.uint32)
            if np.isfinite(expect):
                assert r["path_length"] == p["path_length"], i
    ws.close()


def test_server_rejects_oversized_and_invalid(device):
    hg, dg = _graph(device)
    server = bfn.Server(device, dg, budget_bytes=2 << 20, max_active_queries=4)
    try:
        with pytest.raises(bfn.BfnError):
            server.submit(1, 0, 5, np.arange(hg.view()["tile_count"], dtype=np.uint32))
        with pytest.raises(bfn.BfnError):
            server.submit(2, 0, 5, np.array([3, 2], np.uint32))  # not sorted
    finally:
        server.close()
