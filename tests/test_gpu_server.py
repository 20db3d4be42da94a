"""Serving loop (bfn_server_*): requests admitted into a running batch.

The reference has no serving loop; the bar is the engine's own: a request's
result must not depend on what else is in flight.  Each response -- distance
and compact path -- is compared with the same query run alone through
bfn_batch_submit and with the oracle, on a small graph with staggered
arrivals and on the 64x64x256 serving configuration with Poisson arrivals at
2k and 10k requests/s and the 70/25/5 small/medium/large box mix.
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


def _drain(server, want, deadline_s=60.0, got=None):
    got = {} if got is None else got
    end = time.time() + deadline_s
    while len(got) < want and time.time() < end:
        for r in server.poll_paths(capacity=256, timeout_ms=200):
            got[r["request_id"]] = r
    return got


def _check_against_alone_and_oracle(device, dg, view, engine, qs, tiles, got, alone_every=1):
    """Every response equals the oracle's distance; every alone_every-th
    response equals the same request run alone (distance and compact path)."""
    ws = bfn.Workspace(device)
    for i, q in enumerate(qs):
        ref, _ = oracle_query(view, [q.source], tiles[i])
        r = got[i]
        expect = np.float32(ref[q.target])
        assert np.float32(r["distance"]).view(np.uint32) == expect.view(np.uint32), i
        assert r["reached"] == int(np.isfinite(expect)), i
        assert r["complete_ns"] >= r["submit_ns"]
        if np.isfinite(expect):
            assert len(r["vertices"]) == r["path_length"] + 1, i
            assert r["vertices"][0] == q.source and r["vertices"][-1] == q.target, i
        else:
            assert len(r["vertices"]) == 0, i
        if i % alone_every == 0:
            alone = bfn.Batch([[q.source]], [[q.target]], [tiles[i]])
            ws.submit(dg, alone, bfn.run_options(engine=engine), bfn.tuning(tile_sweeps=0))
            p = ws.paths()[0]
            assert np.float32(p["distance"]).view(np.uint32) == expect.view(np.uint32), i
            if np.isfinite(expect):
                assert r["path_length"] == p["path_length"], i
                assert list(r["vertices"]) == list(p["vertices"]), i
                assert list(r["edge_ids"]) == list(p["edge_ids"]), i
    ws.close()


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
    _check_against_alone_and_oracle(device, dg, view, engine, qs, tiles, got)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("rate", [2000, 10000])
def test_c4_serving_at_size(device, rate):
    """The serving configuration: the 64x64x256 graph resident, requests
    arriving as a Poisson process at `rate` per second with the 70/25/5 box
    mix (sides 4-8, 16-32, 64 or unbounded); every response equals the
    oracle, and every response equals its request run alone."""
    hg = bfn.HostGraph.synthetic(64, 64, 256, 2, host_graph=True)
    view = hg.view()
    dg = bfn.DeviceGraph(device, hg.layout())
    n = 300
    qs = hg.synthetic_queries(n, seed=1000 + rate, mix=SERVING_MIX)
    tiles = [hg.query_tiles(q) for q in qs]
    classes = [len(t) for t in tiles]
    assert max(classes) >= 64 * 64 and min(classes) <= 64, "the mix spans small and large boxes"
    offsets = np.cumsum(np.random.default_rng(rate).exponential(1.0 / rate, n))
    print(f"c4 serving rate={rate} requests={n}", flush=True)
    server = bfn.Server(device, dg, budget_bytes=4 << 30, max_active_queries=512,
                        rounds_per_chunk=2, engine=bfn.FRONTIER_PUSH, tile_sweeps=0)
    got = {}
    try:
        t0 = time.monotonic()
        i = 0
        while i < n:
            while i < n and offsets[i] <= time.monotonic() - t0:
                server.submit(i, qs[i].source, qs[i].target, tiles[i])
                i += 1
            for r in server.poll_paths(capacity=512, timeout_ms=0):
                got[r["request_id"]] = r
            time.sleep(0.0002)
        _drain(server, n, deadline_s=120.0, got=got)
    finally:
        server.close()
    assert len(got) == n, f"{len(got)} of {n} responses"
    print(f"c4 serving rate={rate}: checking {n} responses", flush=True)
    _check_against_alone_and_oracle(device, dg, view, bfn.FRONTIER_PUSH, qs, tiles, got)
    dg.close()


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
