"""bfn_batch_run_static: the batch run that allocates nothing (its device
state lives in a static per-GPU array) against bfn_batch_submit on the same
batches, through the C ABI: per-query outcome and every target's distance
bit for bit, for each engine and control mode, on two calls in a row where
the second batch is smaller (the array keeps the first call's contents), and
a batch larger than the array is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import bfnew_amd as bfn
import oracle
from helpers import bits, oracle_query

pytestmark = pytest.mark.gpu


def run_static(dg, batch, options, tune):
    d = batch.desc()
    Q = len(batch.sources)
    nt = int(batch.target_offsets[-1])
    per = (bfn.QueryStatus * Q)()
    out = np.full(max(nt, 1), np.nan, np.float32)
    bfn._check(bfn.hip_lib().bfn_batch_run_static(dg.h, C.byref(d), C.byref(options),
                                                   C.byref(tune), per, bfn._ptr(out)))
    return [per[i] for i in range(Q)], out[:nt]


def make_batch(hg, count, seed, side_max):
    qs = hg.synthetic_queries(count, seed=seed, side_min=4, side_max=side_max)
    return qs, bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs],
                         [hg.query_tiles(q) for q in qs])


def test_static_run_equals_submit(device):
    assert bfn.hip_lib().bfn_static_scratch_bytes() == 64 << 20
    hg = bfn.HostGraph.synthetic(32, 32, 128, 6, host_graph=True)
    view = hg.view()
    dg = bfn.DeviceGraph(device, hg.layout())
    configs = [(bfn.FRONTIER_PUSH, bfn.PERSISTENT), (bfn.DENSE_CHAOTIC_PUSH, bfn.CHUNKED),
               (bfn.JACOBI_PULL, bfn.PER_ROUND), (bfn.JACOBI_PULL, bfn.PERSISTENT)]
    for count, seed, side in ((24, 5, 16), (6, 9, 8)):  # the second call is smaller
        qs, batch = make_batch(hg, count, seed, side)
        for engine, control in configs:
            label = (count, engine, control)
            print("static run", label, flush=True)
            opts = bfn.run_options(engine=engine, control_mode=control, rounds_per_chunk=4)
            tune = bfn.tuning(tile_sweeps=0, boundary_mode=1)
            per, dist = run_static(dg, batch, opts, tune)
            ws = bfn.Workspace(device)
            ws.submit(dg, batch, opts, tune)
            _, wper, _ = ws.status()
            for i, q in enumerate(qs):
                for field in ("converged", "reached_all_targets", "bounding_box_miss",
                              "stop_reason", "error_bits", "engine"):
                    assert getattr(per[i], field) == getattr(wper[i], field), (label, i, field)
                assert per[i].converged == 1, (label, i)
                assert bits(dist[i:i + 1])[0] == bits(ws.distances(i)[q.target:q.target + 1])[0], \
                    (label, i)
            ws.close()
        for i, q in enumerate(qs):  # and the last call against the oracle
            ref, _ = oracle_query(view, [q.source], batch.tiles[i])
            assert bits(dist[i:i + 1])[0] == bits(ref[q.target:q.target + 1])[0], i
    dg.close()


def test_static_run_refuses_a_batch_beyond_its_scratch(device):
    hg = bfn.HostGraph.synthetic(64, 64, 256, 2, host_graph=False)
    dg = bfn.DeviceGraph(device, hg.layout())
    # 64 unbounded queries: 64 x 4097 tiles x 2 x 256 labels x 4 B >> 64 MiB
    tiles = list(range(64 * 64 + 1))
    qs = hg.synthetic_queries(64, seed=3, side_min=4, side_max=8)
    batch = bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], [tiles] * 64)
    with pytest.raises(bfn.BfnError) as e:
        run_static(dg, batch, bfn.run_options(engine=bfn.FRONTIER_PUSH), bfn.tuning())
    assert e.value.code == 3
    # the refusal leaves the entry point usable
    _, small = make_batch(hg, 2, 1, 4)
    per, _ = run_static(dg, small, bfn.run_options(engine=bfn.FRONTIER_PUSH), bfn.tuning())
    assert all(p.converged == 1 for p in per)
    dg.close()
