"""Negotiated-cost rounds: route, count path occupancy, re-cost, route again.

Every edge u->v weighs cost[v] = base[v] * (1 + 0.5 * occupancy[v]), with the
occupancy of the previous round's complete paths.  Each round is checked bit
for bit against the oracle run on weights recomputed on the host, and the
device occupancy against the count over the returned paths.
"""
import numpy as np
import pytest

import bfnew_amd as bfn
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", [bfn.FRONTIER_PUSH, bfn.JACOBI_PULL])
def test_negotiated_rounds_bit_exact(device, engine):
    hg = bfn.HostGraph.synthetic(32, 32, 128, 9, host_graph=True)
    view = hg.view()
    V, T = view["vertex_count"], view["tile_count"]
    dg = bfn.DeviceGraph(device, hg.layout())
    base = view["node_costs"].astype(np.float32)
    qs = hg.synthetic_queries(24, seed=31, side_min=4, side_max=16)
    tiles = [hg.query_tiles(q) for q in qs]
    batch = bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], tiles)
    col_vertex = np.repeat(np.arange(V, dtype=np.uint32),
                           np.diff(view["column_offsets"]).astype(np.int64))
    occ = np.zeros(V, np.int32)
    dg.update_node_costs(base)  # occupancy 0: the generator's own weights
    ws = bfn.Workspace(device)
    for rnd in range(3):
        print(f"negotiated round {rnd} engine {engine}", flush=True)
        ws.submit(dg, batch, bfn.run_options(engine=engine), bfn.tuning(tile_sweeps=0))
        st, _, _ = ws.status()
        assert st.converged == 1
        cost = base * (np.float32(1.0) + np.float32(0.5) * occ.astype(np.float32))
        w_csr = cost[view["destinations"]]
        w_csc = cost[col_vertex]
        paths = ws.paths()
        host_occ = np.zeros(V, np.int32)
        for i, q in enumerate(qs):
            sel = np.zeros(T, np.uint8)
            sel[tiles[i]] = 1
            ref, _ = oracle.dijkstra(V, view["row_offsets"], view["destinations"], w_csr,
                                     view["owner_tiles"], sel, [q.source])
            assert np.array_equal(ws.distances(i).view(np.uint32), ref.view(np.uint32)), (rnd, i)
            status, ov, oe = oracle.reconstruct(V, view["column_offsets"], view["sources"], w_csc,
                                                view["in_edge_ids"], view["owner_tiles"], sel, ref,
                                                [q.source], q.target)
            p = paths[i]
            if status == "complete":
                assert list(p["vertices"]) == list(ov) and list(p["edge_ids"]) == list(oe)
                host_occ[np.asarray(p["vertices"], np.int64)] += 1
        assert np.array_equal(ws.occupancy(), host_occ), rnd
        occ = host_occ
        dg.update_node_costs_from_workspace(ws)
    ws.close()
    dg.close()
