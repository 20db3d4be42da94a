"""Pins the oracle: reference test literals, golden vectors generated from the
reference's own compiled sources, the reference build itself (oracle/_ref),
and independent algorithms (Bellman-Ford vs Dijkstra, brute force)."""
import itertools

import numpy as np
import pytest

import oracle
from helpers import (bits, core_case, f32_from_bits, fbits, fixture_arrays, plain_vertices,
                     spatial_case, tile_mask)

INF = np.float32(np.inf)


def build(case):
    v, e = fixture_arrays(case)
    return oracle.build_graph(v, e, case["config"])


def test_core_fixture_literal_distances():
    # tests/sssp_test.cpp:108-117 (unordered build_weighted_graph, source 0)
    v, e = fixture_arrays(core_case())
    g = oracle.build_graph(v, e, None)
    d, _ = oracle.graph_sssp(g, [0])
    expected = np.array([0.0, 0.5, 0.5, 2.0, 4.0, 4.25, np.inf], np.float32)
    assert np.array_equal(bits(d), bits(expected))
    bf = oracle.bellman_ford(g.vertex_count, g.row_offsets, g.destinations, g.weights,
                             g.owner_tiles, None, [0])
    assert np.array_equal(bits(bf), bits(expected))
    # tests/sssp_test.cpp:122-133: the path to 5 follows the zero edge.
    status, pv, pe = oracle.reconstruct(g.vertex_count, g.column_offsets, g.sources,
                                        g.in_weights, g.in_edge_ids, g.owner_tiles, None, d,
                                        [0], 5)
    assert status == "complete" and list(pv) == [0, 1, 2, 3, 4, 5]


def test_tie_and_zero_cycle_literal_paths():
    v = plain_vertices(4)
    # tests/sssp_test.cpp:174-200: equal-cost tie -> smaller stable EdgeId.
    e = np.zeros(4, oracle.EDGE_DTYPE)
    for i, (s, d_, w, r) in enumerate([(0, 1, 0.5, 200), (0, 2, 0.5, 201), (1, 3, 0.5, 202),
                                       (2, 3, 0.5, 203)]):
        e[i] = (s, d_, w, 1, 1, 0, r)
    g = oracle.build_graph(v, e, None)
    d = oracle.bellman_ford(g.vertex_count, g.row_offsets, g.destinations, g.weights,
                            g.owner_tiles, None, [0])
    st, pv, _ = oracle.reconstruct(g.vertex_count, g.column_offsets, g.sources, g.in_weights,
                                   g.in_edge_ids, g.owner_tiles, None, d, [0], 3)
    assert st == "complete" and list(pv) == [0, 1, 3]
    # tests/sssp_test.cpp:202-227: zero-weight cycle forces backtracking.
    for i, (s, d_, w, r) in enumerate([(0, 1, 0.0, 300), (1, 0, 0.0, 301), (1, 2, 1.0, 302),
                                       (3, 1, 1.0, 303)]):
        e[i] = (s, d_, w, 1, 1, 0, r)
    g = oracle.build_graph(v, e, None)
    d, _ = oracle.graph_sssp(g, [3])
    st, pv, _ = oracle.reconstruct(g.vertex_count, g.column_offsets, g.sources, g.in_weights,
                                   g.in_edge_ids, g.owner_tiles, None, d, [3], 2)
    assert st == "complete" and list(pv) == [3, 1, 2] and d[2] == 2.0


def test_reorder_literals():
    # tests/reorder_test.cpp:96-160
    g = build(spatial_case())
    assert list(g.new_to_old) == [1, 7, 8, 3, 2, 4, 5, 0, 9, 6]
    assert list(g.tile_vertex_offsets) == [0, 4, 6, 8, 10]
    assert g.tile_count == 4 and list(g.tile_has_location) == [1, 1, 1, 0]
    assert list(g.tile_x[:3]) == [0, 1, 2]


def test_engine_fixtures_match_golden(golden):
    for case in golden["engine_cases"]:
        g = build(case)
        exp = case["expected"]
        assert list(g.new_to_old) == exp["new_to_old"], case["name"]
        src = [int(g.old_to_new[s]) for s in case["sources"]]
        sel = tile_mask(g.tile_count, case["selected_tiles"])
        d, relax = oracle.graph_sssp(g, src, sel)
        assert list(bits(d)) == exp["distance_bits"], case["name"]
        assert relax == exp["relaxations"]
        rounds, jd = oracle.jacobi_rounds(g.vertex_count, g.column_offsets, g.sources,
                                          g.in_weights, g.owner_tiles, sel, src)
        assert rounds == exp["jacobi_rounds"], case["name"]
        assert np.array_equal(bits(jd), bits(d))
        if "expected_rounds" in case:
            assert rounds == case["expected_rounds"]
        pv, pe = oracle.predecessors(g.vertex_count, g.column_offsets, g.sources, g.in_weights,
                                     g.in_edge_ids, g.owner_tiles, sel, d, src)
        assert [int(x) for x in pv] == exp["predecessors"], case["name"]
        for t, p in exp["paths"].items():
            st, pvx, pex = oracle.reconstruct(g.vertex_count, g.column_offsets, g.sources,
                                              g.in_weights, g.in_edge_ids, g.owner_tiles, sel,
                                              d, src, int(t))
            assert st == p["status"] and list(pvx) == p["vertices"], (case["name"], t)
            assert list(pex) == p["edge_ids"]
        if "expected_bounded_target" in case:
            tgt = int(g.old_to_new[case["targets"][0]])
            assert bits(d)[tgt] == fbits(case["expected_bounded_target"])
            full, _ = oracle.graph_sssp(g, src)
            assert bits(full)[tgt] == fbits(case["expected_unbounded_target"])


def test_batched_lanes_match_golden(golden):
    bc = golden["batched"]
    g = build(bc)
    literal = {1400: 1, 1401: 2, 1402: 6, 1403: 1}  # batched_jacobi_fixture_suite.hpp:43-48
    for lane in bc["lanes"]:
        src = [int(g.old_to_new[s]) for s in lane["sources"]]
        sel = tile_mask(g.tile_count, lane["selected_tiles"])
        rounds, d = oracle.jacobi_rounds(g.vertex_count, g.column_offsets, g.sources,
                                         g.in_weights, g.owner_tiles, sel, src)
        assert list(bits(d)) == lane["distance_bits"]
        assert rounds == lane["jacobi_rounds"]
        if lane["query_id"] in literal:
            assert rounds == literal[lane["query_id"]]


def test_compact_fixture_literals(golden):
    cc = golden["compact"]
    g = build(cc)
    for q in cc["queries"]:
        src = [int(g.old_to_new[s]) for s in q["sources"]]
        sel = tile_mask(g.tile_count, q["selected_tiles"])
        d, _ = oracle.graph_sssp(g, src, sel)
        assert list(bits(d)) == q["distance_bits"]
        for old_t, lit in q["expected_paths"].items():
            t = int(g.old_to_new[int(old_t)])
            st, pv, _ = oracle.reconstruct(g.vertex_count, g.column_offsets, g.sources,
                                           g.in_weights, g.in_edge_ids, g.owner_tiles, sel, d,
                                           src, t)
            if lit is None:
                assert st == "unreachable"
            else:
                assert st == "complete"
                assert [int(g.new_to_old[x]) for x in pv] == lit["vertices"]
                assert [float(d[x]) for x in pv] == lit["labels"]


def test_maximum_rounds_stops_jacobi(golden):
    # the 33-vertex chain needs 33 rounds; a limit of 2 stops after 2
    case = next(c for c in golden["engine_cases"] if c["name"].startswith("long_chain"))
    g = build(case)
    rounds, d = oracle.jacobi_rounds(g.vertex_count, g.column_offsets, g.sources, g.in_weights,
                                     g.owner_tiles, None, [0], maximum_rounds=2)
    assert rounds == 2 and np.isfinite(d[2]) and not np.isfinite(d[3])


def _random_graph(rng, n, m, zero_fraction=0.0):
    v = np.zeros(n, oracle.VERTEX_DTYPE)
    for i in range(n):
        v[i] = (int(rng.integers(0, 4)), int(rng.integers(0, 3)), 1, int(rng.integers(1, 3)))
    e = np.zeros(m, oracle.EDGE_DTYPE)
    for i in range(m):
        w = 0.0 if rng.random() < zero_fraction else float(np.float32(rng.uniform(0.25, 3.0)))
        e[i] = (int(rng.integers(0, n)), int(rng.integers(0, n)), w, 1, 1, 0, i)
    return v, e


def test_dijkstra_vs_brute_force_paths():
    """Brute force: on tiny graphs the oracle label equals the minimum over
    every simple path of the left-to-right fp32 path sum."""
    rng = np.random.default_rng(7)
    for trial in range(30):
        n = 6
        v, e = _random_graph(rng, n, 12)
        g = oracle.build_graph(v, e, None)
        d, _ = oracle.graph_sssp(g, [0])
        best = np.full(n, np.inf, np.float32)
        best[0] = 0.0
        adj = {}
        for s in range(n):
            for p in range(int(g.row_offsets[s]), int(g.row_offsets[s + 1])):
                adj.setdefault(s, []).append((int(g.destinations[p]), g.weights[p]))

        def walk(u, cost, seen):
            for (x, w) in adj.get(u, []):
                if x in seen:
                    continue
                c = np.float32(cost + w)
                if c < best[x]:
                    best[x] = c
                walk(x, c, seen | {x})
        walk(0, np.float32(0.0), {0})
        assert np.array_equal(bits(d), bits(best)), trial


def test_dijkstra_matches_bellman_ford_random():
    rng = np.random.default_rng(11)
    for trial in range(25):
        v, e = _random_graph(rng, 40, 160, zero_fraction=0.1)
        g = oracle.build_graph(v, e, (0, 0, 2, 2))
        tiles = rng.random(g.tile_count) < 0.7
        srcs = [s for s in range(g.vertex_count) if tiles[g.owner_tiles[s]]][:2]
        if not srcs:
            continue
        sel = tiles.astype(np.uint8)
        d, _ = oracle.graph_sssp(g, srcs, sel)
        bf = oracle.bellman_ford(g.vertex_count, g.row_offsets, g.destinations, g.weights,
                                 g.owner_tiles, sel, srcs)
        _, jd = oracle.jacobi_rounds(g.vertex_count, g.column_offsets, g.sources,
                                     g.in_weights, g.owner_tiles, sel, srcs)
        assert np.array_equal(bits(d), bits(bf)) and np.array_equal(bits(d), bits(jd))


def test_planted_mistake_is_caught():
    """A plausible mistake in the relaxation (dropping the edge weight of one
    edge) must change the golden result, so the golden checks have teeth."""
    v, e = fixture_arrays(core_case())
    e2 = e.copy()
    e2["weight"][5] = 0.0  # 2 -> 3 weight 1.5 dropped
    g, g2 = oracle.build_graph(v, e, None), oracle.build_graph(v, e2, None)
    d, _ = oracle.graph_sssp(g, [0])
    d2, _ = oracle.graph_sssp(g2, [0])
    assert not np.array_equal(bits(d), bits(d2))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
def test_oracle_matches_reference_build_random():
    rng = np.random.default_rng(3)
    for trial in range(20):
        v, e = _random_graph(rng, 30, 120, zero_fraction=0.05)
        cfg = (0, 0, 2, 1)
        g = oracle.build_graph(v, e, cfg)
        ref = oracle.RefGraph(v, e, cfg)
        a = ref.arrays()
        for k in ("row_offsets", "destinations", "edge_ids", "column_offsets", "sources",
                  "in_edge_ids", "new_to_old", "owner_tiles", "tile_vertex_offsets"):
            assert np.array_equal(a[k], getattr(g, k)), (trial, k)
        assert np.array_equal(bits(a["weights"]), bits(g.weights))
        runs = oracle.tile_runs(g)
        rr = ref.runs()
        for k in rr:
            assert np.array_equal(rr[k], runs[k]), (trial, k)
        sel = (rng.random(g.tile_count) < 0.8).astype(np.uint8)
        srcs = [s for s in range(g.vertex_count) if sel[g.owner_tiles[s]]][:1]
        if srcs:
            d, relax = oracle.graph_sssp(g, srcs, sel)
            rd, rrelax, _ = ref.sssp(srcs, sel)
            assert np.array_equal(bits(d), bits(rd)) and relax == rrelax


def _cert_layout(g):
    return dict(tile_count=g.tile_count, tile_vertex_offsets=g.tile_vertex_offsets.astype(np.uint32),
                csc_column_offsets=g.column_offsets.astype(np.uint32), csc_sources=g.sources,
                csc_weights=g.in_weights, owner_tiles=g.owner_tiles)


def test_region_certificate_accepts_dijkstra_and_rejects_planted_mistakes():
    """The size-independent certificate the GPU tests use at the big sizes:
    it accepts the oracle's own Dijkstra labels on the full graph and on a
    tile subset, and rejects each planted mistake: a label one ulp low, one
    ulp high, a nonzero source, a spurious finite label on an unreachable
    cycle of zero or absorbed weights."""
    import bfnew_amd as bfn
    hg = bfn.HostGraph.synthetic(8, 8, 32, 17, host_graph=True)
    lay = hg.layout_arrays()
    view = hg.view()
    T = lay["tile_count"]
    for tiles in (list(range(T)), [9, 10, 11, 17, 18, 19, 25, 26, 27]):
        src = int(lay["tile_vertex_offsets"][tiles[len(tiles) // 2]]) + 3
        sel = tile_mask(T, tiles)
        d, _ = oracle.dijkstra(view["vertex_count"], view["row_offsets"], view["destinations"],
                               view["weights"], view["owner_tiles"], sel, [src])
        comp = d[oracle.region_vertices(lay["tile_vertex_offsets"], tiles)]
        assert oracle.certify_region(lay, tiles, comp, [src])[0] == ""
        finite = np.flatnonzero(np.isfinite(comp) & (comp > 0))
        k = int(finite[len(finite) // 2])
        for delta in (-1, 1):
            bad = comp.copy()
            bad.view(np.uint32)[k] = np.uint32(int(bad.view(np.uint32)[k]) + delta)
            assert oracle.certify_region(lay, tiles, bad, [src])[0] != "", delta
        bad = comp.copy()
        ks = int(np.flatnonzero(comp == 0)[0])
        bad[ks] = np.float32(1.0)
        assert "source" in oracle.certify_region(lay, tiles, bad, [src])[0]
    # An unreachable two-cycle with zero weights, then with weights absorbed
    # by the label (fl(2^25 + 1) == 2^25): equal finite labels on the cycle
    # must not certify.
    v = plain_vertices(4)
    for w in (0.0, 1.0):
        e = np.zeros(3, oracle.EDGE_DTYPE)
        for i, (s, d_, ww, r) in enumerate([(0, 1, 1.0, 1), (2, 3, w, 2), (3, 2, w, 3)]):
            e[i] = (s, d_, ww, 1, 1, 0, r)
        g = oracle.build_graph(v, e, None)
        layout = _cert_layout(g)
        src = [int(g.old_to_new[0])]
        d, _ = oracle.graph_sssp(g, src)
        assert oracle.certify_region(layout, [0], d, src)[0] == ""
        bad = d.copy()
        bad[int(g.old_to_new[2])] = bad[int(g.old_to_new[3])] = np.float32(2.0 ** 25)
        assert "tight" in oracle.certify_region(layout, [0], bad, src)[0], w
