"""The product's host graph build (part 1 of the path, host side) against the
oracle, the reference's own builder (oracle/_ref) and the reference literals."""
import numpy as np
import pytest

import bfnew_amd as bfn
import oracle
from helpers import bits, fixture_arrays, spatial_case


def product_build(case, config="fixture"):
    v, e = fixture_arrays(case)
    cfg = case["config"] if config == "fixture" else config
    return bfn.HostGraph.build(v, e, cfg)


KEYS = ("row_offsets", "destinations", "edge_ids", "column_offsets", "sources", "in_edge_ids",
        "new_to_old", "old_to_new", "owner_tiles", "tile_vertex_offsets")


def assert_same_graph(view, g, label):
    for k in KEYS:
        assert np.array_equal(view[k].astype(np.uint64), getattr(g, k).astype(np.uint64)), (label, k)
    assert np.array_equal(bits(view["weights"]), bits(g.weights)), label
    assert np.array_equal(bits(view["in_weights"]), bits(g.in_weights)), label


def test_reorder_literals_product():
    # tests/reorder_test.cpp:96-160
    view = product_build(spatial_case()).view()
    assert list(view["new_to_old"]) == [1, 7, 8, 3, 2, 4, 5, 0, 9, 6]
    assert list(view["tile_vertex_offsets"]) == [0, 4, 6, 8, 10]
    assert list(view["tile_has_location"]) == [1, 1, 1, 0]


def test_fixture_builds_match_oracle(golden):
    cases = golden["engine_cases"] + [golden["batched"], golden["compact"]]
    for case in cases:
        v, e = fixture_arrays(case)
        g = oracle.build_graph(v, e, case["config"])
        view = product_build(case).view()
        assert_same_graph(view, g, case["name"])
        runs = oracle.tile_runs(g)
        lay = product_build(case).layout_arrays()
        assert np.array_equal(lay["csr_run_edge_offsets"], runs["csr_run_edge_offsets"])
        assert np.array_equal(lay["csc_run_source_tiles"], runs["csc_run_source_tiles"])
        assert np.array_equal(lay["csc_column_run_offsets"], runs["csc_column_run_offsets"])
        assert np.array_equal(lay["csr_run_destination_tiles"], runs["csr_run_destination_tiles"])


def test_unordered_build_matches_oracle(golden):
    for case in golden["engine_cases"][:4]:
        v, e = fixture_arrays(case)
        g = oracle.build_graph(v, e, None)
        view = bfn.HostGraph.build(v, e, None).view()
        for k in ("row_offsets", "destinations", "edge_ids", "column_offsets", "sources",
                  "in_edge_ids"):
            assert np.array_equal(view[k].astype(np.uint64), getattr(g, k).astype(np.uint64))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
def test_fingerprints_match_reference(golden):
    for case in golden["engine_cases"] + [golden["compact"]]:
        v, e = fixture_arrays(case)
        ref = oracle.RefGraph(v, e, case["config"])
        assert product_build(case).fingerprint() == ref.fingerprint(), case["name"]


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
def test_synthetic_c1_matches_reference_builder():
    g = bfn.HostGraph.synthetic(16, 16, 64, 1, host_graph=True)
    v, e = g.synthetic_input()
    ref = oracle.RefGraph(v, e, (0, 0, 1, 1))
    assert g.fingerprint() == ref.fingerprint()
    a = ref.arrays()
    view = g.view()
    for k in ("column_offsets", "sources", "in_edge_ids", "row_offsets", "destinations"):
        assert np.array_equal(a[k].astype(np.uint64), view[k].astype(np.uint64)), k


def test_synthetic_direct_equals_generic_build():
    for dims in ((16, 16, 64), (8, 12, 256), (5, 3, 7)):
        g = bfn.HostGraph.synthetic(*dims, seed=5)
        direct = g.layout_arrays()
        generic = g.generic_rebuild().layout_arrays()
        for k, a in direct.items():
            b = generic[k]
            if isinstance(a, np.ndarray):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (dims, k)
            else:
                assert a == b, (dims, k)
        assert g.fingerprint() == g.generic_rebuild().fingerprint()


def test_synthetic_config_shape():
    g = bfn.HostGraph.synthetic(16, 16, 64, 1, host_graph=True)
    view = g.view()
    assert view["vertex_count"] == 16384 and view["edge_count"] == 163840
    assert view["tile_count"] == 257  # 256 tiles + empty spill tile
    costs = view["node_costs"]
    assert costs.min() >= 1.0 and costs.max() <= 4.0
    # every edge into v weighs cost[v]
    dst = view["destinations"]
    assert np.array_equal(bits(view["weights"]), bits(costs[dst]))
    # 4 intra + 6 inter edges per node
    assert np.all(np.diff(view["row_offsets"]) == 10)
    owner = view["owner_tiles"]
    src = np.repeat(np.arange(view["vertex_count"]), 10)
    intra = (owner[src] == owner[dst]).reshape(-1, 10).sum(axis=1)
    assert np.all(intra >= 4)


def test_synthetic_queries_are_in_box():
    g = bfn.HostGraph.synthetic(64, 64, 256, 2)
    qs = g.synthetic_queries(200, seed=9, side_min=4, side_max=32)
    for q in qs:
        assert 4 <= q.box_side <= 32
        tiles = set(int(t) for t in g.query_tiles(q))
        assert len(tiles) == q.box_side ** 2
        assert q.source // 256 in tiles and q.target // 256 in tiles and q.source != q.target
    serving = g.synthetic_queries(400, seed=3, mix=1)
    sides = np.array([q.box_side for q in serving])
    assert 0.6 < np.mean(sides <= 8) < 0.8


def test_route_query_tiles_match_fixture(golden):
    for case in golden["engine_cases"]:
        if case["selection"] != "rect":
            continue
        hg = product_build(case)
        view = hg.view()
        terms = [int(view["old_to_new"][t]) for t in case["sources"] + case["targets"]]
        assert list(hg.route_tiles(terms)) == case["selected_tiles"]


def test_invalid_input_is_rejected():
    v = np.zeros(2, bfn.VERTEX_DTYPE)
    e = np.zeros(1, bfn.EDGE_DTYPE)
    e[0] = (0, 1, -1.0, 1, 1, 0, 0)
    with pytest.raises(bfn.BfnError) as err:
        bfn.HostGraph.build(v, e, (0, 0, 1, 1))
    assert err.value.code == 1 and "nonnegative" in str(err.value)
    e[0] = (0, 5, 1.0, 1, 1, 0, 0)
    with pytest.raises(bfn.BfnError):
        bfn.HostGraph.build(v, e, (0, 0, 1, 1))
