"""GPU parity: every engine and control mode through the C ABI against the
oracle and the golden vectors (bit-exact distances, reference round counts,
predecessors and compact paths)."""
import itertools

import numpy as np
import pytest

import bfnew_amd as bfn
import oracle
from helpers import (bits, certify_batch, csr_view, fbits, fixture_arrays,
                     oracle_query, oracle_region, stratified, tile_mask)

pytestmark = pytest.mark.gpu

ENGINES = (bfn.JACOBI_PULL, bfn.DENSE_CHAOTIC_PUSH, bfn.FRONTIER_PUSH)
CONTROLS = (bfn.PERSISTENT, bfn.CHUNKED, bfn.PER_ROUND)


def configurations():
    """(engine, control, knobs): Jacobi in point and block mode, the push
    engines with LDS boundary proposals and with direct atomics."""
    for eng, ctl in itertools.product(ENGINES, CONTROLS):
        if eng == bfn.JACOBI_PULL:
            for sweeps in (1, 0):
                yield eng, ctl, {"sweeps": sweeps}
        else:
            for mode in (0, 1):
                yield eng, ctl, {"boundary_mode": mode}


def fixture_graph(device, case):
    v, e = fixture_arrays(case)
    hg = bfn.HostGraph.build(v, e, case["config"])
    return hg, bfn.DeviceGraph(device, hg.layout())


def run(device, dg, batch, engine, control, sweeps=1, max_rounds=1 << 30, block=256, **kw):
    ws = bfn.Workspace(device)
    ws.submit(dg, batch, bfn.run_options(engine=engine, control_mode=control, rounds_per_chunk=3,
                                         maximum_rounds=max_rounds, block_size=block),
              bfn.tuning(tile_sweeps=sweeps, **kw))
    return ws


def test_fixture_suite_all_engines(device, golden):
    for case in golden["engine_cases"]:
        hg, dg = fixture_graph(device, case)
        view = hg.view()
        exp = case["expected"]
        src = [int(view["old_to_new"][s]) for s in case["sources"]]
        tgt = [int(view["old_to_new"][t]) for t in case["targets"]]
        batch = bfn.Batch(sources=[src], targets=[tgt], tiles=[case["selected_tiles"]])
        for eng, ctl, knobs in configurations():
            label = (case["name"], eng, ctl, knobs)
            ws = run(device, dg, batch, eng, ctl, **knobs)
            st, per, _ = ws.status()
            assert st.converged == 1 and per[0].converged == 1, label
            d = ws.distances(0)
            assert list(bits(d)) == exp["distance_bits"], label
            if eng == bfn.JACOBI_PULL and knobs["sweeps"] == 1:
                # point Jacobi keeps the reference's round count
                assert per[0].rounds_completed == exp["jacobi_rounds"], label
                assert st.rounds_completed == exp["jacobi_rounds"], label
            assert per[0].reached_all_targets == exp["reached"], label
            assert per[0].bounding_box_miss == 1 - exp["reached"], label
            if ctl == bfn.PER_ROUND:
                assert list(ws.predecessors(0)) == exp["predecessors"], label
                paths = ws.paths()
                for p in paths:
                    g = exp["paths"][str(p["target"])]
                    if g["status"] == "complete":
                        assert p["reconstruction"] == 0 and p["reached"] == 1, label
                        assert list(p["vertices"]) == g["vertices"], label
                        assert list(p["edge_ids"]) == g["edge_ids"], label
                        assert p["path_length"] == len(g["edge_ids"])
                        assert list(bits(p["labels"])) == [exp["distance_bits"][x] for x in g["vertices"]]
                    else:
                        assert p["reconstruction"] == 1 and p["reached"] == 0, label
            ws.close()
        dg.close()


def test_reference_round_literals(device, golden):
    # tests/jacobi_fixture_suite.hpp:195-355 literal round counts
    for case in golden["engine_cases"]:
        if "expected_rounds" not in case:
            continue
        hg, dg = fixture_graph(device, case)
        view = hg.view()
        src = [int(view["old_to_new"][s]) for s in case["sources"]]
        tgt = [int(view["old_to_new"][t]) for t in case["targets"]]
        ws = run(device, dg, bfn.Batch([src], [tgt], [case["selected_tiles"]]), bfn.JACOBI_PULL,
                 bfn.PER_ROUND)
        st, per, _ = ws.status()
        assert st.rounds_completed == case["expected_rounds"], case["name"]
        if "expected_bounded_target" in case:
            d = ws.distances(0)
            assert bits(d)[tgt[0]] == fbits(case["expected_bounded_target"])
            assert st.reached_target_mask == case["expected_reached"]
            assert st.bounding_box_miss_mask == case["expected_miss"]


def test_maximum_rounds_exhaustion(device, golden):
    case = next(c for c in golden["engine_cases"] if c["name"].startswith("long_chain"))
    hg, dg = fixture_graph(device, case)
    batch = bfn.Batch([[0]], [[1]], [case["selected_tiles"]])
    for ctl in CONTROLS:
        ws = run(device, dg, batch, bfn.JACOBI_PULL, ctl, max_rounds=2)
        st, per, _ = ws.status()
        assert st.stop_reason == 2 and st.converged == 0 and st.rounds_completed == 2, ctl
        assert st.reached_target_mask == 0 and st.bounding_box_miss_mask == 0
        assert per[0].converged == 0
        paths = ws.paths()
        assert paths[0]["reconstruction"] == 2  # query_terminal_failure


def test_batched_mixed_duration_lanes(device, golden):
    bc = golden["batched"]
    hg, dg = fixture_graph(device, bc)
    view = hg.view()
    lanes = bc["lanes"]
    batch = bfn.Batch(sources=[[int(view["old_to_new"][s]) for s in l["sources"]] for l in lanes],
                      targets=[[int(view["old_to_new"][t]) for t in l["targets"]] for l in lanes],
                      tiles=[l["selected_tiles"] for l in lanes])
    for eng, ctl, knobs in configurations():
        ws = run(device, dg, batch, eng, ctl, **knobs)
        st, per, _ = ws.status()
        for q, lane in enumerate(lanes):
            assert list(bits(ws.distances(q))) == lane["distance_bits"], (eng, ctl, q)
            if eng == bfn.JACOBI_PULL and knobs["sweeps"] == 1:
                assert per[q].rounds_completed == lane["jacobi_rounds"], (ctl, q)
            assert per[q].reached_all_targets == lane["reached"]
        if eng == bfn.JACOBI_PULL and knobs["sweeps"] == 1:
            assert st.rounds_completed == max(l["jacobi_rounds"] for l in lanes)
        assert st.valid_lane_mask == 0b11111
        ws.close()


def test_compact_path_fixture(device, golden):
    cc = golden["compact"]
    hg, dg = fixture_graph(device, cc)
    view = hg.view()
    qs = cc["queries"]
    o2n = view["old_to_new"]
    batch = bfn.Batch(sources=[[int(o2n[s]) for s in q["sources"]] for q in qs],
                      targets=[[int(o2n[t]) for t in q["targets"]] for q in qs],
                      tiles=[q["selected_tiles"] for q in qs])
    for eng in ENGINES:
        ws = run(device, dg, batch, eng, bfn.PER_ROUND)
        paths = ws.paths()
        by = {(qi, p["target"]): p for qi, q in enumerate(qs) for p in
              paths[batch.target_offsets[qi]:batch.target_offsets[qi + 1]]}
        for qi, q in enumerate(qs):
            assert list(bits(ws.distances(qi))) == q["distance_bits"]
            for old_t, lit in q["expected_paths"].items():
                p = by[(qi, int(o2n[int(old_t)]))]
                if lit is None:
                    assert p["reconstruction"] == 1 and p["reached"] == 0
                    assert p["has_selected_source"] == 0 and np.isinf(p["distance"])
                    continue
                assert p["reconstruction"] == 0 and p["has_selected_source"] == 1
                assert [int(view["new_to_old"][x]) for x in p["vertices"]] == lit["vertices"]
                assert [float(x) for x in p["labels"]] == lit["labels"]
                assert p["selected_source"] == p["vertices"][0]
                g = q["paths"][str(int(o2n[int(old_t)]))]
                assert list(p["edge_ids"]) == g["edge_ids"]
        ws.close()


def same_path_arrays(a, b):
    return (a["summaries"].tobytes() == b["summaries"].tobytes() and
            all(np.array_equal(a[k], b[k]) for k in ("offsets", "vertices", "labels", "edge_ids")))


def test_compact_paths_repeat_and_rerun(device, golden):
    """The path kernel's marks stay in the workspace between passes (each
    pass marks with its own stamps): repeated path calls on one completed
    batch, and the calls after a rerun, return the same summaries and
    arenas, and the paths still match the reference's literal paths."""
    cc = golden["compact"]
    hg, dg = fixture_graph(device, cc)
    view = hg.view()
    qs = cc["queries"]
    o2n = view["old_to_new"]
    batch = bfn.Batch(sources=[[int(o2n[s]) for s in q["sources"]] for q in qs],
                      targets=[[int(o2n[t]) for t in q["targets"]] for q in qs],
                      tiles=[q["selected_tiles"] for q in qs])
    hs, ds = synthetic(device, 32, 32, 128, 6)
    sq, sbatch = c2_batch(hs, 8, seed=7)
    for graph, b in ((dg, batch), (ds, sbatch)):
        for eng in ENGINES:
            ws = run(device, graph, b, eng, bfn.PER_ROUND, 0)
            first = ws.path_arrays()
            assert same_path_arrays(ws.path_arrays(), first), eng
            for _ in range(2):
                ws.rerun()
                assert same_path_arrays(ws.path_arrays(), first), eng
                assert same_path_arrays(ws.path_arrays(), first), eng
            ws.close()
    ws = run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.PER_ROUND)
    for _ in range(3):
        ws.path_arrays()
    paths = ws.paths()
    for qi, q in enumerate(qs):
        for p in paths[batch.target_offsets[qi]:batch.target_offsets[qi + 1]]:
            lit = q["expected_paths"].get(str(int(view["new_to_old"][p["target"]])))
            if lit is not None:
                assert [int(view["new_to_old"][x]) for x in p["vertices"]] == lit["vertices"]
    ws.close()


def synthetic(device, tx, ty, npt, seed, host_graph=True):
    hg = bfn.HostGraph.synthetic(tx, ty, npt, seed, host_graph=host_graph)
    return hg, bfn.DeviceGraph(device, hg.layout())


def test_c1_full_graph_distances_and_predecessors(device):
    hg, dg = synthetic(device, 16, 16, 64, 1)
    view = hg.view()
    T = view["tile_count"]
    src, tgt = 4321, 9876
    ref, _ = oracle.dijkstra(view["vertex_count"], view["row_offsets"], view["destinations"],
                             view["weights"], view["owner_tiles"], None, [src])
    pv, _ = oracle.predecessors(view["vertex_count"], view["column_offsets"], view["sources"],
                                view["in_weights"], view["in_edge_ids"], view["owner_tiles"], None,
                                ref, [src])
    batch = bfn.Batch([[src]], [[tgt]], [list(range(T))])
    for eng, ctl, knobs in configurations():
        ws = run(device, dg, batch, eng, ctl, **knobs)
        assert np.array_equal(bits(ws.distances(0)), bits(ref)), (eng, ctl, knobs)
        assert np.array_equal(ws.predecessors(0), pv)
        p = ws.paths()[0]
        st, ov, oe = oracle.reconstruct(view["vertex_count"], view["column_offsets"],
                                        view["sources"], view["in_weights"], view["in_edge_ids"],
                                        view["owner_tiles"], None, ref, [src], tgt)
        assert list(p["vertices"]) == list(ov) and list(p["edge_ids"]) == list(oe)
        ws.close()


def c2_batch(hg, count, seed=2):
    qs = hg.synthetic_queries(count, seed=seed, side_min=4, side_max=32)
    tiles = [hg.query_tiles(q) for q in qs]
    return qs, bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs], tiles)


@pytest.mark.timeout(600)
def test_c2_batched_and_plain_bit_exact(device):
    """Config 2 (64x64x256, 64 queries, sides 4-32): the whole batch under
    every engine and boundary/occupancy variant, and plain mode -- each query
    alone in its own persistent launch, for each of the three engines --
    bit-exact against oracle Dijkstra for all 64 queries."""
    hg, dg = synthetic(device, 64, 64, 256, 2)
    view = hg.view()
    qs, batch = c2_batch(hg, 64)
    refs = [oracle_query(view, [q.source], batch.tiles[i])[0] for i, q in enumerate(qs)]
    for eng, knobs in ((0, {"sweeps": 0}), (0, {"sweeps": 1}), (1, {"boundary_mode": 0}),
                       (1, {"boundary_mode": 1}), (2, {"boundary_mode": 0}),
                       (2, {"boundary_mode": 1}), (2, {"boundary_mode": 1, "min_waves": 6}),
                       (1, {"boundary_mode": 1, "min_waves": 8})):
        ws = run(device, dg, batch, eng, bfn.PERSISTENT, **knobs)
        for i in range(len(qs)):
            assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), (eng, knobs, i)
        ws.close()
    # plain mode: one query per launch, one workspace per engine reused by
    # every query (each submit rebinds it)
    for eng in ENGINES:
        ws = bfn.Workspace(device)
        for i in range(len(qs)):
            one = bfn.Batch([[qs[i].source]], [[qs[i].target]], [batch.tiles[i]])
            ws.submit(dg, one, bfn.run_options(engine=eng, control_mode=bfn.PERSISTENT),
                      bfn.tuning(tile_sweeps=0, boundary_mode=1))
            assert ws.status()[0].converged == 1, (eng, i)
            assert np.array_equal(bits(ws.distances(0)), bits(refs[i])), (eng, i)
        ws.close()
    # the reference's point-Jacobi round semantics, host-polled per round
    for i in range(4):
        one = bfn.Batch([[qs[i].source]], [[qs[i].target]], [batch.tiles[i]])
        for eng in ENGINES:
            ws = run(device, dg, one, eng, bfn.PER_ROUND)
            assert np.array_equal(bits(ws.distances(0)), bits(refs[i]))
            ws.close()


def test_mixed_engines_and_rerun(device):
    hg, dg = synthetic(device, 32, 32, 128, 4)
    view = hg.view()
    qs, batch = c2_batch(hg, 12, seed=5)
    batch.engines = [i % 3 for i in range(len(qs))]
    refs = [oracle_query(view, [q.source], batch.tiles[i])[0] for i, q in enumerate(qs)]
    ws = run(device, dg, batch, bfn.JACOBI_PULL, bfn.PERSISTENT, 0)
    first = [ws.distances(i) for i in range(len(qs))]
    for i in range(len(qs)):
        assert np.array_equal(bits(first[i]), bits(refs[i])), i
    for _ in range(3):
        ws.rerun()
        for i in range(len(qs)):
            assert np.array_equal(bits(ws.distances(i)), bits(first[i]))
    _, per, _ = ws.status()
    assert [p.engine for p in per] == batch.engines


def test_hipgraph_replay(device):
    hg, dg = synthetic(device, 32, 32, 128, 6)
    view = hg.view()
    qs, batch = c2_batch(hg, 8, seed=7)
    refs = [oracle_query(view, [q.source], batch.tiles[i])[0] for i, q in enumerate(qs)]
    for eng in ENGINES:
        ws = run(device, dg, batch, eng, bfn.CHUNKED, 0, use_graph=1, graph_rounds=3)
        for rep in range(3):
            for i in range(len(qs)):
                assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), (eng, rep, i)
            ws.rerun()
        ws.close()


def test_slot_order_claims(device):
    """claim_order 1 (queue bits compacted between rounds into slot-order
    segments claimed per XCD group): every engine, per-round and chunked
    control with hipGraph replay, one and three items per claim, bit-exact
    against the oracle on two runs; point Jacobi takes exactly the rounds of
    arrival-order claims (its rounds do not depend on the order of visits);
    a round limit stops it the same way; persistent control refuses it."""
    hg, dg = synthetic(device, 32, 32, 128, 6)
    view = hg.view()
    qs, batch = c2_batch(hg, 12, seed=13)
    refs = [oracle_query(view, [q.source], batch.tiles[i])[0] for i, q in enumerate(qs)]
    for eng in ENGINES:
        for ctl, graph in ((bfn.PER_ROUND, 0), (bfn.CHUNKED, 1)):
            for ipc in (1, 3):
                ws = run(device, dg, batch, eng, ctl, 0, items_per_claim=ipc, use_graph=graph,
                         graph_rounds=4 if graph else 0, boundary_mode=1, claim_order=1)
                for rep in range(2):
                    assert ws.status()[0].converged == 1, (eng, ctl, ipc, rep)
                    for i in range(len(qs)):
                        assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), (eng, ctl, ipc, rep, i)
                    ws.rerun()
                ws.close()
    rounds = []
    for order in (0, 1):
        ws = run(device, dg, batch, bfn.JACOBI_PULL, bfn.PER_ROUND, 1, claim_order=order)
        st, per, _ = ws.status()
        rounds.append((st.rounds_completed, [p.rounds_completed for p in per]))
        ws.close()
    assert rounds[0] == rounds[1]
    ws = run(device, dg, batch, bfn.JACOBI_PULL, bfn.CHUNKED, 1, max_rounds=2, claim_order=1)
    st, _, _ = ws.status()
    assert st.stop_reason == 2 and st.converged == 0 and st.rounds_completed == 2
    ws.close()
    with pytest.raises(bfn.BfnError) as refused:
        run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.PERSISTENT, 0, claim_order=1)
    assert refused.value.code == 1


def test_push_window(device):
    """push_window w > 0 (a round visits only the queued push slots whose
    smallest pushed-in label is within w of the queue's smallest; the others
    stay queued): both push engines, persistent, per-round and chunked
    control with hipGraph replay (arrival and slot order), narrow and wide
    windows, bit-exact against the oracle on two runs; a mixed batch keeps
    its Jacobi queries on reference rounds; a negative or non-finite window
    is refused."""
    hg, dg = synthetic(device, 32, 32, 128, 6)
    view = hg.view()
    qs, batch = c2_batch(hg, 12, seed=17)
    refs = [oracle_query(view, [q.source], batch.tiles[i])[0] for i, q in enumerate(qs)]
    controls = ((bfn.PERSISTENT, 0, 0), (bfn.PER_ROUND, 0, 0), (bfn.CHUNKED, 1, 0), (bfn.CHUNKED, 1, 1))
    for eng in (bfn.DENSE_CHAOTIC_PUSH, bfn.FRONTIER_PUSH):
        for ctl, graph, order in controls:
            for window in (0.25, 4.0):
                ws = run(device, dg, batch, eng, ctl, 0, items_per_claim=2, use_graph=graph,
                         graph_rounds=6 if graph else 0, boundary_mode=1, claim_order=order,
                         push_window=window)
                for rep in range(2):
                    assert ws.status()[0].converged == 1, (eng, ctl, order, window, rep)
                    for i in range(len(qs)):
                        assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), (eng, ctl, order, window, rep, i)
                    ws.rerun()
                ws.close()
    batch.engines = [bfn.JACOBI_PULL if i % 2 else bfn.FRONTIER_PUSH for i in range(len(qs))]
    ws = run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.PERSISTENT, 1, push_window=1.0)
    alone = run(device, dg, bfn.Batch([[qs[1].source]], [[qs[1].target]], [batch.tiles[1]]),
                bfn.JACOBI_PULL, bfn.PERSISTENT, 1)
    assert ws.status()[1][1].rounds_completed == alone.status()[1][0].rounds_completed
    for i in range(len(qs)):
        assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), i
    ws.close()
    alone.close()
    batch.engines = None
    for bad in (-1.0, float("inf"), float("nan")):
        with pytest.raises(bfn.BfnError) as refused:
            run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.PERSISTENT, 0, push_window=bad)
        assert refused.value.code == 1


def test_update_weights_orientations(device):
    """bfn_graph_update_weights: new weights in both orientations (each edge's
    weight doubled) give the oracle's distances on the doubled weights for a
    push and a pull engine; orientations that do not hold the same weights,
    or arrays of the wrong length, are refused and leave the weights as they
    were."""
    hg, dg = synthetic(device, 16, 16, 64, 5)
    view = hg.view()
    lay = hg.layout_arrays()
    qs, batch = c2_batch(hg, 6, seed=21)
    csr_w = (lay["csr_weights"] * np.float32(2)).astype(np.float32)
    csc_w = (lay["csc_weights"] * np.float32(2)).astype(np.float32)
    dg.update_weights(csr_w, csc_w)
    assert dg.info().has_zero_weight == 0
    doubled = dict(view, weights=(view["weights"] * np.float32(2)).astype(np.float32))
    refs = [oracle_query(doubled, [q.source], batch.tiles[i])[0] for i, q in enumerate(qs)]
    for eng in (bfn.FRONTIER_PUSH, bfn.JACOBI_PULL):
        ws = run(device, dg, batch, eng, bfn.PERSISTENT, 0)
        for i in range(len(qs)):
            assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), (eng, i)
        ws.close()
    bad = csc_w.copy()
    bad[0] = np.float32(bad[0] + 1)
    with pytest.raises(bfn.BfnError) as refused:
        dg.update_weights(csr_w, bad)
    assert refused.value.code == 1
    with pytest.raises(ValueError):
        dg.update_weights(csr_w[:-1], csc_w[:-1])
    ws = run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.PERSISTENT, 0)
    for i in range(len(qs)):
        assert np.array_equal(bits(ws.distances(i)), bits(refs[i])), i
    ws.close()


def test_relaxation_counts_match_oracle(device):
    hg, dg = synthetic(device, 32, 32, 128, 8)
    view = hg.view()
    qs, batch = c2_batch(hg, 6, seed=9)
    ws = run(device, dg, batch, bfn.JACOBI_PULL, bfn.PERSISTENT, 0, count_relaxations=1)
    _, per, _ = ws.status()
    for i, q in enumerate(qs):
        _, relax = oracle_query(view, [q.source], batch.tiles[i])
        assert per[i].relaxations == relax


@pytest.mark.timeout(600)
def test_c3_batch_certificate(device):
    """Config 3 at full size (128x128x512, 256 queries, sides 4-64, 1%
    unbounded): every query of three runs -- block Jacobi in one persistent
    cooperative launch, the tuned frontier push replayed through a hipGraph,
    and dense chaotic push -- certified exact by the oracle's region
    certificate, and a stratified sample (smallest, median and largest
    boxes, plus the unbounded queries) bit-exact against oracle Dijkstra."""
    hg, dg = synthetic(device, 128, 128, 512, 3, host_graph=False)
    lay = hg.layout_arrays()
    view = csr_view(lay)
    qs = hg.synthetic_queries(256, seed=33, side_min=4, side_max=64, unbounded_fraction=0.01)
    tiles = [hg.query_tiles(q) for q in qs]
    sources = [[q.source] for q in qs]
    batch = bfn.Batch(sources, [[q.target] for q in qs], tiles)
    runs = {
        "jacobi_persistent": run(device, dg, batch, bfn.JACOBI_PULL, bfn.PERSISTENT, 0),
        "frontier_tuned": run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.CHUNKED, 0, block=256,
                              items_per_claim=2, use_graph=1, graph_rounds=36, boundary_mode=1),
        "dense_persistent": run(device, dg, batch, bfn.DENSE_CHAOTIC_PUSH, bfn.PERSISTENT, 0,
                                boundary_mode=1),
        "frontier_slot_order": run(device, dg, batch, bfn.FRONTIER_PUSH, bfn.CHUNKED, 0, block=256,
                                   items_per_claim=2, use_graph=1, graph_rounds=36, boundary_mode=1,
                                   claim_order=1),
    }
    for name, ws in runs.items():
        assert ws.status()[0].converged == 1, name
        failures = certify_batch(lay, ws, sources, tiles)
        assert not failures, (name, failures[:5])
    sample = stratified([len(t) for t in tiles]) + [i for i, q in enumerate(qs) if q.unbounded]
    for i in sorted(set(sample)):
        ref = oracle_region(view, lay, qs[i].source, tiles[i])
        for name, ws in runs.items():
            assert np.array_equal(bits(ws.region_distances(i)), bits(ref)), (name, i)
    jac = runs["jacobi_persistent"].paths()
    for name in ("frontier_tuned", "dense_persistent", "frontier_slot_order"):
        assert [p["distance"] for p in runs[name].paths()] == [p["distance"] for p in jac], name
    for ws in runs.values():
        ws.close()
    dg.close()


def test_refused_submit_leaves_no_batch(device):
    """A submit refused after planning (here: a tile neighbourhood larger than
    the LDS of one workgroup) leaves the workspace without a batch -- status,
    results and rerun report BFN_E_STATE instead of mixing the previous
    batch's labels with the new batch's description -- and the workspace
    takes the next valid batch as usual."""
    hg, dg = synthetic(device, 8, 8, 64, 4)
    view = hg.view()
    qs = hg.synthetic_queries(3, seed=3, side_min=4, side_max=8)
    batch = bfn.Batch([[q.source] for q in qs], [[q.target] for q in qs],
                      [hg.query_tiles(q) for q in qs])
    ws = run(device, dg, batch, bfn.JACOBI_PULL, bfn.PERSISTENT, 0)
    assert ws.status()[0].converged == 1
    # 8x8 tiles of 4096 vertices: an interior tile reads 13 in-neighbour slabs
    # (itself and +-1, +-2, +-6 in x and y), 13 * 16 KiB > the 160 KiB of LDS.
    big_hg, big = synthetic(device, 8, 8, 4096, 1, host_graph=False)
    bq = big_hg.synthetic_queries(1, seed=1, side_min=8, side_max=8)
    big_batch = bfn.Batch([[bq[0].source]], [[bq[0].target]], [big_hg.query_tiles(bq[0])])
    with pytest.raises(bfn.BfnError) as refused:
        ws.submit(big, big_batch, bfn.run_options(engine=bfn.JACOBI_PULL), bfn.tuning(tile_sweeps=0))
    assert refused.value.code == 4
    for call in (ws.status, ws.rerun, lambda: ws.distances(0), lambda: ws.region_distances(0)):
        with pytest.raises(bfn.BfnError) as e:
            call()
        assert e.value.code == 5
    ws.submit(dg, batch, bfn.run_options(engine=bfn.FRONTIER_PUSH), bfn.tuning(tile_sweeps=0))
    for i, q in enumerate(qs):
        ref, _ = oracle_query(view, [q.source], batch.tiles[i])
        assert np.array_equal(bits(ws.distances(i)), bits(ref)), i
    ws.close()
    big.close()
    dg.close()
