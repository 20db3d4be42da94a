"""The reference-named C++ API (tests/cpp/api_test.cpp driver).

CPU: run options / controller / run status rules, resident-graph plans,
batch descriptions and the batch planner.  GPU: every batched engine, both control modes, compact
status, compact paths and the three single-query engines on the 16x16x64
graph (seed 1); the driver checks the engines against each other and this
test checks the labels and paths it writes against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
import bfnew_amd as bfn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bf-new_amd", "lib", "bfnew_api_test")


def _run(*args, timeout=120, env=None):
    if not os.path.exists(DRIVER):
        pytest.fail("bf-new_amd/lib/bfnew_api_test is not built (run __graft_entry__.build())")
    p = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    return p


def test_cpp_api_rules_and_layouts():
    out = _run("cpu")
    assert "0 failure(s)" in out.stdout


def test_cpp_batch_planner_golden():
    """Greedy overlapping planner on the reference's four-query fixture
    (tests/batch_plan_test.cpp:80-411): lane order, union masks, estimates,
    the literal serialised plan (tests/golden/batch_plan_golden.tsv, whose
    FNV-1a fingerprint is the reference's 16700197086403570866), damaged-plan
    rejection codes and a random 70-query corpus at widths 8/16/32."""
    golden = os.path.join(ROOT, "tests", "golden", "batch_plan_golden.tsv")
    data = open(golden, "rb").read()
    h = 14695981039346656037
    for c in data:
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    assert h == 16700197086403570866
    out = _run("planner", golden)
    assert "planner: 0 failure(s)" in out.stdout


def _read(path):
    data = np.fromfile(path, np.uint32)
    pos = 0

    def take(n):
        nonlocal pos
        a = data[pos:pos + n]
        pos += n
        return a

    V, W, L = (int(x) for x in take(3))
    lanes = []
    for _ in range(L):
        s, t, n = (int(x) for x in take(3))
        lanes.append((s, t, take(n).copy()))
    dist = take(V * W).view(np.float32).reshape(V, W)
    paths = []
    for _ in range(L):
        status, n = (int(x) for x in take(2))
        verts = take(n).copy()
        edges = take(max(0, n - 1)).copy()
        labels = take(n).view(np.float32).copy()
        paths.append((status, verts, edges, labels))
    assert pos == len(data)
    return V, W, lanes, dist, paths


def check_api_output(path):
    """The lanes, labels and paths the driver's GPU mode wrote, against the
    oracle on the 16x16x64 graph (seed 1)."""
    V, W, lanes, dist, paths = _read(path)
    hg = bfn.HostGraph.synthetic(16, 16, 64, 1, host_graph=True)
    view = hg.view()
    assert view["vertex_count"] == V
    for lane, (s, t, tiles) in enumerate(lanes):
        sel = np.zeros(view["tile_count"], np.uint8)
        sel[tiles] = 1
        ref, _ = oracle.dijkstra(V, view["row_offsets"], view["destinations"], view["weights"],
                                 view["owner_tiles"], sel, [s])
        assert np.array_equal(dist[:, lane].view(np.uint32), ref.view(np.uint32)), lane
        status, verts, edges, labels = paths[lane]
        st, ov, oe = oracle.reconstruct(V, view["column_offsets"], view["sources"],
                                        view["in_weights"], view["in_edge_ids"],
                                        view["owner_tiles"], sel, ref, [s], t)
        if np.isfinite(ref[t]):
            assert status == 0
            assert list(verts) == list(ov) and list(edges) == list(oe)
            assert np.array_equal(labels.view(np.uint32), ref[verts].view(np.uint32))
        else:
            assert status == 1 and len(verts) == 0


@pytest.mark.gpu
def test_cpp_api_engines_match_oracle(tmp_path):
    out = tmp_path / "api.bin"
    _run("gpu", str(out))
    check_api_output(out)
