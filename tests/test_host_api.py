"""Host-only parts of the reference-named C++ API (tests/cpp/host_api_test.cpp).

- tiles: the tile directory of the reference's spatial fixture against the
  literals of its tests/tile_test.cpp:80-262 (counts, classification,
  neighbours, long and spill edges, halos, diagonal neighbours,
  determinism), and validate_tile_directory on planted mistakes.
- paths: reconstruct_path_from_distances, validate_reconstructed_path and the
  compact-path host API (reconstruct_compact_path_payload, ...query_paths,
  make_failed_compact_query_result, extract_host_compact_paths) on the
  literal fixtures of the reference's sssp_test.cpp:108-227 and
  compact_paths_test.cpp:114-205 (zero cycle, EdgeId tie, bounded region).
- workspace: the batch_workspace byte model (estimate_batch_workspace, the
  2 x 2 strategy comparison, reservations, decisions and their record) on the
  reference's workspace fixture, against the literal totals of its
  tests/batch_workspace_test.cpp:275-960 (dense widths 8/16/32, the sparse
  compact union 960/720/948/708 bytes, cold 24/28/76 preparation bytes,
  warmed 152/184/184/216, budget boundaries, overflow and planted decision
  mistakes).
- random graphs with zero-weight edges (so the search backtracks): the
  library's reconstruction equals the oracle's path by path, from oracle
  labels.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bf-new_amd", "lib", "bfnew_host_api_test")


def _run(*args, timeout=120):
    if not os.path.exists(DRIVER):
        pytest.fail("bf-new_amd/lib/bfnew_host_api_test is not built (run __graft_entry__.build())")
    p = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failure(s)" in p.stdout, p.stdout + p.stderr
    return p


def test_tile_directory_reference_literals():
    _run("tiles")


def test_path_reconstruction_reference_literals():
    _run("paths")


def test_batch_workspace_reference_literals():
    _run("workspace")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_path_reconstruction_equals_oracle(tmp_path, seed):
    rng = np.random.default_rng(seed)
    V, E = 80, 320
    v = np.zeros(V, oracle.VERTEX_DTYPE)
    for i in range(V):
        v[i] = (i % 9, i // 9, 1, 1)
    e = np.zeros(E, oracle.EDGE_DTYPE)
    for i in range(E):
        s, d = int(rng.integers(V)), int(rng.integers(V))
        w = float(rng.choice([0.0, 0.0, 0.5, 1.0, 1.25]))
        e[i] = (s, d, w, 1, 1, 0, 5000 + i)
    g = oracle.build_graph(v, e, None)
    sources_old = sorted({int(x) for x in rng.choice(V, 2, replace=False)})
    src_new = sorted(int(g.old_to_new[s]) for s in sources_old)
    dist, _ = oracle.graph_sssp(g, src_new)
    targets = list(range(V))
    blob = bytearray(struct.pack("<II", V, E))
    for x in v:
        blob += struct.pack("<iiII", int(x["x"]), int(x["y"]), int(x["has_location"]),
                            int(x["resource_class"]))
    for x in e:
        blob += struct.pack("<IIfQ", int(x["source"]), int(x["destination"]), float(x["weight"]),
                            int(x["provenance_record"]))
    blob += struct.pack("<I", len(sources_old)) + struct.pack(f"<{len(sources_old)}I", *sources_old)
    blob += np.asarray(dist, np.float32).tobytes()
    blob += struct.pack("<I", len(targets)) + struct.pack(f"<{len(targets)}I", *targets)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    inp.write_bytes(bytes(blob))
    _run("paths-file", str(inp), str(out))
    data = out.read_bytes()
    pos = 0
    complete = 0
    for t in targets:
        (status,) = struct.unpack_from("<I", data, pos)
        pos += 4
        ost, ov, oe = oracle.reconstruct(g.vertex_count, g.column_offsets, g.sources, g.in_weights,
                                         g.in_edge_ids, g.owner_tiles, None, dist, src_new, t)
        if status == 1:
            assert ost == "unreachable", t
            continue
        (n,) = struct.unpack_from("<I", data, pos)
        pos += 4
        verts = list(struct.unpack_from(f"<{n}I", data, pos))
        pos += 4 * n
        edges = list(struct.unpack_from(f"<{n - 1}Q", data, pos))
        pos += 8 * (n - 1)
        (cost,) = struct.unpack_from("<f", data, pos)
        pos += 4
        assert ost == "complete", t
        assert verts == [int(x) for x in ov] and edges == [int(x) for x in oe], t
        assert np.float32(cost) == dist[t], t
        complete += 1
    assert pos == len(data) and complete > V // 4
