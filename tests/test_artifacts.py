"""BFGRPH07 / BFQRY007 workload artifacts (tests/cpp/artifact_test.cpp).

The reference's reader and writer (src/io/workload_artifact.cpp) cannot be
built here (its query header needs a newer libstdc++), and its tests hold no
artifact files, so the format is pinned by a second codec written here with
`struct` from the layout the reference writes (:230-340): the files the
library writes for a fixture decode field by field to the fixture's values
and consume every byte, the fixture's graph encoded here is byte-identical to
the library's file, and a file encoded here reads back in the library.
The C++ driver checks round trips and every rejection path.
"""
import json
import os
import struct
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bf-new_amd", "lib", "bfnew_artifact_test")


def _driver(*args):
    if not os.path.exists(DRIVER):
        pytest.fail("bf-new_amd/lib/bfnew_artifact_test is not built (run __graft_entry__.build())")
    p = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


class Reader:
    def __init__(self, data):
        self.data, self.at = data, 0

    def take(self, fmt):
        (v,) = struct.unpack_from("<" + fmt, self.data, self.at)
        self.at += struct.calcsize(fmt)
        return v

    def text(self):
        n = self.take("I")
        s = list(self.data[self.at:self.at + n])
        self.at += n
        return s

    def u32s(self):
        return [self.take("I") for _ in range(self.take("I"))]


def encode_graph(vertices, edges):
    out = b"BFGRPH07" + struct.pack("<IIQ", 1, len(vertices), len(edges))
    for x, y, located, rc in vertices:
        out += struct.pack("<iiBH", x, y, located, rc)
    for s, d, wbits, domain, kind, record in edges:
        out += struct.pack("<IIIIIQ", s, d, wbits, domain, kind, record)
    return out


def decode_terminal(r):
    return {"site": r.text(), "pin": r.text(), "original_vertex": r.take("I"), "vertex": r.take("I"),
            "x": r.take("i"), "y": r.take("i"), "has_location": r.take("B"), "owner_tile": r.take("I")}


def test_artifacts_match_an_independent_codec(tmp_path):
    fixture = json.loads(_driver("write", str(tmp_path)).strip().splitlines()[-1])
    graph = open(tmp_path / "graph.bfgrph", "rb").read()
    # The graph file is exactly the fixture encoded here, and decodes back.
    assert graph == encode_graph(fixture["vertices"], fixture["edges"])
    r = Reader(graph)
    assert r.data[:8] == b"BFGRPH07"
    r.at = 8
    assert r.take("I") == 1
    nv, ne = r.take("I"), r.take("Q")
    assert (nv, ne) == (len(fixture["vertices"]), len(fixture["edges"]))
    assert [[r.take("i"), r.take("i"), r.take("B"), r.take("H")] for _ in range(nv)] == fixture["vertices"]
    assert [[r.take(f) for f in "IIIIIQ"] for _ in range(ne)] == fixture["edges"]
    assert r.at == len(graph)

    data = open(tmp_path / "queries.bfqry", "rb").read()
    r = Reader(data)
    assert data[:8] == b"BFQRY007"
    r.at = 8
    assert r.take("I") == 1
    decoded = []
    for _ in range(r.take("Q")):
        q = {"net_name": r.text(), "query_id": r.take("I"), "source_terminals": r.u32s(),
             "target_terminals": r.u32s(), "sources": r.u32s(), "targets": r.u32s(),
             "source_terminal_to_source": r.u32s(), "target_terminal_to_target": r.u32s(),
             "selected_tiles": r.u32s(), "expansion_generation": r.take("I")}
        q["source_descriptors"] = [decode_terminal(r) for _ in range(r.take("I"))]
        q["target_descriptors"] = [decode_terminal(r) for _ in range(r.take("I"))]
        decoded.append(q)
    assert r.at == len(data)
    assert decoded == fixture["queries"]
    # The fixture exercises what the layout must carry: an unlocated vertex,
    # a 16-bit resource class, a zero and a tiny weight, a 64-bit record, a
    # duplicated terminal, an empty and a non-ASCII net name.
    assert any(v[2] == 0 for v in fixture["vertices"]) and any(v[3] > 255 for v in fixture["vertices"])
    assert any(e[2] == 0 for e in fixture["edges"]) and any(e[5] >= 1 << 32 for e in fixture["edges"])
    assert any(len(set(q["source_terminals"])) < len(q["source_terminals"]) for q in fixture["queries"])
    assert any(q["net_name"] == [] for q in fixture["queries"])
    assert any(c > 127 for q in fixture["queries"] for c in q["net_name"])


def test_artifact_round_trips_and_rejections(tmp_path):
    assert "check: 0 failure(s)" in _driver("check", str(tmp_path))


def test_graph_file_encoded_here_reads_back(tmp_path):
    """A graph encoded by the codec here loads in the library and writes back
    byte for byte; a -0 weight reads as +0 (the canonical weight), and a
    corrupted boolean is refused."""
    fixture = json.loads(_driver("write", str(tmp_path)).strip().splitlines()[-1])
    canonical = encode_graph(fixture["vertices"], fixture["edges"])
    edges = [list(e) for e in fixture["edges"]]
    zero = next(i for i, e in enumerate(edges) if e[2] == 0)
    edges[zero][2] = 0x80000000  # -0.0
    src, dst = tmp_path / "here.bfgrph", tmp_path / "back.bfgrph"
    src.write_bytes(encode_graph(fixture["vertices"], edges))
    _driver("recode", str(src), str(dst))
    assert dst.read_bytes() == canonical
    bad = bytearray(canonical)
    bad[24 + 8] = 3
    src.write_bytes(bytes(bad))
    p = subprocess.run([DRIVER, "recode", str(src), str(dst)], capture_output=True, text=True, timeout=60)
    assert p.returncode != 0 and "boolean" in p.stderr
