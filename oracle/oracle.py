"""ctypes wrapper over the C oracle and the reference-built ``_ref`` library.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of bench.py, always as the checker.  The product path
never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbfnew_ref.so")

VERTEX_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("has_location", "<u4"),
                         ("resource_class", "<u4")])
EDGE_DTYPE = np.dtype([("source", "<u4"), ("destination", "<u4"), ("weight", "<f4"),
                       ("provenance_domain", "<u4"), ("provenance_kind", "<u4"),
                       ("pad", "<u4"), ("provenance_record", "<u8")])
NONE32 = 0xFFFFFFFF


class _Graph(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32), ("edge_count", C.c_uint64), ("tile_count", C.c_uint32),
        ("row_offsets", C.POINTER(C.c_uint64)), ("destinations", C.POINTER(C.c_uint32)),
        ("weights", C.POINTER(C.c_float)), ("edge_ids", C.POINTER(C.c_uint64)),
        ("column_offsets", C.POINTER(C.c_uint64)), ("sources", C.POINTER(C.c_uint32)),
        ("in_weights", C.POINTER(C.c_float)), ("in_edge_ids", C.POINTER(C.c_uint64)),
        ("old_to_new", C.POINTER(C.c_uint32)), ("new_to_old", C.POINTER(C.c_uint32)),
        ("owner_tiles", C.POINTER(C.c_uint32)), ("tile_x", C.POINTER(C.c_int64)),
        ("tile_y", C.POINTER(C.c_int64)), ("tile_has_location", C.POINTER(C.c_uint32)),
        ("tile_vertex_offsets", C.POINTER(C.c_uint64)),
    ]


class _Runs(C.Structure):
    _fields_ = [
        ("csr_row_run_offsets", C.POINTER(C.c_uint64)),
        ("csr_run_edge_offsets", C.POINTER(C.c_uint64)),
        ("csr_run_tiles", C.POINTER(C.c_uint32)), ("csr_run_count", C.c_uint64),
        ("csc_column_run_offsets", C.POINTER(C.c_uint64)),
        ("csc_run_edge_offsets", C.POINTER(C.c_uint64)),
        ("csc_run_tiles", C.POINTER(C.c_uint32)), ("csc_run_count", C.c_uint64),
    ]


class _Spatial(C.Structure):
    _fields_ = [("origin_x", C.c_int32), ("origin_y", C.c_int32),
                ("tile_width", C.c_uint32), ("tile_height", C.c_uint32)]


_lib = None


def build() -> None:
    """Compile the oracle (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        _lib.orc_build_graph.argtypes = [P, C.c_uint32, P, C.c_uint64, P, C.POINTER(_Graph)]
        _lib.orc_graph_free.argtypes = [C.POINTER(_Graph)]
        _lib.orc_tile_runs.argtypes = [C.POINTER(_Graph), C.POINTER(_Runs)]
        _lib.orc_tile_runs_free.argtypes = [C.POINTER(_Runs)]
        _lib.orc_dijkstra.argtypes = [C.c_uint32, P, P, P, P, P, P, C.c_uint32, P, P]
        _lib.orc_bellman_ford.argtypes = [C.c_uint32, P, P, P, P, P, P, C.c_uint32, P]
        _lib.orc_jacobi_rounds.argtypes = [C.c_uint32, P, P, P, P, P, P, C.c_uint32,
                                           C.c_uint64, P]
        _lib.orc_jacobi_rounds.restype = C.c_int64
        _lib.orc_predecessors.argtypes = [C.c_uint32, P, P, P, P, P, P, P, P, C.c_uint32, P, P]
        _lib.orc_reconstruct.argtypes = [C.c_uint32, P, P, P, P, P, P, P, P, C.c_uint32,
                                         C.c_uint32, C.c_uint32, P, P, P]
        _lib.orc_certify_region.argtypes = [C.c_uint32, P, P, P, P, P, C.c_uint32, P, P,
                                            C.c_uint32, P, C.POINTER(C.c_uint64)]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


@dataclass
class Graph:
    """A reordered weighted graph (reference numbering)."""
    vertex_count: int
    edge_count: int
    tile_count: int
    row_offsets: np.ndarray
    destinations: np.ndarray
    weights: np.ndarray
    edge_ids: np.ndarray
    column_offsets: np.ndarray
    sources: np.ndarray
    in_weights: np.ndarray
    in_edge_ids: np.ndarray
    old_to_new: np.ndarray
    new_to_old: np.ndarray
    owner_tiles: np.ndarray
    tile_x: np.ndarray
    tile_y: np.ndarray
    tile_has_location: np.ndarray
    tile_vertex_offsets: np.ndarray


def build_graph(vertices: np.ndarray, edges: np.ndarray, config=None) -> Graph:
    vertices = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
    edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
    g = _Graph()
    cfg = None if config is None else C.byref(_Spatial(*config))
    rc = lib().orc_build_graph(_p(vertices), len(vertices), _p(edges), len(edges), cfg,
                               C.byref(g))
    if rc != 0:
        raise ValueError(f"oracle graph build failed: {rc}")
    V, E, T = g.vertex_count, g.edge_count, g.tile_count
    out = Graph(
        V, E, T,
        _copy(g.row_offsets, V + 1, np.uint64), _copy(g.destinations, E, np.uint32),
        _copy(g.weights, E, np.float32), _copy(g.edge_ids, E, np.uint64),
        _copy(g.column_offsets, V + 1, np.uint64), _copy(g.sources, E, np.uint32),
        _copy(g.in_weights, E, np.float32), _copy(g.in_edge_ids, E, np.uint64),
        _copy(g.old_to_new, V, np.uint32), _copy(g.new_to_old, V, np.uint32),
        _copy(g.owner_tiles, V, np.uint32), _copy(g.tile_x, T, np.int64),
        _copy(g.tile_y, T, np.int64), _copy(g.tile_has_location, T, np.uint32),
        _copy(g.tile_vertex_offsets, T + 1, np.uint64))
    lib().orc_graph_free(C.byref(g))
    return out


def tile_runs(g: Graph) -> dict:
    cg = _Graph(g.vertex_count, g.edge_count, g.tile_count,
                *[a.ctypes.data_as(C.POINTER(t)) for a, t in [
                    (g.row_offsets, C.c_uint64), (g.destinations, C.c_uint32),
                    (g.weights, C.c_float), (g.edge_ids, C.c_uint64),
                    (g.column_offsets, C.c_uint64), (g.sources, C.c_uint32),
                    (g.in_weights, C.c_float), (g.in_edge_ids, C.c_uint64),
                    (g.old_to_new, C.c_uint32), (g.new_to_old, C.c_uint32),
                    (g.owner_tiles, C.c_uint32), (g.tile_x, C.c_int64),
                    (g.tile_y, C.c_int64), (g.tile_has_location, C.c_uint32),
                    (g.tile_vertex_offsets, C.c_uint64)]])
    r = _Runs()
    if lib().orc_tile_runs(C.byref(cg), C.byref(r)) != 0:
        raise ValueError("oracle tile runs failed")
    V = g.vertex_count
    out = dict(
        csr_row_run_offsets=_copy(r.csr_row_run_offsets, V + 1, np.uint64),
        csr_run_edge_offsets=_copy(r.csr_run_edge_offsets, r.csr_run_count + 1, np.uint64),
        csr_run_destination_tiles=_copy(r.csr_run_tiles, r.csr_run_count, np.uint32),
        csc_column_run_offsets=_copy(r.csc_column_run_offsets, V + 1, np.uint64),
        csc_run_edge_offsets=_copy(r.csc_run_edge_offsets, r.csc_run_count + 1, np.uint64),
        csc_run_source_tiles=_copy(r.csc_run_tiles, r.csc_run_count, np.uint32))
    lib().orc_tile_runs_free(C.byref(r))
    return out


def _csr(row_offsets, destinations, weights):
    return (np.ascontiguousarray(row_offsets, np.uint64),
            np.ascontiguousarray(destinations, np.uint32),
            np.ascontiguousarray(weights, np.float32))


def _sel(tile_selected):
    return None if tile_selected is None else np.ascontiguousarray(tile_selected, np.uint8)


def dijkstra(V, row_offsets, destinations, weights, owner_tiles, tile_selected, sources):
    ro, de, w = _csr(row_offsets, destinations, weights)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    src = np.ascontiguousarray(sources, np.uint32)
    sel = _sel(tile_selected)
    dist = np.empty(V, np.float32)
    relax = C.c_uint64(0)
    rc = lib().orc_dijkstra(V, _p(ro), _p(de), _p(w), _p(own), _p(sel), _p(src), len(src),
                            _p(dist), C.byref(relax))
    if rc != 0:
        raise ValueError(f"oracle dijkstra failed: {rc}")
    return dist, relax.value


def bellman_ford(V, row_offsets, destinations, weights, owner_tiles, tile_selected, sources):
    ro, de, w = _csr(row_offsets, destinations, weights)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    src = np.ascontiguousarray(sources, np.uint32)
    sel = _sel(tile_selected)
    dist = np.empty(V, np.float32)
    rc = lib().orc_bellman_ford(V, _p(ro), _p(de), _p(w), _p(own), _p(sel), _p(src),
                                len(src), _p(dist))
    if rc != 0:
        raise ValueError(f"oracle bellman-ford failed: {rc}")
    return dist


def jacobi_rounds(V, column_offsets, sources_csc, in_weights, owner_tiles, tile_selected,
                  sources, maximum_rounds=0):
    co, sc, w = _csr(column_offsets, sources_csc, in_weights)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    src = np.ascontiguousarray(sources, np.uint32)
    sel = _sel(tile_selected)
    dist = np.empty(V, np.float32)
    rounds = lib().orc_jacobi_rounds(V, _p(co), _p(sc), _p(w), _p(own), _p(sel), _p(src),
                                     len(src), maximum_rounds, _p(dist))
    if rounds < 0:
        raise ValueError("oracle jacobi failed")
    return int(rounds), dist


def predecessors(V, column_offsets, sources_csc, in_weights, in_edge_ids, owner_tiles,
                 tile_selected, distances, sources):
    co, sc, w = _csr(column_offsets, sources_csc, in_weights)
    ids = np.ascontiguousarray(in_edge_ids, np.uint64)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    d = np.ascontiguousarray(distances, np.float32)
    src = np.ascontiguousarray(sources, np.uint32)
    sel = _sel(tile_selected)
    pv = np.empty(V, np.uint32)
    pe = np.empty(V, np.uint64)
    rc = lib().orc_predecessors(V, _p(co), _p(sc), _p(w), _p(ids), _p(own), _p(sel), _p(d),
                                _p(src), len(src), _p(pv), _p(pe))
    if rc != 0:
        raise ValueError("oracle predecessors failed")
    return pv, pe


def reconstruct(V, column_offsets, sources_csc, in_weights, in_edge_ids, owner_tiles,
                tile_selected, distances, sources, target):
    """Returns (status, vertices, edge_ids); status 'complete', 'unreachable' or
    'no_tight_path'."""
    co, sc, w = _csr(column_offsets, sources_csc, in_weights)
    ids = np.ascontiguousarray(in_edge_ids, np.uint64)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    d = np.ascontiguousarray(distances, np.float32)
    src = np.ascontiguousarray(sources, np.uint32)
    sel = _sel(tile_selected)
    cap = V + 1
    pv = np.empty(cap + 1, np.uint32)
    pe = np.empty(cap, np.uint64)
    n = C.c_uint32(0)
    rc = lib().orc_reconstruct(V, _p(co), _p(sc), _p(w), _p(ids), _p(own), _p(sel), _p(d),
                               _p(src), len(src), int(target), cap, _p(pv), _p(pe),
                               C.byref(n))
    if rc == 1:
        return "unreachable", np.zeros(0, np.uint32), np.zeros(0, np.uint64)
    if rc == -3:
        return "no_tight_path", np.zeros(0, np.uint32), np.zeros(0, np.uint64)
    if rc != 0:
        raise ValueError(f"oracle reconstruct failed: {rc}")
    L = n.value
    return "complete", pv[:L + 1].copy(), pe[:L].copy()


CERTIFY_RULES = {0: "", 1: "a source label is not +0 or a label is not a nonnegative float",
                 2: "an admitted edge still lowers a label (not a fixed point)",
                 3: "a finite label has no tight admitted in-edge from a smaller label"}


def region_vertices(tile_vertex_offsets, tiles):
    """Global vertex ids of the selected tiles, in the compact (tile by tile)
    order of ``certify_region``'s labels."""
    tvo = np.asarray(tile_vertex_offsets, np.int64)
    t = np.asarray(tiles, np.int64)
    return np.concatenate([np.arange(tvo[x], tvo[x + 1], dtype=np.int64) for x in t]) \
        if len(t) else np.zeros(0, np.int64)


def certify_region(layout, tiles, labels, sources):
    """orc_certify_region over a product layout (``HostGraph.layout_arrays()``):
    labels tile by tile over the sorted selected tiles.  Returns (message,
    compact index); message '' means certified exact.  Releases the GIL, so
    queries can be certified on several threads."""
    tvo = np.ascontiguousarray(layout["tile_vertex_offsets"], np.uint32)
    col = np.ascontiguousarray(layout["csc_column_offsets"], np.uint32)
    srcs = np.ascontiguousarray(layout["csc_sources"], np.uint32)
    w = np.ascontiguousarray(layout["csc_weights"], np.float32)
    own = np.ascontiguousarray(layout["owner_tiles"], np.uint32)
    t = np.ascontiguousarray(tiles, np.uint32)
    lab = np.ascontiguousarray(labels, np.float32)
    s = np.ascontiguousarray(sources, np.uint32)
    where = C.c_uint64(0)
    rc = lib().orc_certify_region(layout["tile_count"], _p(tvo), _p(col), _p(srcs), _p(w), _p(own),
                                  len(t), _p(t), _p(lab), len(s), _p(s), C.byref(where))
    if rc < 0:
        raise ValueError(f"oracle certificate failed: {rc}")
    return CERTIFY_RULES[rc], where.value


def graph_sssp(g: Graph, sources, tile_selected=None):
    return dijkstra(g.vertex_count, g.row_offsets, g.destinations, g.weights, g.owner_tiles,
                    tile_selected, sources)


# ---------------------------------------------------------------------------
# The reference's own CPU sources, compiled into oracle/_ref.
# ---------------------------------------------------------------------------

_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        P = C.c_void_p
        _ref.ref_build.argtypes = [P, C.c_uint32, P, C.c_uint64, C.c_int32, C.c_int32,
                                   C.c_uint32, C.c_uint32]
        _ref.ref_build.restype = C.c_void_p
        _ref.ref_free.argtypes = [P]
        _ref.ref_counts.argtypes = [P, P]
        _ref.ref_graph_arrays.argtypes = [P] * 13
        _ref.ref_run_arrays.argtypes = [P] * 7
        _ref.ref_fingerprint.argtypes = [P, P]
        _ref.ref_sssp.argtypes = [P, P, P, C.c_uint32, C.c_int, P, P]
        _ref.ref_sssp.restype = C.c_int64
        _ref.ref_reconstruct.argtypes = [P, P, C.c_uint32, P, C.c_uint32, P, P, C.c_uint32]
        _ref.ref_reconstruct.restype = C.c_int64
    return _ref


def ref_csr_sssp(V, row_offsets, destinations, weights, owner_tiles, tile_selected, sources):
    """The reference's dijkstra_oracle on the induced subgraph (built by the
    reference's build_weighted_graph, untimed).  Returns (ns, relaxations)."""
    lib = ref_lib()
    if not hasattr(lib, "_csr_bound"):
        P = C.c_void_p
        lib.ref_csr_sssp.argtypes = [C.c_uint32, P, P, P, P, P, P, C.c_uint32, P]
        lib.ref_csr_sssp.restype = C.c_int64
        lib._csr_bound = True
    ro = np.ascontiguousarray(row_offsets, np.uint64)
    de = np.ascontiguousarray(destinations, np.uint32)
    w = np.ascontiguousarray(weights, np.float32)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    sel = np.ascontiguousarray(tile_selected, np.uint8)
    src = np.ascontiguousarray(sources, np.uint32)
    relax = C.c_uint64(0)
    ns = lib.ref_csr_sssp(V, _p(ro), _p(de), _p(w), _p(own), _p(sel), _p(src), len(src),
                          C.byref(relax))
    if ns < 0:
        raise ValueError(f"reference csr sssp failed: {ns}")
    return ns, relax.value


def ref_csr_sssp_parallel(V, row_offsets, destinations, weights, owner_tiles, tile_count,
                          query_tiles, sources, threads):
    """The reference's dijkstra_oracle over many single-source queries, one
    query per worker thread in waves of `threads` (induced-subgraph
    construction untimed).  Returns (summed wave wall ns, relaxations)."""
    lib = ref_lib()
    if not hasattr(lib, "_csr_parallel_bound"):
        P = C.c_void_p
        lib.ref_csr_sssp_parallel.argtypes = [C.c_uint32, P, P, P, P, C.c_uint32, C.c_uint32, P,
                                              P, P, C.c_uint32, P]
        lib.ref_csr_sssp_parallel.restype = C.c_int64
        lib._csr_parallel_bound = True
    ro = np.ascontiguousarray(row_offsets, np.uint64)
    de = np.ascontiguousarray(destinations, np.uint32)
    w = np.ascontiguousarray(weights, np.float32)
    own = np.ascontiguousarray(owner_tiles, np.uint32)
    off = np.zeros(len(query_tiles) + 1, np.uint32)
    off[1:] = np.cumsum([len(t) for t in query_tiles])
    tl = np.ascontiguousarray(np.concatenate([np.asarray(t, np.uint32) for t in query_tiles]),
                              np.uint32)
    src = np.ascontiguousarray(sources, np.uint32)
    relax = C.c_uint64(0)
    ns = lib.ref_csr_sssp_parallel(V, _p(ro), _p(de), _p(w), _p(own), tile_count, len(src),
                                   _p(off), _p(tl), _p(src), threads, C.byref(relax))
    if ns < 0:
        raise ValueError(f"reference parallel csr sssp failed: {ns}")
    return ns, relax.value


class RefGraph:
    """Handle over the reference's build_spatially_ordered_graph result."""

    def __init__(self, vertices, edges, config=None):
        vertices = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
        edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
        ox, oy, tw, th = config if config is not None else (0, 0, 0, 0)
        self.h = ref_lib().ref_build(_p(vertices), len(vertices), _p(edges), len(edges),
                                     ox, oy, tw, th)
        if not self.h:
            raise ValueError("reference graph build threw")
        counts = np.zeros(6, np.uint64)
        ref_lib().ref_counts(self.h, _p(counts))
        self.V, self.E, self.T, self.Rcsr, self.Rcsc, _ = (int(c) for c in counts)

    def __del__(self):
        if getattr(self, "h", None):
            ref_lib().ref_free(self.h)
            self.h = None

    def arrays(self) -> dict:
        V, E, T = self.V, self.E, self.T
        a = dict(row_offsets=np.zeros(V + 1, np.uint64), destinations=np.zeros(E, np.uint32),
                 weights=np.zeros(E, np.float32), edge_ids=np.zeros(E, np.uint64),
                 column_offsets=np.zeros(V + 1, np.uint64), sources=np.zeros(E, np.uint32),
                 in_weights=np.zeros(E, np.float32), in_edge_ids=np.zeros(E, np.uint64),
                 old_to_new=np.zeros(V, np.uint32), new_to_old=np.zeros(V, np.uint32),
                 owner_tiles=np.zeros(V, np.uint32),
                 tile_vertex_offsets=np.zeros(T + 1, np.uint64))
        ref_lib().ref_graph_arrays(self.h, *[_p(v) for v in a.values()])
        return a

    def runs(self) -> dict:
        V = self.V
        r = dict(csr_row_run_offsets=np.zeros(V + 1, np.uint64),
                 csr_run_edge_offsets=np.zeros(self.Rcsr + 1, np.uint64),
                 csr_run_destination_tiles=np.zeros(self.Rcsr, np.uint32),
                 csc_column_run_offsets=np.zeros(V + 1, np.uint64),
                 csc_run_edge_offsets=np.zeros(self.Rcsc + 1, np.uint64),
                 csc_run_source_tiles=np.zeros(self.Rcsc, np.uint32))
        ref_lib().ref_run_arrays(self.h, *[_p(v) for v in r.values()])
        return r

    def fingerprint(self):
        out = np.zeros(2, np.uint64)
        ref_lib().ref_fingerprint(self.h, _p(out))
        return int(out[0]), int(out[1])

    def sssp(self, sources, tile_selected=None, algorithm=0):
        src = np.ascontiguousarray(sources, np.uint32)
        sel = _sel(tile_selected)
        dist = np.empty(self.V, np.float32)
        relax = C.c_uint64(0)
        ns = ref_lib().ref_sssp(self.h, _p(sel), _p(src), len(src), algorithm, _p(dist),
                                C.byref(relax))
        if ns < 0:
            raise ValueError(f"reference sssp failed: {ns}")
        return dist, relax.value, ns

    def reconstruct(self, sources, distances, target):
        src = np.ascontiguousarray(sources, np.uint32)
        d = np.ascontiguousarray(distances, np.float32)
        cap = self.V + 1
        pv = np.zeros(cap + 1, np.uint32)
        pe = np.zeros(cap, np.uint64)
        n = ref_lib().ref_reconstruct(self.h, _p(src), len(src), _p(d), int(target), _p(pv),
                                      _p(pe), cap)
        if n == -1:
            return None
        if n < 0:
            raise ValueError("reference reconstruct failed")
        return pv[:n + 1].copy(), pe[:n].copy()
