"""ctypes binding of the bfnew MI355X engine's C ABI (include/bfnew_amd.h) and
of the host graph library (include/bfnew_host.h).

The compute path is the HIP library ``lib/libbfnew_hip.so``; this module only
marshals arrays.  Loading fails loudly when a library is missing: there is no
fallback implementation behind it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB_PATH = os.path.join(LIB_DIR, "libbfnew_hip.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "libbfnew.so")

JACOBI_PULL, DENSE_CHAOTIC_PUSH, FRONTIER_PUSH = 0, 1, 2
PERSISTENT, CHUNKED, PER_ROUND = 0, 1, 2
ENGINE_NAMES = {JACOBI_PULL: "jacobi_pull", DENSE_CHAOTIC_PUSH: "dense_chaotic_push",
                FRONTIER_PUSH: "frontier_push"}
CONTROL_NAMES = {PERSISTENT: "persistent_cooperative", CHUNKED: "chunked_host_poll",
                 PER_ROUND: "per_round_host_poll"}
NONE32 = 0xFFFFFFFF
INF_BITS = 0x7F800000


class BfnError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"bfnew error {code}: {message}")
        self.code = code


class RunOptions(C.Structure):
    _fields_ = [("engine", C.c_uint32), ("control_mode", C.c_uint32),
                ("rounds_per_chunk", C.c_uint32), ("block_size", C.c_uint32),
                ("grid_policy", C.c_uint32), ("blocks_per_wgp", C.c_uint32),
                ("instrumentation", C.c_uint32), ("pad0", C.c_uint32),
                ("maximum_rounds", C.c_uint64), ("enable_per_lane_convergence", C.c_uint32),
                ("pad1", C.c_uint32)]


class RunStatus(C.Structure):
    _fields_ = [("final_distance_slot", C.c_uint32), ("converged", C.c_uint32),
                ("rounds_completed", C.c_uint64), ("reached_target_mask", C.c_uint32),
                ("bounding_box_miss_mask", C.c_uint32), ("valid_lane_mask", C.c_uint32),
                ("active_lane_mask", C.c_uint32), ("converged_lane_mask", C.c_uint32),
                ("stop_reason", C.c_uint32), ("error_bits", C.c_uint32), ("pad", C.c_uint32)]


class WorkStatistics(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "edges_examined", "successful_decreases", "active_vertices", "active_lane_rounds",
        "maximum_queue_size", "host_checks", "host_synchronizations", "controller_copies",
        "kernel_dispatches", "expansion_count", "atomic_attempts", "successful_atomic_updates",
        "queue_claims", "duplicate_suppressions", "mask_operations", "overflow_events",
        "high_contention_destinations", "changed_flag_updates", "full_edge_rounds",
        "empty_frontier_rounds", "small_frontier_rounds")]


class TargetSummary(C.Structure):
    _fields_ = [("target", C.c_uint32), ("selected_source", C.c_uint32),
                ("distance", C.c_float), ("path_length", C.c_uint32), ("reached", C.c_uint32),
                ("reconstruction", C.c_uint32), ("has_selected_source", C.c_uint32)]


class QueryStatus(C.Structure):
    _fields_ = [("rounds_completed", C.c_uint64), ("relaxations", C.c_uint64),
                ("converged", C.c_uint32), ("reached_all_targets", C.c_uint32),
                ("bounding_box_miss", C.c_uint32), ("stop_reason", C.c_uint32),
                ("error_bits", C.c_uint32), ("engine", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("tile_sweeps", C.c_uint32), ("items_per_claim", C.c_uint32),
                ("graph_rounds", C.c_uint32), ("use_graph", C.c_uint32),
                ("count_relaxations", C.c_uint32), ("boundary_mode", C.c_uint32),
                ("min_waves", C.c_uint32), ("claim_order", C.c_uint32),
                ("push_window", C.c_float)]


_P32 = C.POINTER(C.c_uint32)
_PF = C.POINTER(C.c_float)


class GraphLayout32(C.Structure):
    _fields_ = [("vertex_count", C.c_uint32), ("edge_count", C.c_uint32),
                ("tile_count", C.c_uint32), ("csr_run_count", C.c_uint32),
                ("csc_run_count", C.c_uint32), ("pad", C.c_uint32),
                ("owner_tiles", _P32), ("tile_vertex_offsets", _P32),
                ("csr_row_offsets", _P32), ("csr_destinations", _P32), ("csr_weights", _PF),
                ("csr_row_run_offsets", _P32), ("csr_run_edge_offsets", _P32),
                ("csr_run_destination_tiles", _P32), ("csc_column_offsets", _P32),
                ("csc_sources", _P32), ("csc_weights", _PF), ("csc_edge_ids", _P32),
                ("csc_column_run_offsets", _P32), ("csc_run_edge_offsets", _P32),
                ("csc_run_source_tiles", _P32)]


class BatchDesc(C.Structure):
    _fields_ = [("query_count", C.c_uint32), ("pad", C.c_uint32), ("query_ids", _P32),
                ("source_offsets", _P32), ("sources", _P32), ("target_offsets", _P32),
                ("targets", _P32), ("tile_offsets", _P32), ("selected_tiles", _P32),
                ("engine_by_query", _P32)]


class GraphInfo(C.Structure):
    _fields_ = [("device_bytes", C.c_uint64), ("max_tile_vertices", C.c_uint32),
                ("max_tile_in_edges", C.c_uint32), ("max_tile_out_edges", C.c_uint32),
                ("max_in_neighbor_tiles", C.c_uint32), ("max_out_neighbor_tiles", C.c_uint32),
                ("max_stage_floats", C.c_uint32), ("has_zero_weight", C.c_uint32),
                ("pad", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("compute_units", C.c_uint32), ("xcd_count", C.c_uint32),
                ("wavefront_size", C.c_uint32), ("lds_bytes_per_cu", C.c_uint32),
                ("global_memory_bytes", C.c_uint64), ("arch", C.c_char * 32)]


class HostGraphView(C.Structure):
    _fields_ = [("vertex_count", C.c_uint32), ("tile_count", C.c_uint32),
                ("edge_count", C.c_uint64), ("row_offsets", C.POINTER(C.c_uint64)),
                ("destinations", _P32), ("weights", _PF), ("edge_ids", C.POINTER(C.c_uint64)),
                ("column_offsets", C.POINTER(C.c_uint64)), ("sources", _P32),
                ("in_weights", _PF), ("in_edge_ids", C.POINTER(C.c_uint64)),
                ("old_to_new", _P32), ("new_to_old", _P32), ("owner_tiles", _P32),
                ("tile_vertex_offsets", C.POINTER(C.c_uint64)),
                ("tile_x", C.POINTER(C.c_int64)), ("tile_y", C.POINTER(C.c_int64)),
                ("tile_has_location", _P32), ("node_costs", _PF)]


class HostQuery(C.Structure):
    _fields_ = [("source", C.c_uint32), ("target", C.c_uint32), ("box_x", C.c_uint32),
                ("box_y", C.c_uint32), ("box_side", C.c_uint32), ("unbounded", C.c_uint32)]


class ServerOptions(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("max_active_queries", C.c_uint32),
                ("rounds_per_chunk", C.c_uint32), ("engine", C.c_uint32),
                ("tile_sweeps", C.c_uint32), ("block_size", C.c_uint32),
                ("want_paths", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("source", C.c_uint32), ("target", C.c_uint32),
                ("tile_count", C.c_uint32), ("pad", C.c_uint32), ("selected_tiles", _P32)]


class Response(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("distance", C.c_float),
                ("path_length", C.c_uint32), ("reached", C.c_uint32), ("pad", C.c_uint32),
                ("submit_ns", C.c_double), ("complete_ns", C.c_double)]


PHASE_NAMES = ("setup", "internal", "closure", "boundary", "writeback")


class EngineCounters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("tile_visits", "edges_relaxed", "algorithmic_bytes",
                                          "atomic_attempts", "atomic_successes", "maximum_list",
                                          "rounds_executed", "closure_sweeps")] + [
        ("phase_cycles", C.c_uint64 * len(PHASE_NAMES)), ("candidate_evaluations", C.c_uint64)]


VERTEX_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("has_location", "<u4"),
                         ("resource_class", "<u4")])
EDGE_DTYPE = np.dtype([("source", "<u4"), ("destination", "<u4"), ("weight", "<f4"),
                       ("provenance_domain", "<u4"), ("provenance_kind", "<u4"),
                       ("pad", "<u4"), ("provenance_record", "<u8")])

_hip = None
_host = None

HIP_SIGNATURES = {
    "bfn_last_error": ([], C.c_char_p),
    "bfn_abi_version": ([], C.c_uint32),
    "bfn_device_open": ([C.c_int32, C.POINTER(C.c_void_p)], C.c_int32),
    "bfn_device_close": ([C.c_void_p], C.c_int32),
    "bfn_device_query": ([C.c_void_p, C.POINTER(DeviceInfo)], C.c_int32),
    "bfn_device_synchronize": ([C.c_void_p], C.c_int32),
    "bfn_graph_upload": ([C.c_void_p, C.POINTER(GraphLayout32), C.POINTER(C.c_void_p)], C.c_int32),
    "bfn_graph_validate": ([C.POINTER(GraphLayout32)], C.c_int32),
    "bfn_batch_region_distances": ([C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64)], C.c_int32),
    "bfn_graph_free": ([C.c_void_p], C.c_int32),
    "bfn_graph_query": ([C.c_void_p, C.POINTER(GraphInfo)], C.c_int32),
    "bfn_graph_update_weights": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "bfn_graph_update_node_costs": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "bfn_graph_update_node_costs_from_workspace": ([C.c_void_p, C.c_void_p, C.c_void_p],
                                                   C.c_int32),
    "bfn_workspace_create": ([C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int32),
    "bfn_workspace_free": ([C.c_void_p], C.c_int32),
    "bfn_tuning_default": ([C.POINTER(Tuning)], None),
    "bfn_batch_estimate": ([C.c_void_p, C.POINTER(BatchDesc), C.POINTER(RunOptions),
                            C.POINTER(C.c_uint64)], C.c_int32),
    "bfn_batch_submit": ([C.c_void_p, C.c_void_p, C.POINTER(BatchDesc), C.POINTER(RunOptions),
                          C.POINTER(Tuning)], C.c_int32),
    "bfn_batch_rerun": ([C.c_void_p], C.c_int32),
    "bfn_batch_wait": ([C.c_void_p], C.c_int32),
    "bfn_batch_status": ([C.c_void_p, C.POINTER(RunStatus), C.c_void_p,
                          C.POINTER(WorkStatistics)], C.c_int32),
    "bfn_batch_distances": ([C.c_void_p, C.c_uint32, C.c_void_p], C.c_int32),
    "bfn_batch_predecessors": ([C.c_void_p, C.c_uint32, C.c_void_p], C.c_int32),
    "bfn_batch_paths": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_uint64, C.POINTER(C.c_uint64)], C.c_int32),
    "bfn_batch_occupancy": ([C.c_void_p, C.c_void_p], C.c_int32),
    "bfn_batch_counters": ([C.c_void_p, C.POINTER(EngineCounters)], C.c_int32),
    "bfn_batch_run_static": ([C.c_void_p, C.POINTER(BatchDesc), C.POINTER(RunOptions),
                              C.POINTER(Tuning), C.c_void_p, C.c_void_p], C.c_int32),
    "bfn_static_scratch_bytes": ([], C.c_uint64),
    "bfn_batch_timing": ([C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                          C.POINTER(C.c_uint64)], C.c_int32),
    "bfn_server_create": ([C.c_void_p, C.c_void_p, C.POINTER(ServerOptions),
                           C.POINTER(C.c_void_p)], C.c_int32),
    "bfn_server_submit": ([C.c_void_p, C.POINTER(Request)], C.c_int32),
    "bfn_server_poll": ([C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32],
                        C.c_int32),
    "bfn_server_poll_paths": ([C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                               C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p],
                              C.c_int32),
    "bfn_server_stats": ([C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                          C.POINTER(C.c_uint64)], C.c_int32),
    "bfn_server_destroy": ([C.c_void_p], C.c_int32),
}

HOST_SIGNATURES = {
    "bfnh_last_error": ([], C.c_char_p),
    "bfnh_graph_build": ([C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_int32, C.c_int32,
                          C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)], C.c_int32),
    "bfnh_synthetic_graph": ([C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                              C.c_uint32, C.POINTER(C.c_void_p)], C.c_int32),
    "bfnh_synthetic_generic": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int32),
    "bfnh_synthetic_input": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "bfnh_graph_free": ([C.c_void_p], C.c_int32),
    "bfnh_graph_view_get": ([C.c_void_p, C.POINTER(HostGraphView)], C.c_int32),
    "bfnh_graph_layout": ([C.c_void_p, C.POINTER(GraphLayout32)], C.c_int32),
    "bfnh_graph_fingerprint": ([C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
                               C.c_int32),
    "bfnh_route_query_tiles": ([C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                C.c_uint32, C.POINTER(C.c_uint32)], C.c_int32),
    "bfnh_synthetic_queries": ([C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                C.c_double, C.c_uint32, C.POINTER(HostQuery)], C.c_int32),
    "bfnh_synthetic_query_tiles": ([C.c_void_p, C.POINTER(HostQuery), C.c_void_p, C.c_uint32,
                                    C.POINTER(C.c_uint32)], C.c_int32),
}


def _bind(lib, signatures):
    for name, (args, res) in signatures.items():
        if not hasattr(lib, name):
            continue  # calling it raises AttributeError; the ABI test lists it
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def hip_lib():
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise ImportError(f"engine library missing: {HIP_LIB_PATH} (run build())")
        _hip = C.CDLL(HIP_LIB_PATH, mode=C.RTLD_GLOBAL)
        _bind(_hip, HIP_SIGNATURES)
    return _hip


def host_lib():
    global _host
    if _host is None:
        hip_lib()
        if not os.path.exists(HOST_LIB_PATH):
            raise ImportError(f"host library missing: {HOST_LIB_PATH} (run build())")
        _host = C.CDLL(HOST_LIB_PATH)
        _bind(_host, HOST_SIGNATURES)
    return _host


def _check(rc, host=False):
    if rc != 0:
        msg = (host_lib().bfnh_last_error() if host else hip_lib().bfn_last_error()) or b""
        raise BfnError(rc, msg.decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _p32(a):
    return a.ctypes.data_as(_P32)


def run_options(engine=JACOBI_PULL, control_mode=PERSISTENT, rounds_per_chunk=8, block_size=256,
                maximum_rounds=1 << 40, per_lane=1, instrumentation=0, grid_policy=0,
                blocks_per_wgp=0):
    o = RunOptions()
    o.engine, o.control_mode, o.rounds_per_chunk = engine, control_mode, rounds_per_chunk
    o.block_size, o.maximum_rounds = block_size, maximum_rounds
    o.enable_per_lane_convergence, o.instrumentation = per_lane, instrumentation
    o.grid_policy, o.blocks_per_wgp = grid_policy, blocks_per_wgp
    return o


def tuning(tile_sweeps=1, items_per_claim=1, graph_rounds=0, use_graph=0, count_relaxations=0,
           boundary_mode=0, min_waves=0, claim_order=0, push_window=0.0):
    t = Tuning()
    hip_lib().bfn_tuning_default(C.byref(t))
    t.tile_sweeps, t.items_per_claim = tile_sweeps, items_per_claim
    t.graph_rounds, t.use_graph, t.count_relaxations = graph_rounds, use_graph, count_relaxations
    t.boundary_mode, t.min_waves, t.claim_order = boundary_mode, min_waves, claim_order
    t.push_window = push_window
    return t


# ---------------------------------------------------------------------------
# Host graphs
# ---------------------------------------------------------------------------
def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


_LAYOUT_ARRAYS = [name for name, _ in GraphLayout32._fields_[6:]]


def layout_from_arrays(arrays: dict) -> GraphLayout32:
    """A GraphLayout32 over the numpy arrays of ``HostGraph.layout_arrays()``
    (or a modified copy); the record keeps the arrays alive."""
    L = GraphLayout32()
    L.vertex_count, L.edge_count, L.tile_count = (arrays["vertex_count"], arrays["edge_count"],
                                                  arrays["tile_count"])
    L.csr_run_count = len(arrays["csr_run_destination_tiles"])
    L.csc_run_count = len(arrays["csc_run_source_tiles"])
    keep = []
    for name in _LAYOUT_ARRAYS:
        dtype = np.float32 if name.endswith("weights") else np.uint32
        a = np.ascontiguousarray(arrays[name], dtype=dtype)
        keep.append(a)
        setattr(L, name, a.ctypes.data_as(_PF if dtype == np.float32 else _P32))
    L._keep = keep
    return L


def validate_layout(layout: GraphLayout32):
    """(status, message) of bfn_graph_validate: the host checks that
    bfn_graph_upload runs before it touches the device."""
    lib = hip_lib()
    rc = lib.bfn_graph_validate(C.byref(layout))
    return rc, (lib.bfn_last_error().decode() if rc else "")


class HostGraph:
    """A graph built by the host library (reference-order build or synthetic)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self._layout = None

    @classmethod
    def build(cls, vertices, edges, config=None):
        v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
        e = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
        ox, oy, tw, th = config if config is not None else (0, 0, 0, 0)
        h = C.c_void_p()
        _check(host_lib().bfnh_graph_build(_ptr(v), len(v), _ptr(e), len(e), ox, oy, tw, th,
                                           C.byref(h)), host=True)
        return cls(h)

    @classmethod
    def synthetic(cls, tiles_x, tiles_y, nodes_per_tile, seed, threads=0, host_graph=False):
        h = C.c_void_p()
        _check(host_lib().bfnh_synthetic_graph(tiles_x, tiles_y, nodes_per_tile, seed, threads,
                                               1 if host_graph else 0, C.byref(h)), host=True)
        g = cls(h)
        g.grid = (tiles_x, tiles_y, nodes_per_tile, seed)
        return g

    def generic_rebuild(self):
        h = C.c_void_p()
        _check(host_lib().bfnh_synthetic_generic(self.h, C.byref(h)), host=True)
        return HostGraph(h)

    def synthetic_input(self):
        v = self.view()
        verts = np.zeros(v["vertex_count"], VERTEX_DTYPE)
        edges = np.zeros(v["edge_count"], EDGE_DTYPE)
        _check(host_lib().bfnh_synthetic_input(self.h, _ptr(verts), _ptr(edges)), host=True)
        return verts, edges

    def close(self):
        if self.h:
            host_lib().bfnh_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self) -> dict:
        v = HostGraphView()
        _check(host_lib().bfnh_graph_view_get(self.h, C.byref(v)), host=True)
        V, E, T = v.vertex_count, v.edge_count, v.tile_count
        out = dict(vertex_count=V, edge_count=E, tile_count=T,
                   owner_tiles=_arr(v.owner_tiles, V, np.uint32).copy(),
                   tile_vertex_offsets=_arr(v.tile_vertex_offsets, T + 1, np.uint64).copy(),
                   tile_x=_arr(v.tile_x, T, np.int64).copy(),
                   tile_y=_arr(v.tile_y, T, np.int64).copy(),
                   tile_has_location=_arr(v.tile_has_location, T, np.uint32).copy())
        if v.row_offsets:
            out.update(row_offsets=_arr(v.row_offsets, V + 1, np.uint64).copy(),
                       destinations=_arr(v.destinations, E, np.uint32).copy(),
                       weights=_arr(v.weights, E, np.float32).copy(),
                       edge_ids=_arr(v.edge_ids, E, np.uint64).copy(),
                       column_offsets=_arr(v.column_offsets, V + 1, np.uint64).copy(),
                       sources=_arr(v.sources, E, np.uint32).copy(),
                       in_weights=_arr(v.in_weights, E, np.float32).copy(),
                       in_edge_ids=_arr(v.in_edge_ids, E, np.uint64).copy(),
                       old_to_new=_arr(v.old_to_new, V, np.uint32).copy(),
                       new_to_old=_arr(v.new_to_old, V, np.uint32).copy())
        if v.node_costs:
            out["node_costs"] = _arr(v.node_costs, V, np.float32).copy()
        return out

    def layout(self) -> GraphLayout32:
        L = GraphLayout32()
        _check(host_lib().bfnh_graph_layout(self.h, C.byref(L)), host=True)
        return L

    def layout_arrays(self) -> dict:
        L = self.layout()
        V, E, T = L.vertex_count, L.edge_count, L.tile_count
        Rr, Rc = L.csr_run_count, L.csc_run_count
        a = lambda p, n, dt=np.uint32: _arr(p, n, dt).copy()  # noqa: E731
        return dict(vertex_count=V, edge_count=E, tile_count=T,
                    owner_tiles=a(L.owner_tiles, V), tile_vertex_offsets=a(L.tile_vertex_offsets, T + 1),
                    csr_row_offsets=a(L.csr_row_offsets, V + 1),
                    csr_destinations=a(L.csr_destinations, E),
                    csr_weights=a(L.csr_weights, E, np.float32),
                    csr_row_run_offsets=a(L.csr_row_run_offsets, V + 1),
                    csr_run_edge_offsets=a(L.csr_run_edge_offsets, Rr + 1),
                    csr_run_destination_tiles=a(L.csr_run_destination_tiles, Rr),
                    csc_column_offsets=a(L.csc_column_offsets, V + 1),
                    csc_sources=a(L.csc_sources, E), csc_weights=a(L.csc_weights, E, np.float32),
                    csc_edge_ids=a(L.csc_edge_ids, E),
                    csc_column_run_offsets=a(L.csc_column_run_offsets, V + 1),
                    csc_run_edge_offsets=a(L.csc_run_edge_offsets, Rc + 1),
                    csc_run_source_tiles=a(L.csc_run_source_tiles, Rc))

    def fingerprint(self):
        a, b = C.c_uint64(), C.c_uint64()
        _check(host_lib().bfnh_graph_fingerprint(self.h, C.byref(a), C.byref(b)), host=True)
        return a.value, b.value

    def route_tiles(self, terminals, padding=0):
        t = np.ascontiguousarray(terminals, np.uint32)
        n = C.c_uint32()
        _check(host_lib().bfnh_route_query_tiles(self.h, _ptr(t), len(t), padding, None, 0,
                                                 C.byref(n)), host=True)
        out = np.zeros(n.value, np.uint32)
        _check(host_lib().bfnh_route_query_tiles(self.h, _ptr(t), len(t), padding, _ptr(out),
                                                 n.value, C.byref(n)), host=True)
        return out

    def synthetic_queries(self, count, seed, side_min=4, side_max=32, unbounded_fraction=0.0,
                          mix=0):
        arr = (HostQuery * max(count, 1))()
        _check(host_lib().bfnh_synthetic_queries(self.h, count, seed, side_min, side_max,
                                                 unbounded_fraction, mix, arr), host=True)
        return [arr[i] for i in range(count)]

    def query_tiles(self, q: HostQuery):
        n = C.c_uint32()
        _check(host_lib().bfnh_synthetic_query_tiles(self.h, C.byref(q), None, 0, C.byref(n)),
               host=True)
        out = np.zeros(n.value, np.uint32)
        _check(host_lib().bfnh_synthetic_query_tiles(self.h, C.byref(q), _ptr(out), n.value,
                                                     C.byref(n)), host=True)
        return out


# ---------------------------------------------------------------------------
# Device side
# ---------------------------------------------------------------------------
class Device:
    def __init__(self, ordinal=0):
        self.h = C.c_void_p()
        _check(hip_lib().bfn_device_open(ordinal, C.byref(self.h)))

    def info(self) -> DeviceInfo:
        d = DeviceInfo()
        _check(hip_lib().bfn_device_query(self.h, C.byref(d)))
        return d

    def close(self):
        if self.h:
            hip_lib().bfn_device_close(self.h)
            self.h = None


class DeviceGraph:
    def __init__(self, device: Device, layout: GraphLayout32):
        self.device = device
        self.vertex_count = layout.vertex_count
        self.edge_count = layout.edge_count
        self.h = C.c_void_p()
        _check(hip_lib().bfn_graph_upload(device.h, C.byref(layout), C.byref(self.h)))

    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(hip_lib().bfn_graph_query(self.h, C.byref(i)))
        return i

    def update_node_costs(self, base, occupancy=None):
        b = np.ascontiguousarray(base, np.float32)
        o = None if occupancy is None else np.ascontiguousarray(occupancy, np.int32)
        _check(hip_lib().bfn_graph_update_node_costs(self.h, _ptr(b), _ptr(o)))

    def update_node_costs_from_workspace(self, workspace, base=None):
        b = None if base is None else np.ascontiguousarray(base, np.float32)
        _check(hip_lib().bfn_graph_update_node_costs_from_workspace(self.h, workspace.h, _ptr(b)))

    def update_weights(self, csr_weights, csc_weights):
        a = np.ascontiguousarray(csr_weights, np.float32)
        b = np.ascontiguousarray(csc_weights, np.float32)
        if a.shape != (self.edge_count,) or b.shape != (self.edge_count,):
            raise ValueError(f"update_weights: both orientations need {self.edge_count} weights")
        _check(hip_lib().bfn_graph_update_weights(self.h, _ptr(a), _ptr(b)))

    def close(self):
        if self.h:
            hip_lib().bfn_graph_free(self.h)
            self.h = None


@dataclass
class Batch:
    """Queries for one submit: per query canonical sources, targets and sorted
    selected tiles (global ids)."""
    sources: list
    targets: list
    tiles: list
    engines: list | None = None
    query_ids: list | None = None
    _keep: list = field(default_factory=list)

    def desc(self) -> BatchDesc:
        Q = len(self.sources)

        def flat(lists):
            off = np.zeros(Q + 1, np.uint32)
            off[1:] = np.cumsum([len(x) for x in lists])
            vals = (np.concatenate([np.asarray(x, np.uint32) for x in lists])
                    if Q else np.zeros(0, np.uint32))
            return off, np.ascontiguousarray(vals, np.uint32)

        so, sv = flat([sorted(set(int(v) for v in s)) for s in self.sources])
        to, tv = flat([sorted(set(int(v) for v in t)) for t in self.targets])
        ko, kv = flat(self.tiles)
        qid = np.arange(Q, dtype=np.uint32) if self.query_ids is None else np.asarray(
            self.query_ids, np.uint32)
        eng = None if self.engines is None else np.asarray(self.engines, np.uint32)
        self._keep = [so, sv, to, tv, ko, kv, qid, eng]
        d = BatchDesc()
        d.query_count = Q
        d.query_ids = _p32(qid)
        d.source_offsets, d.sources = _p32(so), _p32(sv if len(sv) else np.zeros(1, np.uint32))
        d.target_offsets, d.targets = _p32(to), _p32(tv if len(tv) else np.zeros(1, np.uint32))
        d.tile_offsets, d.selected_tiles = _p32(ko), _p32(kv if len(kv) else np.zeros(1, np.uint32))
        d.engine_by_query = None if eng is None else _p32(eng)
        self.target_offsets = to
        return d


class Workspace:
    def __init__(self, device: Device, budget_bytes=0):
        self.device = device
        self.h = C.c_void_p()
        _check(hip_lib().bfn_workspace_create(device.h, budget_bytes, C.byref(self.h)))
        self.batch = None
        self.vertex_count = 0

    def estimate(self, graph: DeviceGraph, batch: Batch, options: RunOptions) -> int:
        d = batch.desc()
        n = C.c_uint64()
        _check(hip_lib().bfn_batch_estimate(graph.h, C.byref(d), C.byref(options), C.byref(n)))
        return n.value

    def submit(self, graph: DeviceGraph, batch: Batch, options: RunOptions, tune: Tuning = None,
               wait=True):
        d = batch.desc()
        self.batch = batch
        self.desc = d
        self.vertex_count = graph.vertex_count
        t = tune if tune is not None else tuning()
        _check(hip_lib().bfn_batch_submit(self.h, graph.h, C.byref(d), C.byref(options),
                                          C.byref(t)))
        if wait:
            self.wait()

    def rerun(self, wait=True):
        _check(hip_lib().bfn_batch_rerun(self.h))
        if wait:
            self.wait()

    def wait(self):
        _check(hip_lib().bfn_batch_wait(self.h))

    def status(self):
        Q = len(self.batch.sources)
        st = RunStatus()
        per = (QueryStatus * Q)()
        work = WorkStatistics()
        _check(hip_lib().bfn_batch_status(self.h, C.byref(st), per, C.byref(work)))
        return st, [per[i] for i in range(Q)], work

    def distances(self, query: int) -> np.ndarray:
        out = np.empty(self.vertex_count, np.float32)
        _check(hip_lib().bfn_batch_distances(self.h, query, _ptr(out)))
        return out

    def region_distances(self, query: int) -> np.ndarray:
        """Labels of the query's selected tiles, tile by tile (sorted tiles)."""
        n = C.c_uint64()
        _check(hip_lib().bfn_batch_region_distances(self.h, query, None, 0, C.byref(n)))
        out = np.empty(max(n.value, 1), np.float32)
        _check(hip_lib().bfn_batch_region_distances(self.h, query, _ptr(out), n.value,
                                                    C.byref(n)))
        return out[:n.value]

    def predecessors(self, query: int) -> np.ndarray:
        out = np.empty(self.vertex_count, np.uint32)
        _check(hip_lib().bfn_batch_predecessors(self.h, query, _ptr(out)))
        return out

    def path_arrays(self) -> dict:
        """Compact paths of every target as bfn_batch_paths returns them: the
        target summaries (a record array with the bfn_target_summary fields),
        the path offsets (target i's path is [offsets[i], offsets[i + 1]) of
        the arenas) and the vertex, label and edge-id arenas."""
        NT = int(self.batch.target_offsets[-1])
        summ = (TargetSummary * max(NT, 1))()
        offs = np.zeros(NT + 1, np.uint64)
        total = C.c_uint64()
        _check(hip_lib().bfn_batch_paths(self.h, summ, _ptr(offs), None, None, None, 0,
                                         C.byref(total)))
        n = max(total.value, 1)
        verts = np.empty(n, np.uint32)
        labels = np.empty(n, np.float32)
        edges = np.empty(n, np.uint32)
        _check(hip_lib().bfn_batch_paths(self.h, summ, _ptr(offs), _ptr(verts), _ptr(labels),
                                         _ptr(edges), n, C.byref(total)))
        return dict(summaries=np.ctypeslib.as_array(summ)[:NT], offsets=offs, vertices=verts,
                    labels=labels, edge_ids=edges)

    def paths(self):
        """Compact paths of every target, one dict per target."""
        a = self.path_arrays()
        summ, offs = a["summaries"], a["offsets"]
        out = []
        for i in range(len(summ)):
            s = summ[i]
            lo, hi = int(offs[i]), int(offs[i + 1])
            out.append(dict(target=int(s["target"]), selected_source=int(s["selected_source"]),
                            distance=float(s["distance"]), path_length=int(s["path_length"]),
                            reached=int(s["reached"]), reconstruction=int(s["reconstruction"]),
                            has_selected_source=int(s["has_selected_source"]),
                            vertices=a["vertices"][lo:hi].copy(), labels=a["labels"][lo:hi].copy(),
                            edge_ids=a["edge_ids"][lo:lo + max(0, hi - lo - 1)].copy()))
        return out

    def occupancy(self) -> np.ndarray:
        out = np.zeros(self.vertex_count, np.int32)
        _check(hip_lib().bfn_batch_occupancy(self.h, _ptr(out)))
        return out

    def counters(self) -> EngineCounters:
        c = EngineCounters()
        _check(hip_lib().bfn_batch_counters(self.h, C.byref(c)))
        return c

    def timing(self):
        r, t, n = C.c_float(), C.c_float(), C.c_uint64()
        _check(hip_lib().bfn_batch_timing(self.h, C.byref(r), C.byref(t), C.byref(n)))
        return r.value, t.value, n.value

    def close(self):
        if self.h:
            hip_lib().bfn_workspace_free(self.h)
            self.h = None


class Server:
    """Continuous-batching route server over a resident graph (bfn_server_*).
    Requests join the running batch between chunks of rounds and come back
    as soon as their own work has drained."""

    def __init__(self, device: Device, graph: DeviceGraph, budget_bytes=1 << 30,
                 max_active_queries=256, rounds_per_chunk=4, engine=FRONTIER_PUSH, tile_sweeps=0,
                 block_size=256, want_paths=1):
        o = ServerOptions()
        o.budget_bytes, o.max_active_queries = budget_bytes, max_active_queries
        o.rounds_per_chunk, o.engine, o.tile_sweeps = rounds_per_chunk, engine, tile_sweeps
        o.block_size, o.want_paths = block_size, want_paths
        self.device, self.graph = device, graph
        self.h = C.c_void_p()
        _check(hip_lib().bfn_server_create(device.h, graph.h, C.byref(o), C.byref(self.h)))

    def submit(self, request_id, source, target, tiles):
        t = np.ascontiguousarray(tiles, np.uint32)
        r = Request()
        r.request_id, r.source, r.target, r.tile_count = request_id, source, target, len(t)
        r.selected_tiles = _p32(t)
        _check(hip_lib().bfn_server_submit(self.h, C.byref(r)))

    def poll(self, capacity=1024, timeout_ms=100):
        buf = (Response * capacity)()
        n = C.c_uint32()
        _check(hip_lib().bfn_server_poll(self.h, buf, capacity, C.byref(n), timeout_ms))
        return [dict(request_id=buf[i].request_id, distance=buf[i].distance,
                     path_length=buf[i].path_length, reached=buf[i].reached,
                     submit_ns=buf[i].submit_ns, complete_ns=buf[i].complete_ns)
                for i in range(n.value)]

    def poll_paths(self, capacity=1024, timeout_ms=100, vertex_capacity=1 << 22):
        """poll() with each response's compact path: 'vertices' (source
        first) and 'edge_ids' (empty when unreached)."""
        buf = (Response * capacity)()
        verts = np.zeros(vertex_capacity, np.uint32)
        edges = np.zeros(vertex_capacity, np.uint32)
        offs = np.zeros(capacity + 1, np.uint64)
        n = C.c_uint32()
        _check(hip_lib().bfn_server_poll_paths(self.h, buf, capacity, C.byref(n), timeout_ms,
                                               _ptr(verts), _ptr(edges), vertex_capacity,
                                               _ptr(offs)))
        out = []
        for i in range(n.value):
            a, b = int(offs[i]), int(offs[i + 1])
            out.append(dict(request_id=buf[i].request_id, distance=buf[i].distance,
                            path_length=buf[i].path_length, reached=buf[i].reached,
                            submit_ns=buf[i].submit_ns, complete_ns=buf[i].complete_ns,
                            vertices=verts[a:b].copy(), edge_ids=edges[a:max(a, b - 1)].copy()))
        return out

    def stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(hip_lib().bfn_server_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(launches=a.value, admitted=b.value, completed=c.value)

    def close(self):
        if self.h:
            hip_lib().bfn_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return bool(torch.cuda.is_available())
    except Exception:
        return False
