"""The C ABI boundary: libraries load, every declared entry point is exported,
record layouts keep the reference's sizes, and errors surface as status codes
(no GPU compute here)."""
import ctypes as C
import os
import re

import bfnew_amd as bfn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bfnh?_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    hip = C.CDLL(bfn.HIP_LIB_PATH, mode=C.RTLD_GLOBAL)
    host = C.CDLL(bfn.HOST_LIB_PATH)
    amd = declared("bfnew_amd.h")
    hosts = declared("bfnew_host.h")
    assert "bfn_batch_submit" in amd and "bfnh_graph_build" in hosts
    missing = [n for n in amd if not hasattr(hip, n)] + [n for n in hosts if not hasattr(host, n)]
    assert not missing, missing


def test_record_sizes_match_reference_abi():
    # gpu_api.hpp:240-251, compact_path_types.hpp:44-47
    assert C.sizeof(bfn.RunOptions) == 48
    assert C.sizeof(bfn.RunStatus) == 48
    assert C.sizeof(bfn.WorkStatistics) == 168
    assert C.sizeof(bfn.TargetSummary) == 28
    assert bfn.RunOptions.maximum_rounds.offset == 32
    assert bfn.RunOptions.enable_per_lane_convergence.offset == 40
    assert bfn.RunStatus.rounds_completed.offset == 8
    assert bfn.RunStatus.reached_target_mask.offset == 16
    assert bfn.RunStatus.error_bits.offset == 40
    assert bfn.WorkStatistics.high_contention_destinations.offset == 128


def test_no_device_reports_hip_error():
    if bfn.gpu_available():
        return
    try:
        bfn.Device(0)
    except bfn.BfnError as e:
        assert e.code == 2
    else:
        raise AssertionError("device open succeeded without a GPU")


def test_tuning_defaults():
    t = bfn.tuning()
    assert t.tile_sweeps == 1 and t.items_per_claim == 1 and t.use_graph == 0


def test_null_arguments_are_rejected():
    lib = bfn.hip_lib()
    assert lib.bfn_graph_upload(None, None, None) == 1
    assert lib.bfn_batch_wait(None) == 5
    assert b"no batch" in lib.bfn_last_error()


def _small_layout():
    hg = bfn.HostGraph.synthetic(4, 4, 16, 9)
    return hg.layout_arrays()


def test_validate_accepts_builder_layout():
    rc, msg = bfn.validate_layout(bfn.layout_from_arrays(_small_layout()))
    assert rc == 0, msg


def test_validate_rejects_bad_tile_runs():
    """The reference's tile-run checks (device_layout.cpp:280-322): a run tile
    out of range, a run naming a tile that is not its edges' endpoint tile,
    an empty run and negative weights are refused before any device call."""
    base = _small_layout()
    T = base["tile_count"]
    cases = []
    a = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in base.items()}
    a["csr_run_destination_tiles"][3] = T + 5
    cases.append((a, "out of range"))
    a = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in base.items()}
    r = a["csc_run_source_tiles"]
    r[0] = (r[0] + 1) % T
    cases.append((a, "run's tile"))
    a = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in base.items()}
    # collapse the first CSR run of vertex 0 onto its start: it becomes empty
    ro, re_ = a["csr_row_run_offsets"], a["csr_run_edge_offsets"]
    assert ro[1] - ro[0] >= 2
    re_[1] = re_[0]
    cases.append((a, "run"))
    a = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in base.items()}
    a["csc_weights"][7] = -1.0
    cases.append((a, "nonnegative"))
    for arrays, needle in cases:
        rc, msg = bfn.validate_layout(bfn.layout_from_arrays(arrays))
        assert rc == 1 and needle in msg, (rc, msg)
