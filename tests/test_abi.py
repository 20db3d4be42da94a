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
