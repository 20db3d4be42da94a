"""Reachability expansion over batched queries (tests/cpp/expansion_test.cpp).

CPU: run_batched_expansion over a host batch runner (fp32 Dijkstra restricted
to each lane's tiles) on the seven-route fixture of the reference's
tests/batched_expansion_fixture_suite.hpp: the retry count of every route
under each of the four schedules (the reference test's literals), the
full-region fallback, the expansion limit, region stalls, the spill tile's
entry through real adjacency, identity limits, measured retry work, the
runner contract, option validation, evidence-only schedule selection, and the
compact payload rules.

GPU: the same orchestrator over hip::BatchedExpansionExecutor for every
engine, schedule and transfer mode; outcomes, traces and compact payloads
equal the host runner's.  The emulated variant runs that GPU driver on the
CPU emulation of the HIP subset (tests/emu).
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "bf-new_amd", "lib", "bfnew_expansion_test")


def _run(mode, timeout=120, env=None):
    if not os.path.exists(DRIVER):
        pytest.fail("bf-new_amd/lib/bfnew_expansion_test is not built (run __graft_entry__.build())")
    p = subprocess.run([DRIVER, mode], capture_output=True, text=True, timeout=timeout, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "%s: 0 failure(s)" % mode in p.stdout, p.stdout + p.stderr
    return p


def test_expansion_orchestrator_host_runner():
    _run("host")


def test_expansion_orchestrator_emulated_executor():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    lib = build_emu.build()
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.dirname(lib) + (
        ":" + env["LD_LIBRARY_PATH"] if env.get("LD_LIBRARY_PATH") else "")
    env.update(EMU_CUS="2", EMU_OCC="2", EMU_YIELD_PCT="10")
    _run("gpu", timeout=900, env=env)


@pytest.mark.gpu
def test_expansion_orchestrator_executor():
    _run("gpu")


@pytest.mark.parametrize("runner", ["batch", "pass"])
def test_expansion_bench_driver_emulated(runner):
    """The bench driver (bench.py --expansion) end to end on the CPU emulation,
    at a size the emulation finishes in seconds: every query is reached, the
    misses of the terminal-tiles start expand once, and the records add up."""
    import json
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    lib = build_emu.build()
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.dirname(lib) + (
        ":" + env["LD_LIBRARY_PATH"] if env.get("LD_LIBRARY_PATH") else "")
    env.update(EMU_CUS="2", EMU_OCC="2", EMU_YIELD_PCT="0")
    driver = os.path.join(ROOT, "bf-new_amd", "lib", "bfnew_expansion_bench")
    p = subprocess.run([driver, "--tiles", "8", "8", "16", "--queries", "24", "--lane-width", "8",
                        "--reps", "1", "--runner", runner], capture_output=True, text=True, timeout=900,
                       env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    out = json.loads(p.stdout.strip().splitlines()[-1])
    assert out["workload"] == "expansion_8x8x16_q24" and len(out["schedules"]) == 4
    for row in out["schedules"]:
        assert row["reached"] == 24 and row["other_failures"] == 0
        assert sum(row["histogram"]) == 24
        assert row["scheduled_expansions"] + row["full_region_fallbacks"] == sum(
            i * n for i, n in enumerate(row["histogram"]))
        assert row["initial_reached"] + sum(row["histogram"][1:]) == 24
        assert row["path_vertices"] > 0 and row["compact_d2h_bytes"] > 0
