"""The serving loop (bfn_server_*) on the CPU emulation of the kernels.

tests/emu/server_driver.py runs the server of the emulated engine library
(tests/emu: lanes are fibers, atomics may switch lanes, device memory
starts as garbage) with more requests than active places, so requests are
admitted while others are in flight and rounds advance across chunks, and
checks every response against the oracle.  This covers the server's own
control-block set-up and its round protocol, which the batch API tests do
not reach.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))

import build_emu  # noqa: E402


@pytest.mark.parametrize("engine", [2, 0])
def test_emulated_server_matches_oracle(engine):
    lib = build_emu.build()
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.dirname(lib) + (
        ":" + env["LD_LIBRARY_PATH"] if env.get("LD_LIBRARY_PATH") else "")
    env.update(EMU_CUS="3", EMU_OCC="2", EMU_YIELD_PCT="15")
    p = subprocess.run([sys.executable, os.path.join(HERE, "emu", "server_driver.py"), lib,
                        str(engine)], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    out = json.loads(p.stdout.strip().splitlines()[-1])
    assert out["failures"] == [] and out["responses"] == 6, out
