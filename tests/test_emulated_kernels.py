"""The kernels' logic on the CPU.

bf-new_amd/csrc/hip is compiled for the host against a fiber emulation of
the HIP subset it uses (tests/emu: lanes are fibers, workgroups share LDS and
barriers, cooperative launches run every workgroup together, atomics may
switch lanes).  The C++ API driver's GPU mode -- every batched engine, every
control mode including the persistent grid barrier, compact paths, the
single-query engines and the expansion runner -- then runs on that library
and its output is checked against the oracle, as on a GPU.  This checks the
kernels' algorithm and the host/device protocol; it does not check the GPU
memory model, occupancy or code generation.
"""
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))

import build_emu  # noqa: E402
from test_cpp_api import _run, check_api_output  # noqa: E402


@pytest.mark.parametrize("yield_pct", [0, 15])
def test_emulated_engines_match_oracle(tmp_path, yield_pct):
    lib = build_emu.build()
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.dirname(lib) + (
        ":" + env["LD_LIBRARY_PATH"] if env.get("LD_LIBRARY_PATH") else "")
    env.update(EMU_CUS="3", EMU_OCC="2", EMU_YIELD_PCT=str(yield_pct))
    out = tmp_path / "api.bin"
    p = _run("gpu", str(out), timeout=600, env=env)
    assert "gpu: 0 failure(s)" in p.stdout
    check_api_output(out)
