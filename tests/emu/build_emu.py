"""Builds tests/emu/_build/libbfnew_hip.so: the sources of bf-new_amd/csrc/hip
compiled for the host against the HIP emulation of this directory.

Copies of the sources are compiled, with two mechanical substitutions: the
dynamic LDS array becomes the emulated workgroup's LDS, and cooperative
launches through a kernel pointer call the kernel template directly.
Rebuilds only when a source is newer than the library.
"""
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "bf-new_amd", "csrc", "hip")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libbfnew_hip.so")


def _inputs():
    return (glob.glob(os.path.join(SRC, "*")) + glob.glob(os.path.join(HERE, "*.cpp")) +
            glob.glob(os.path.join(HERE, "shim", "hip", "*.h")) + [__file__] +
            [os.path.join(ROOT, "include", "bfnew_amd.h")])


def build(cxx="g++", opt="-O1"):
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(f) for f in _inputs()):
        return LIB
    src = os.path.join(OUT, "src")
    obj = os.path.join(OUT, "obj")
    os.makedirs(src, exist_ok=True)
    os.makedirs(obj, exist_ok=True)
    units = []
    for f in sorted(glob.glob(os.path.join(SRC, "*"))):
        s = open(f).read()
        s = s.replace("extern __shared__ std::uint32_t lds_words[];",
                      "std::uint32_t* lds_words = emu::block_lds();")
        m = re.search(r"const void\* kernel = reinterpret_cast<const void\*>\((&relax_persistent_kernel<[^>]*>)\);", s)
        if m:
            fn = m.group(1)
            s = re.sub(r"return hipLaunchCooperativeKernel\(kernel, dim3\(grid\), dim3\(B\), args,\s*"
                       r"static_cast<unsigned>\(lds_bytes\), stream\);",
                       "return emu_launch_ptr(%s, dim3(grid), dim3(B), args, lds_bytes, stream, true);" % fn, s)
            s = s.replace("return hipLaunchKernel(kernel, dim3(grid), dim3(B), args, lds_bytes, stream);",
                          "return emu_launch_ptr(%s, dim3(grid), dim3(B), args, lds_bytes, stream, true);" % fn)
        dst = os.path.join(src, os.path.basename(f))
        with open(dst, "w") as out:
            out.write(s)
        if dst.endswith(".hip"):
            units.append(dst)
    units.append(os.path.join(HERE, "emu_rt.cpp"))
    flags = ["-std=c++20", opt, "-g", "-fPIC", "-pthread", "-ffp-contract=off", "-w",
             "-I" + os.path.join(HERE, "shim"), "-I" + os.path.join(ROOT, "include"), "-I" + src]
    procs, objs = [], []
    for u in units:
        o = os.path.join(obj, os.path.splitext(os.path.basename(u))[0] + ".o")
        objs.append(o)
        procs.append(subprocess.Popen([cxx, *flags, "-x", "c++", "-c", u, "-o", o]))
    if any(p.wait() != 0 for p in procs):
        raise RuntimeError("emulated HIP build failed")
    subprocess.run([cxx, "-shared", "-pthread", "-o", LIB, *objs], check=True)
    return LIB


if __name__ == "__main__":
    print(build(*sys.argv[1:]))
