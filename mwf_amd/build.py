"""Build libmwf.so (HIP, gfx950) in-tree.

    python -m mwf_amd.build        # or: from mwf_amd.build import build; build()

The library is compiled by hipcc straight from mwf_amd/csrc; the Galerkin
header is regenerated first.  Packed FP32 instructions are disabled for all
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmwf.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["mwf.hip", "ic.cpp"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture",
         "-Wno-unknown-pragmas"]


def build(verbose=False):
    sys.path.insert(0, CSRC)
    try:
        import gen_galerkin
        gen_galerkin.main()
    finally:
        sys.path.pop(0)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
