"""Build libmwf.so (HIP, gfx950) in-tree.

    python -m mwf_amd.build        # or: from mwf_amd.build import build; build()

The library is compiled by hipcc straight from mwf_amd/csrc; the Galerkin
header is regenerated first.  The kernels are split into translation units by
precision (x kernels) and by precision and wall-normal basis (z kernels),
compiled in parallel into mwf_amd/_build and linked into mwf_amd/libmwf.so.
Packed FP32 instructions are disabled for all device code
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmwf.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-lambda-capture",
         "-Wno-unknown-pragmas"]

# (object, source, defines)
UNITS = [("mwf.o", "mwf.hip", []),
         ("ic.o", "ic.cpp", []),
         ("plan_x_f64.o", "plan_x.hip", ["-DMWF_R=double"]),
         ("plan_x_f32.o", "plan_x.hip", ["-DMWF_R=float"]),
         ("plan_z_f64_m4.o", "plan_z.hip", ["-DMWF_R=double", "-DMWF_K0=4"]),
         ("plan_z_f64_m6.o", "plan_z.hip", ["-DMWF_R=double", "-DMWF_K0=6"]),
         ("plan_z_f32_m4.o", "plan_z.hip", ["-DMWF_R=float", "-DMWF_K0=4"]),
         ("plan_z_f32_m6.o", "plan_z.hip", ["-DMWF_R=float", "-DMWF_K0=6"])]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-FP32 feature
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if err.strip():
        print(err, file=sys.stderr, flush=True)
    if r.returncode:
        raise subprocess.CalledProcessError(r.returncode, cmd)


def build(verbose=False, jobs=None):
    sys.path.insert(0, CSRC)
    try:
        import gen_galerkin
        gen_galerkin.main()
    finally:
        sys.path.pop(0)
    os.makedirs(OBJ, exist_ok=True)
    cmds = [[HIPCC] + FLAGS + d + ["-c", "-o", os.path.join(OBJ, o), os.path.join(CSRC, s)] for o, s, d in UNITS]
    jobs = jobs or min(len(cmds), os.cpu_count() or 1, 8)
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] +
         [os.path.join(OBJ, o) for o, _, _ in UNITS], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
