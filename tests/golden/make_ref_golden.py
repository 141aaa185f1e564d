"""Write tests/golden/ref_m4_kat.npz from the compiled reference.

Builds oracle/_ref/ref_kat (the reference's own mpi.f90, parameters.f90,
modes.M4.f90 and variables.f90 compiled where they lie, plus the driver
oracle/ref_kat.f90), feeds it the SplitMix64 inputs of katio.py and stores
its outputs.  Run from the repository root:  python tests/golden/make_ref_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import katio  # noqa: E402


def run_reference():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_kat")
    with tempfile.TemporaryDirectory() as td:
        katio.write_inputs(os.path.join(td, "kat_in.bin"))
        subprocess.run([exe], cwd=td, check=True)
        return katio.read_outputs(os.path.join(td, "kat_out.bin"))


def main():
    out = run_reference()
    keep = {}
    for k, v in out.items():
        if v.ndim == 3:
            keep[k] = v[:katio.NSUB].copy()
        else:
            keep[k] = v
    path = os.path.join(HERE, "ref_m4_kat.npz")
    np.savez_compressed(path, **keep)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
