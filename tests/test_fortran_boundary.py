"""The Fortran host side: the GPU variants of module velocity and module
transform.

CPU: they compile against the reference's own parameters/variables/modes
modules (when oracle/_ref has them) and against the host harness.
GPU: the driver with main.f90's loop order runs the M4 64^2 case for 100
steps; vel_energy.dat (printed '(2e20.12)') and the final state match the
oracle, and the transform variant's tra_spec2phys / tra_phys2spec match the
oracle's transforms.
"""
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "mwf_amd", "fortran")
DRIVER = os.path.join(FDIR, "_build", "mwf_drive")


def _have(exe):
    return shutil.which(exe) is not None


@pytest.mark.skipif(not _have("amdflang"), reason="amdflang absent")
def test_variant_compiles_against_host_harness(tmp_path):
    for src in ("host/host_modules.f90", "mwf_capi.f90", "velocity.mwf.f90", "transform.mwf.f90"):
        subprocess.run(["amdflang", "-c", os.path.join(FDIR, src), "-J", str(tmp_path)], cwd=tmp_path, check=True)


@pytest.mark.skipif(not (_have("amdflang") and os.path.exists(os.path.join(ROOT, "oracle", "_ref", "variables.mod"))),
                    reason="reference modules not built here")
def test_variant_compiles_against_reference_modules(tmp_path):
    subprocess.run(["amdflang", "-c", os.path.join(FDIR, "mwf_capi.f90"), "-J", str(tmp_path)], cwd=tmp_path, check=True)
    for src in ("velocity.mwf.f90", "transform.mwf.f90"):
        subprocess.run(["amdflang", "-c", os.path.join(FDIR, src), "-I", os.path.join(ROOT, "oracle", "_ref"),
                        "-J", str(tmp_path)], cwd=tmp_path, check=True)
    assert os.path.exists(tmp_path / "velocity.mod") and os.path.exists(tmp_path / "transform.mod")


@pytest.mark.gpu
def test_fortran_driver_matches_oracle(tmp_path):
    import mwf_amd as mw
    import oracle
    assert os.path.exists(DRIVER), "build with: make -C mwf_amd/fortran"
    nsteps = 100
    print(f"[fortran] mwf_drive {nsteps} steps, M4 64^2 modes, seed 1", flush=True)
    subprocess.run([DRIVER, str(nsteps), "1", "0.1", "4.0"], cwd=tmp_path, check=True, timeout=120)
    got = np.loadtxt(tmp_path / "vel_energy.dat")
    p = mw.make_params(4, 32, 32, 8 * math.pi, 8 * math.pi, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 1, 0.1, 4.0)
    o = oracle.Oracle(4, 32, 32, p.alpha, p.gamma, 200.0, 0.02)
    o.set_state(re, im)
    want = []
    for s in range(nsteps):
        want.append(o.energy())
        o.step(1)
    want = np.array(want)
    assert got.shape == (nsteps, 2)
    np.testing.assert_allclose(got[:, 0], 0.02 * np.arange(nsteps), rtol=0, atol=1e-10)
    # '(2e20.12)' prints a 12-digit mantissa 0.ddd...: half a printed unit is
    # 0.5e-12 * 10**(exponent), up to 5e-12 relative (program/io.f90:460)
    half_ulp = 0.5e-12 * 10.0 ** (np.floor(np.log10(want)) + 1)
    err = np.abs(got[:, 1] - want)
    print(f"[fortran] worst energy error {np.max(err / half_ulp):.2f} printed half-units, "
          f"{np.max(err / want):.2e} relative")
    assert np.all(err <= half_ulp * 1.05 + 1e-14 * want)
    fin = np.fromfile(tmp_path / "final_state.bin", dtype=np.float64)
    K, M, NN = 7, 63, 32
    fre, fim = fin[:K * M * NN].reshape(NN, M, K), fin[K * M * NN:].reshape(NN, M, K)
    ore, oim = o.get_state()
    err = math.sqrt((np.sum((fre - ore) ** 2) + np.sum((fim - oim) ** 2)) / (np.sum(ore ** 2) + np.sum(oim ** 2)))
    print(f"[fortran] final state rel L2 {err:.2e}")
    assert err <= 1e-12

    # the device diagnostics and the AB2 history through the Fortran bindings:
    # a Python handle on the same final state gives the same numbers
    d = np.fromfile(tmp_path / "diag.bin", dtype=np.float64)
    v = mw.Velocity(p)
    v.set_state(fre, fim)
    sizes = [1, 32, 32, 4 * 32, 3 * 16 * 16, K * M * NN, K * M * NN, 1]
    parts = np.split(d, np.cumsum(sizes)[:-1])
    assert d.size == sum(sizes)
    assert abs(parts[0][0] - o.energy()) <= 1e-12 * o.energy()
    m_max, n_max = v.spectrum_max()
    assert np.array_equal(parts[1], m_max) and np.array_equal(parts[2], n_max)
    assert np.array_equal(parts[3], v.vel_history(32, 0.25).ravel())
    assert np.array_equal(parts[4], v.turb_sample(6, 6).ravel())
    w = mw.Velocity(p)
    w.set_state(re, im)
    w.vel_TS(nsteps)
    hre, him, first = w.get_onl()
    # the Fortran variant tunes its plan (mwf_autotune), the handle here uses the
    # rule-based one: equal to rounding, not necessarily bit for bit
    herr = math.sqrt((np.sum((parts[5].reshape(NN, M, K) - hre) ** 2) + np.sum((parts[6].reshape(NN, M, K) - him) ** 2))
                     / (np.sum(hre ** 2) + np.sum(him ** 2)))
    assert herr <= 1e-12, herr
    assert parts[7][0] == 0.0 and not first
    print("[fortran] device diagnostics and vel_onl match the Python handle")

    # the transform variant (tra_spec2phys, tra_phys2spec) against the oracle
    KK, P, Q = 11, 96, 96
    t = np.fromfile(tmp_path / "transform.bin", dtype=np.float64)
    assert t.size == KK * Q * P + 2 * KK * M * NN
    tp = t[:KK * Q * P].reshape(P, Q, KK)
    sre = t[KK * Q * P:KK * Q * P + KK * M * NN].reshape(NN, M, KK)
    sim = t[KK * Q * P + KK * M * NN:].reshape(NN, M, KK)
    n_, m_, k_ = np.meshgrid(np.arange(NN), np.arange(M), np.arange(1, KK + 1), indexing="ij")
    ire, iim = np.sin(k_ + 3.0 * m_ + 7.0 * n_), np.cos(2.0 * k_ - m_ + 5.0 * n_)
    want = o.spec2phys(ire, iim)
    assert np.abs(tp - want).max() <= 1e-13 * np.abs(want).max()
    wr, wi = o.phys2spec(want)
    assert np.abs(sre - wr).max() <= 1e-13 * np.abs(wr).max() and np.abs(sim - wi).max() <= 1e-13 * np.abs(wi).max()
    print("[fortran] transform variant matches the oracle")
