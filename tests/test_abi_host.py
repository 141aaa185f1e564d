"""CPU tests of the C-ABI library and its host-side parts (no GPU calls)."""
import ctypes
import math
import os
import re
import sys

import numpy as np
import pytest

import mwf_amd as mw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mwf_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = mw.lib()
    names = declared_symbols()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n


def test_params_struct_matches_header():
    # six int32 then eight doubles, no padding
    assert ctypes.sizeof(mw.Params) == 6 * 4 + 8 * 8
    text = open(os.path.join(ROOT, "include", "mwf.h")).read()
    body = text[text.index("typedef struct mwf_params"):text.index("} mwf_params;")]
    fields = re.findall(r"(?:int32_t|double)\s+(\w+);", body)
    assert fields == [f[0] for f in mw.Params._fields_]


def test_create_rejects_bad_sizes_without_gpu_work():
    for bad in (dict(k0=5), dict(mm=48), dict(nn=2048), dict(precision=7)):
        kw = dict(k0=4, mm=32, nn=32, lx=8 * math.pi, lz=8 * math.pi, re=200.0, dt=0.02)
        prec = bad.pop("precision", mw.MWF_FP64)
        kw.update(bad)
        p = mw.make_params(precision=prec, **kw)
        h = ctypes.c_void_p()
        rc = mw.lib().mwf_create(ctypes.byref(p), ctypes.byref(h))
        assert rc == -1 and not h.value
        assert mw.lib().mwf_last_error().decode()


def _energy_parseval(p, re, im):
    K0 = p.k0
    NN, M, K = re.shape
    MM = (M + 1) // 2
    c = re + 1j * im
    m = np.where(np.arange(M) < MM, np.arange(M), np.arange(M) - M).astype(float)
    n = np.arange(NN).astype(float)
    kx, kz = p.alpha * m[None, :], p.gamma * n[:, None]
    E = 0.0
    for k in range(K0):
        psi = c[:, :, k]
        phi = c[:, :, K0 - 1 + k] if k > 0 else 0 * psi
        D = (-1) ** k * k * math.pi / 2
        u = -1j * kz * psi + 1j * kx * D * phi
        w = 1j * kx * psi + 1j * kz * D * phi
        v = (kx ** 2 + kz ** 2) * phi
        t = np.abs(u) ** 2 + np.abs(w) ** 2 + (np.abs(v) ** 2 if k > 0 else 0)
        E += (1.0 if k == 0 else 0.5) * (t[0].sum() + 2 * t[1:].sum())
    return E


@pytest.mark.parametrize("k0,mm,nn,L,scl,kic", [(4, 32, 32, 8 * math.pi, 0.1, 4.0),
                                               (6, 16, 32, 6.0, 0.3, 2.0),
                                               (4, 64, 64, 16 * math.pi, 0.2, 8.0)])
def test_ic_gaussian_properties(k0, mm, nn, L, scl, kic):
    p = mw.make_params(k0, mm, nn, L, L, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 1234, scl, kic)
    re2, im2 = mw.ic_gaussian(p, 1234, scl, kic)
    assert np.array_equal(re, re2) and np.array_equal(im, im2)          # deterministic
    re3, _ = mw.ic_gaussian(p, 1235, scl, kic)
    assert not np.array_equal(re, re3)                                   # seeded
    M = 2 * mm - 1
    assert np.all(re[0, 0] == 0) and np.all(im[0, 0] == 0)               # laminar mean
    for m in range(1, mm):                                               # Hermitian n = 0
        assert np.array_equal(re[0, M - m], re[0, m]) and np.array_equal(im[0, M - m], -im[0, m])
    assert _energy_parseval(p, re, im) == pytest.approx(scl * scl, rel=1e-12)   # rms = scl


def test_generated_galerkin_equals_oracle():
    import oracle
    sys.path.insert(0, os.path.join(ROOT, "mwf_amd", "csrc"))
    try:
        import gen_galerkin as G
    finally:
        sys.path.pop(0)
    rng = np.random.default_rng(1)
    for k0 in (4, 6):
        st, y = G.stream_terms(k0)
        kk = 3 * k0 - 1
        for _ in range(50):
            v, vx, vz = rng.standard_normal((3, kk))
            a = np.zeros(kk)
            for s in range(kk):
                for o, b, c in st["x"][s]:
                    a[o] += float(c) * v[b] * vx[s]
                for o, b, c in st["z"][s]:
                    a[o] += float(c) * v[b] * vz[s]
            for o, j, l, c in y:
                a[o] += float(c) * oracle.BETA * v[j] * v[l]
            np.testing.assert_allclose(a, oracle.udotgradu(k0, v, vx, vz), rtol=0, atol=5e-14)
    # the committed header is what the generator writes
    hdr = os.path.join(ROOT, "mwf_amd", "csrc", "galerkin_gen.hpp")
    assert open(hdr).read() == G.emit()


def test_bench_refuses_multi_gpu():
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
