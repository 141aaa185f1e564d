"""CPU tests that pin the oracle (the C restatement in oracle/).

Pins, per function:
* LHSRHS, mpt2spec, spec_grad, spec2mpt: the compiled reference
  (tests/golden/ref_m4_kat.npz, written by tests/golden/make_ref_golden.py
  from the reference's own variables.f90/modes.M4.f90; regenerated and
  compared live when oracle/_ref builds here).
* Galerkin closures (M4 and M6): the compiled reference (M4) and exact
  y-quadrature of the basis products (M4, M6).
* DFT / spec2phys / phys2spec: brute-force DFT, numpy.fft (pocketfft), and the
  band-limited round trip.
* energy: spectral Parseval sum.
* vel_TS: laminar fixed point, mean-mode decay closed form, AB2 update from
  the reference's own lhs/rhs, spec2mpt(mpt2spec(c)) = -rhs * c.
"""
import math
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import katio  # noqa: E402

GOLD = os.path.join(HERE, "golden", "ref_m4_kat.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


@pytest.fixture(scope="module")
def ref_oracle(gold):
    h = gold["hdr"]
    return oracle.Oracle(int(h[0]), int(h[1]), int(h[2]), h[3], h[4], h[6], h[5], drag=h[7],
                         vdrag=h[8], hypo=h[9], dragall=bool(h[11]))


# ---------------------------------------------------------------- reference KAT

def test_fixture_header(gold):
    h = gold["hdr"]
    assert (h[0], h[1], h[2]) == (4, 32, 32)
    assert h[3] == pytest.approx(2 * math.pi / 12) and h[4] == pytest.approx(2 * math.pi / 6)
    assert h[10] == pytest.approx(math.pi / 2)


def test_lhsrhs_matches_reference(gold, ref_oracle):
    l, r = ref_oracle.lhsrhs()
    ns = katio.NSUB
    assert np.array_equal(l[:ns], gold["lhs"])
    assert np.array_equal(r[:ns], gold["rhs"])


def test_mpt2spec_specgrad_spec2mpt_match_reference(gold, ref_oracle):
    Are, Aim, Sre, Sim, V, VX, VZ = katio.inputs()
    ns = katio.NSUB
    sr, si = ref_oracle.mpt2spec(Are, Aim)
    assert np.array_equal(sr[:ns], gold["m2s_re"]) and np.array_equal(si[:ns], gold["m2s_im"])
    g = ref_oracle.spec_grad(Sre, Sim)
    for a, k in zip(g, ["ux_re", "ux_im", "uz_re", "uz_im"]):
        assert np.array_equal(a[:ns], gold[k])
    mr, mi = ref_oracle.spec2mpt(Sre, Sim)
    assert np.array_equal(mr[:ns], gold["s2m_re"]) and np.array_equal(mi[:ns], gold["s2m_im"])


def test_galerkin_matches_reference(gold):
    _, _, _, _, V, VX, VZ = katio.inputs()
    K0 = 4
    for j in range(katio.NV):
        v, vx, vz = V[j], VX[j], VZ[j]
        np.testing.assert_allclose(oracle.udotgradu(K0, v, vx, vz), gold["udotgradu"][j], rtol=0, atol=4e-15)
        np.testing.assert_allclose(oracle.nluw(K0, v[:4], vx[:4]), gold["nluw"][j], rtol=0, atol=1e-15)
        np.testing.assert_allclose(oracle.nlvv(K0, v[4:7]), gold["nlvv"][j], rtol=0, atol=1e-15)
        np.testing.assert_allclose(oracle.nluv(K0, v[:4], vx[4:7]), gold["nluv"][j], rtol=0, atol=1e-15)
        np.testing.assert_allclose(oracle.nluyv(K0, v[:4], v[4:7]), gold["nluyv"][j], rtol=0, atol=2e-15)
        np.testing.assert_allclose(oracle.nlvvy(K0, v[4:7]), gold["nlvvy"][j], rtol=0, atol=2e-15)
    # the hand-expanded M4 closure equals the composed one in the reference itself
    np.testing.assert_allclose(gold["udotgradu"], gold["udotgradu2"], rtol=0, atol=4e-15)


@pytest.mark.skipif(not os.path.isdir("/root/reference/program"), reason="reference sources absent")
def test_fixture_regenerates_from_reference():
    subprocess.run(["make", "-s", "-C", os.path.join(os.path.dirname(HERE), "oracle"), "ref"], check=True)
    exe = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_kat")
    gold = dict(np.load(GOLD))
    with tempfile.TemporaryDirectory() as td:
        katio.write_inputs(os.path.join(td, "kat_in.bin"))
        subprocess.run([exe], cwd=td, check=True)
        out = katio.read_outputs(os.path.join(td, "kat_out.bin"))
    for k, v in gold.items():
        live = out[k][:katio.NSUB] if out[k].ndim == 3 else out[k]
        assert np.array_equal(live, v), k


# ---------------------------------------------------------------- quadrature

def _basis(k0, y, beta=math.pi / 2):
    """h_k (u/w family, k=0..K0-1), g_k (v family, k=1..K0-1) and d/dy of each."""
    h = np.array([np.cos(k * beta * y) if k % 2 == 0 else np.sin(k * beta * y) for k in range(k0)])
    dh = np.array([-k * beta * np.sin(k * beta * y) if k % 2 == 0 else k * beta * np.cos(k * beta * y)
                   for k in range(k0)])
    g = np.array([np.cos(k * beta * y) if k % 2 == 1 else np.sin(k * beta * y) for k in range(1, k0)])
    dg = np.array([-k * beta * np.sin(k * beta * y) if k % 2 == 1 else k * beta * np.cos(k * beta * y)
                   for k in range(1, k0)])
    return h, dh, g, dg


def _proj(f, B, y):
    """Coefficients of f on the rows of B by exact periodic quadrature."""
    nrm = (B * B).mean(axis=1)
    return (B * f[None, :]).mean(axis=1) / nrm


@pytest.mark.parametrize("k0", [4, 6])
def test_galerkin_matches_quadrature(k0):
    rng = np.random.default_rng(k0)
    y = np.arange(256) * (4.0 / 256)  # one period of sin(beta y), beta = pi/2
    h, dh, g, dg = _basis(k0, y)
    kk = 3 * k0 - 1
    for _ in range(20):
        v, vx, vz = rng.standard_normal((3, kk))
        U, Vv, W = v[:k0] @ h, v[k0:2 * k0 - 1] @ g, v[2 * k0 - 1:] @ h
        dU, dV, dW = v[:k0] @ dh, v[k0:2 * k0 - 1] @ dg, v[2 * k0 - 1:] @ dh
        Ux, Vx, Wx = vx[:k0] @ h, vx[k0:2 * k0 - 1] @ g, vx[2 * k0 - 1:] @ h
        Uz, Vz, Wz = vz[:k0] @ h, vz[k0:2 * k0 - 1] @ g, vz[2 * k0 - 1:] @ h
        ans = np.concatenate([_proj(U * Ux + Vv * dU + W * Uz, h, y),
                              _proj(U * Vx + Vv * dV + W * Vz, g, y),
                              _proj(U * Wx + Vv * dW + W * Wz, h, y)])
        np.testing.assert_allclose(oracle.udotgradu(k0, v, vx, vz), ans, rtol=0, atol=2e-13)
        np.testing.assert_allclose(oracle.nluw(k0, v[:k0], vx[:k0]), _proj(U * Ux, h, y), atol=1e-13)
        np.testing.assert_allclose(oracle.nlvv(k0, v[k0:2 * k0 - 1]), _proj(Vv * Vv, h, y), atol=1e-13)
        np.testing.assert_allclose(oracle.epos(k0, v), (U * U + Vv * Vv + W * W).mean(), atol=1e-13)


# ---------------------------------------------------------------- transforms

@pytest.mark.parametrize("n", [7, 10, 12, 45, 96, 192])
def test_dft_matches_bruteforce_and_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    for sign in (+1, -1):
        brute = np.exp(sign * 2j * np.pi * np.outer(j, j) / n) @ x
        got = oracle.dft(x, sign)
        np.testing.assert_allclose(got, brute, rtol=0, atol=1e-12 * np.abs(brute).max())
    np.testing.assert_allclose(oracle.dft(x, -1), np.fft.fft(x), rtol=0, atol=1e-13 * n)


def _numpy_spec2phys(o, sre, sim):
    """Independent spec2phys: numpy ifft in x (+1, unnormalised), irfft in z
    (numpy drops the imaginary part of the DC bin, as FFTW c2r does)."""
    NN, M, KK = sre.shape
    MM = o.mm
    s = sre + 1j * sim
    full = np.zeros((NN, o.P, KK), complex)
    for m in range(-(MM - 1), MM):
        full[:, m % o.P, :] = s[:, m % M, :]
    T = np.fft.ifft(full, axis=1) * o.P                   # (NN, P, KK)
    T = np.transpose(T, (1, 0, 2))                        # (P, NN, KK)
    pad = np.zeros((o.P, o.Q // 2 + 1, KK), complex)
    pad[:, :NN, :] = T
    return np.fft.irfft(pad, n=o.Q, axis=1) * o.Q         # (P, Q, KK)


@pytest.mark.parametrize("k0,mm,nn", [(4, 16, 16), (4, 32, 16), (6, 16, 32)])
def test_spec2phys_matches_numpy_and_roundtrip(k0, mm, nn):
    o = oracle.Oracle(k0, mm, nn, 0.5, 0.7, 200.0, 0.01)
    rng = np.random.default_rng(mm + nn)
    sre, sim = rng.standard_normal((2,) + o.spec_shape)
    ph = o.spec2phys(sre, sim)
    ref = _numpy_spec2phys(o, sre, sim)
    np.testing.assert_allclose(ph, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # round trip on a band-limited field with a Hermitian n = 0 column
    M = o.M
    for m in range(1, mm):
        sre[0, M - m, :] = sre[0, m, :]
        sim[0, M - m, :] = -sim[0, m, :]
    sim[0, 0, :] = 0.0
    br, bi = o.phys2spec(o.spec2phys(sre, sim))
    np.testing.assert_allclose(br, sre, rtol=0, atol=1e-13)
    np.testing.assert_allclose(bi, sim, rtol=0, atol=1e-13)


# ---------------------------------------------------------------- energy, step

def _parseval(o, re, im):
    c = re + 1j * im
    K0, MM, M = o.k0, o.mm, o.M
    m = np.where(np.arange(M) < MM, np.arange(M), np.arange(M) - M).astype(float)
    n = np.arange(o.nn).astype(float)
    a, g = o.params.alpha, o.params.gamma
    kx, kz = a * m[None, :], g * n[:, None]
    E = 0.0
    for k in range(K0):
        psi = c[:, :, k]
        phi = c[:, :, K0 - 1 + k] if k > 0 else 0 * psi
        D = (-1) ** k * k * math.pi / 2
        u = -1j * kz * psi + 1j * kx * D * phi
        w = 1j * kx * psi + 1j * kz * D * phi
        v = (kx ** 2 + kz ** 2) * phi if k > 0 else 0 * psi
        # mean modes at (m, n) = (0, 0): variables.f90:418-433
        u[0, 0] = psi[0, 0] if k > 0 else 0.0
        w[0, 0] = phi[0, 0] if k > 0 else 0.0
        for f in (u, v, w):
            if k > 0 or f is not v:
                herm = 0.5 * (f[0] + np.conj(np.roll(f[0][::-1], 1)))
                yw = 1.0 if k == 0 else 0.5
                E += yw * ((np.abs(herm) ** 2).sum() + 2 * (np.abs(f[1:]) ** 2).sum())
    return E


@pytest.mark.parametrize("k0", [4, 6])
def test_energy_matches_parseval(k0):
    o = oracle.Oracle(k0, 16, 16, 0.25, 0.3, 200.0, 0.02)
    rng = np.random.default_rng(3)
    re, im = 0.1 * rng.standard_normal((2,) + o.mpt_shape)  # non-Hermitian n=0 column on purpose
    o.set_state(re, im)
    assert o.energy() == pytest.approx(_parseval(o, re, im), rel=1e-13)


def test_laminar_fixed_point():
    o = oracle.Oracle(4, 16, 16, 0.25, 0.25, 200.0, 0.02)
    z = np.zeros(o.mpt_shape)
    o.set_state(z, z)
    o.step(5)
    re, im = o.get_state()
    assert np.all(re == 0) and np.all(im == 0)


def test_mean_mode_decay_closed_form():
    """A y-only perturbation has no nonlinear term: c_n = c_0 / tvl^n."""
    K0, dt, Re, drag = 4, 0.02, 200.0, 1e-2
    o = oracle.Oracle(K0, 16, 16, 0.25, 0.25, Re, dt, drag=drag)
    re, im = np.zeros(o.mpt_shape), np.zeros(o.mpt_shape)
    rng = np.random.default_rng(0)
    re[0, 0, 1:] = rng.standard_normal(o.K - 1)
    o.set_state(re, im)
    nsteps = 40
    o.step(nsteps)
    got, _ = o.get_state()
    beta = math.pi / 2
    for slot in range(1, o.K):
        k = slot if slot < K0 else slot - (K0 - 1)
        tvl = 1 - dt * ((-(k * k) * beta * beta) / Re - drag)
        assert got[0, 0, slot] == pytest.approx(re[0, 0, slot] / tvl ** nsteps, rel=1e-13)
    assert np.count_nonzero(got[:, 1:, :]) == 0 and np.count_nonzero(got[1:]) == 0


def test_ab2_update_with_reference_lhsrhs(gold, ref_oracle):
    """Two steps from a random state follow velocity.f90:24-40 with the
    reference's own lhs/rhs (forward Euler first, then AB2)."""
    o = ref_oracle
    rng = np.random.default_rng(5)
    re, im = 0.01 * rng.standard_normal((2,) + o.mpt_shape)
    o.set_state(re, im)
    nl0 = o.nonlinear()
    l, r = o.lhsrhs()
    dt = gold["hdr"][5]
    o.step(1)
    c1 = o.get_state()
    for part in range(2):
        want = (r * (re, im)[part] + dt / 2 * (3 * nl0[part] - nl0[part])) / l
        if part == 0:
            want[0, 0, 0] = 0.0
        np.testing.assert_allclose(c1[part], want, rtol=1e-14, atol=1e-17)
    nl1 = o.nonlinear()
    o.step(1)
    c2 = o.get_state()
    for part in range(2):
        want = (r * c1[part] + dt / 2 * (3 * nl1[part] - nl0[part])) / l
        if part == 0:
            want[0, 0, 0] = 0.0
        np.testing.assert_allclose(c2[part], want, rtol=1e-13, atol=1e-16)
    # lhs/rhs of this oracle are the reference's
    assert np.array_equal(l[:katio.NSUB], gold["lhs"])


def test_spec2mpt_mpt2spec_is_minus_rhs(ref_oracle):
    o = ref_oracle
    rng = np.random.default_rng(9)
    re, im = rng.standard_normal((2,) + o.mpt_shape)
    back = o.spec2mpt(*o.mpt2spec(re, im))
    _, r = o.lhsrhs()
    np.testing.assert_allclose(back[0], -r * re, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(back[1], -r * im, rtol=1e-12, atol=1e-9)
