"""CPU pin of the oracle's composed nonlinear term (vel_nonlinear,
program/velocity.f90:92-113) and, through it, of the oracle's step.

The oracle composes pieces that are each pinned elsewhere (mpt2spec, spec_grad,
spec2mpt and udotgradu against the compiled reference; the transforms against
brute force and numpy).  This file checks the composition itself against an
independent 3-D evaluation: the velocity field is built on an (x, y, z) grid --
x, z on the 3/2-padded grid by numpy's FFTs, y on a periodic quadrature grid
from the wall-normal bases -- with the laminar profile added in physical space,
(u.grad)u is formed pointwise with analytic derivatives, projected onto the
bases by exact quadrature and transformed back with numpy.  Only mpt2spec and
spec2mpt (bit-identical to the compiled reference) are shared with the oracle.

The 3-D evaluation is pinned in turn by the energy budget of the advection
operator: for a divergence-free perturbation u' on the laminar profile U_L(y),
<u', (u.grad)u> = <v', u' dU_L/dy + w' dW_L/dy> exactly (the other terms vanish
by incompressibility and periodicity), and every sum here is an exact integral
of a trigonometric polynomial.  A dropped term, a swapped derivative or a
wrong basis derivative breaks the budget.
"""
import math

import numpy as np
import pytest

import oracle

BETA = math.pi / 2
NY = 64  # y points over one period (length 4): exact for every product here


def _ybases(k0, y):
    """h_k (u/w family, k < K0), g_k (v family, 1 <= k < K0) and their d/dy."""
    h = np.array([np.cos(k * BETA * y) if k % 2 == 0 else np.sin(k * BETA * y) for k in range(k0)])
    dh = np.array([-k * BETA * np.sin(k * BETA * y) if k % 2 == 0 else k * BETA * np.cos(k * BETA * y)
                   for k in range(k0)])
    g = np.array([np.cos(k * BETA * y) if k % 2 == 1 else np.sin(k * BETA * y) for k in range(1, k0)])
    dg = np.array([-k * BETA * np.sin(k * BETA * y) if k % 2 == 1 else k * BETA * np.cos(k * BETA * y)
                   for k in range(1, k0)])
    return h, dh, g, dg


def _to_xz(o, s):
    """spec (NN, M, KK) complex -> physical (P, Q, KK): c2c (+1) in x, c2r in z
    (numpy's irfft ignores the imaginary part of the DC bin, as FFTW's c2r does)."""
    NN, M, KK = s.shape
    full = np.zeros((NN, o.P, KK), complex)
    for mi in range(M):
        m = mi if mi < o.mm else mi - M
        full[:, m % o.P, :] = s[:, mi, :]
    T = np.transpose(np.fft.ifft(full, axis=1) * o.P, (1, 0, 2))
    pad = np.zeros((o.P, o.Q // 2 + 1, KK), complex)
    pad[:, :NN, :] = T
    return np.fft.irfft(pad, n=o.Q, axis=1) * o.Q


def _from_xz(o, f):
    """physical (P, Q, KK) -> spec (NN, M, KK): scale 1/(PQ), r2c in z, c2c (-1)
    in x, truncated to |m| < MM, n < NN."""
    F = np.fft.fft(np.fft.rfft(f, axis=1), axis=0) / (o.P * o.Q)  # (P, Q/2+1, KK)
    out = np.zeros((o.nn, o.M, f.shape[2]), complex)
    for mi in range(o.M):
        m = mi if mi < o.mm else mi - o.M
        out[:, mi, :] = F[m % o.P, :o.nn, :]
    return out


def _wavenumbers(o):
    m = np.array([mi if mi < o.mm else mi - o.M for mi in range(o.M)], float)
    n = np.arange(o.nn, dtype=float)
    return o.params.alpha * m[None, :, None], o.params.gamma * n[:, None, None]


def composed_3d(o, re, im, theta):
    """Independent vel_nonlinear: returns (spec-form NL, budget lhs, budget rhs,
    mean magnitude of the budget's terms)."""
    k0 = o.k0
    sr, si = o.mpt2spec(re, im)
    s = sr + 1j * si
    kx, kz = _wavenumbers(o)
    p, px, pz = _to_xz(o, s), _to_xz(o, 1j * kx * s), _to_xz(o, 1j * kz * s)
    y = np.arange(NY) * (4.0 / NY)
    h, dh, g, dg = _ybases(k0, y)
    su, sv, sw = slice(0, k0), slice(k0, 2 * k0 - 1), slice(2 * k0 - 1, 3 * k0 - 1)

    def field(a, sl, B):
        return np.einsum("xzk,ky->xzy", a[:, :, sl], B)

    # perturbation and its derivatives on the (x, z, y) grid
    U, V, W = field(p, su, h), field(p, sv, g), field(p, sw, h)
    Ux, Vx, Wx = field(px, su, h), field(px, sv, g), field(px, sw, h)
    Uz, Vz, Wz = field(pz, su, h), field(pz, sv, g), field(pz, sw, h)
    Uy, Vy, Wy = field(p, su, dh), field(p, sv, dg), field(p, sw, dh)
    # laminar profile (vel_addlam, program/velocity.f90:76-82): cos/sin(theta) sin(beta y)
    UL = math.cos(theta) * np.sin(BETA * y)
    WL = math.sin(theta) * np.sin(BETA * y)
    dUL = math.cos(theta) * BETA * np.cos(BETA * y)
    dWL = math.sin(theta) * BETA * np.cos(BETA * y)
    Ut, Wt = U + UL, W + WL
    Uyt, Wyt = Uy + dUL, Wy + dWL
    NLu = Ut * Ux + V * Uyt + Wt * Uz
    NLv = Ut * Vx + V * Vy + Wt * Vz
    NLw = Ut * Wx + V * Wyt + Wt * Wz
    # energy budget of the advection operator
    lhs = (U * NLu + V * NLv + W * NLw).mean()
    rhs = (V * (U * dUL + W * dWL)).mean()
    size = (np.abs(U * NLu) + np.abs(V * NLv) + np.abs(W * NLw)).mean()
    # exact projections onto the bases (coefficient of each basis function)
    def proj(f, B):
        return np.einsum("xzy,ky->xzk", f, B) / NY / (B * B).mean(axis=1)
    phys = np.concatenate([proj(NLu, h), proj(NLv, g), proj(NLw, h)], axis=2)
    return _from_xz(o, phys), lhs, rhs, size


def _state(o, seed, herm=True):
    rng = np.random.default_rng(seed)
    re, im = 0.2 * rng.standard_normal((2,) + o.mpt_shape)
    if herm:  # Hermitian n = 0 column, as the initial conditions have
        for mi in range(1, o.mm):
            re[0, o.M - mi, :] = re[0, mi, :]
            im[0, o.M - mi, :] = -im[0, mi, :]
        im[0, 0, :] = 0.0
    re[0, 0, 0] = 0.0  # mpt slot (1,0,0) is zero after every step
    im[0, 0, 0] = 0.0
    return re, im


CASES = [  # k0, mm, nn, theta (rad), alpha, gamma
    (4, 8, 8, 0.0, 0.5, 0.7),
    (4, 8, 16, 0.3, 0.25, 0.4),
    (4, 16, 8, -0.7, 0.3, 0.2),
    (6, 8, 8, 0.2, 0.5, 0.6),
    (6, 16, 8, 1.1, 0.2, 0.35),
]


@pytest.mark.parametrize("k0,mm,nn,theta,alpha,gamma", CASES)
def test_energy_budget_of_3d_composition(k0, mm, nn, theta, alpha, gamma):
    o = oracle.Oracle(k0, mm, nn, alpha, gamma, 200.0, 0.02, theta=theta)
    re, im = _state(o, 11 * k0 + mm + nn)
    _, lhs, rhs, size = composed_3d(o, re, im, theta)
    # the budget's terms cancel down to rhs; rounding scales with their size
    assert abs(lhs - rhs) <= 1e-13 * size, (lhs, rhs, size)
    # the production term stands far above the tolerance, so losing it fails
    assert abs(rhs) > 1e-7 * size


@pytest.mark.parametrize("k0,mm,nn,theta,alpha,gamma", CASES)
@pytest.mark.parametrize("herm", [True, False])
def test_oracle_nonlinear_matches_3d_composition(k0, mm, nn, theta, alpha, gamma, herm):
    o = oracle.Oracle(k0, mm, nn, alpha, gamma, 200.0, 0.02, theta=theta)
    re, im = _state(o, 7 * k0 + mm + 3 * nn, herm)
    o.set_state(re, im)
    got = o.nonlinear()
    nl = composed_3d(o, re, im, theta)[0]
    want = o.spec2mpt(np.ascontiguousarray(nl.real), np.ascontiguousarray(nl.imag))
    scale = max(np.abs(want[0]).max(), np.abs(want[1]).max())
    for part in range(2):
        np.testing.assert_allclose(got[part], want[part], rtol=0, atol=1e-13 * scale)


def test_oracle_two_steps_follow_pinned_nonlinear():
    """With vel_nonlinear pinned, vel_TS (forward Euler, then AB2 and the implicit
    diagonal solve, program/velocity.f90:21-41) is pinned by its update rule."""
    k0, mm, nn, theta = 4, 8, 16, 0.3
    o = oracle.Oracle(k0, mm, nn, 0.25, 0.4, 180.0, 0.02, theta=theta)
    re, im = _state(o, 3)
    l, r = o.lhsrhs()
    dt = 0.02

    def nl_of(a, b):
        nl = composed_3d(o, a, b, theta)[0]
        return o.spec2mpt(np.ascontiguousarray(nl.real), np.ascontiguousarray(nl.imag))

    nl0 = nl_of(re, im)
    c1 = [(r * x + dt / 2 * (3 * n - n)) / l for x, n in zip((re, im), nl0)]
    c1[0][0, 0, 0] = 0.0
    nl1 = nl_of(*c1)
    c2 = [(r * x + dt / 2 * (3 * n1 - n0)) / l for x, n1, n0 in zip(c1, nl1, nl0)]
    c2[0][0, 0, 0] = 0.0
    o.set_state(re, im)
    o.step(2)
    got = o.get_state()
    for part in range(2):
        np.testing.assert_allclose(got[part], c2[part], rtol=0, atol=1e-13 * np.abs(c2[part]).max())


# ---------------------------------------------------------------- flux form

def _gen():
    import os
    import sys
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mwf_amd", "csrc")
    sys.path.insert(0, here)
    try:
        import gen_galerkin
    finally:
        sys.path.pop(0)
    return gen_galerkin


def flux_composition(o, re, im):
    """vel_nonlinear through the flux form of the generated tables (flux_gen.hpp):
    rows of the base values on the 3/2 grid, forward transforms, then
    N_s = i alpha m B_s + A_s with A_s = sum of (s beta) row and i gamma n row."""
    G = _gen()
    k0, kk = o.k0, o.KK
    sr, si = o.mpt2spec(re, im)
    s = (sr + 1j * si).copy()
    s[0, 0, 1] += math.cos(o.params.theta)        # vel_addlam, program/velocity.f90:76-82
    s[0, 0, 2 * k0] += math.sin(o.params.theta)
    p = _to_xz(o, s)
    rows = G.flux_rows(k0)
    R = np.zeros(p.shape[:2] + (len(rows),))
    for r, (_, _, terms) in enumerate(rows):
        for a, b, c in terms:
            R[:, :, r] += float(c) * p[:, :, a] * p[:, :, b]
    Rs = _from_xz(o, R)
    kx, kz = _wavenumbers(o)
    out = np.zeros(s.shape, complex)
    for L, terms in enumerate(G.flux_lines(k0)):
        acc = 0
        for r, kind, sg in terms:
            f = {0: 1.0, 1: sg * BETA, 2: 1j * kz[:, :, 0]}[kind]
            acc = acc + f * Rs[:, :, r]
        if L < kk:
            out[:, :, L] += 1j * kx[:, :, 0] * acc
        else:
            out[:, :, L - kk] += acc
    return out


@pytest.mark.parametrize("k0,mm,nn,theta,alpha,gamma", CASES)
def test_flux_form_equals_advective_form(k0, mm, nn, theta, alpha, gamma):
    """The generated flux-form rows and line assembly reproduce the advective
    nonlinear term of the independent 3-D evaluation, so a flux-form kernel
    computes vel_nonlinear up to rounding."""
    o = oracle.Oracle(k0, mm, nn, alpha, gamma, 200.0, 0.02, theta=theta)
    re, im = _state(o, 5 * k0 + mm + nn)
    adv = composed_3d(o, re, im, theta)[0]
    flux = flux_composition(o, re, im)
    assert np.abs(flux - adv).max() <= 1e-13 * np.abs(adv).max()


@pytest.mark.parametrize("k0", [4, 6])
def test_flux_rows_are_exact_projections(k0):
    """Each generated row equals the quadrature projection of its product."""
    G = _gen()
    rng = np.random.default_rng(k0)
    y = np.arange(NY) * (4.0 / NY)
    h, _, g, _ = _ybases(k0, y)
    su, sv, sw = slice(0, k0), slice(k0, 2 * k0 - 1), slice(2 * k0 - 1, 3 * k0 - 1)
    for _ in range(10):
        v = rng.standard_normal(3 * k0 - 1)
        U, V, W = v[su] @ h, v[sv] @ g, v[sw] @ h
        fields = {"uu": (U * U, h, 0), "uw": (U * W, h, 0), "ww": (W * W, h, 0),
                  "uv": (U * V, g, 1), "vw": (V * W, g, 1), "vv": (V * V, h, 0)}
        for name, i, terms in G.flux_rows(k0):
            f, B, off = fields[name]
            want = (f * B[i - off]).mean() / (B[i - off] ** 2).mean()
            got = sum(float(c) * v[a] * v[b] for a, b, c in terms)
            assert abs(got - want) <= 1e-13, (name, i)
