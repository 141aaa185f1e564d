"""GPU: warm restart with the AB2 history, and the on-device diagnostics of
the kept host I/O (io_save_Uspec, vel_history, turb_sample) against the oracle.

The diagnostics are checked against quantities the oracle pins:
var_mpt2spec (bit-identical to the compiled reference), vel_mpt2phys (DFT and
numpy pins) and vel_energy (Parseval); and against closed forms of the
wall-normal bases (h_0 = 1, h_k = sin/cos(k beta y), g_k = cos/sin(k beta y),
modes.M4.f90:83-108, modes.M6.f90:81-111).
"""
import math

import numpy as np
import pytest

import mwf_amd as mw
import oracle

pytestmark = pytest.mark.gpu


def setup(k0, mm, nn, steps=5, seed=3):
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02)
    v = mw.Velocity(p)
    v.set_state(*mw.ic_gaussian(p, seed, 0.2, 4.0))
    v.vel_TS(steps)
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    return p, v, o


@pytest.mark.parametrize("k0,mm,nn,prec", [(4, 32, 32, mw.MWF_FP64), (6, 64, 32, mw.MWF_FP64),
                                            (4, 512, 512, mw.MWF_FP64), (4, 64, 64, mw.MWF_MIXED),
                                            (6, 512, 256, mw.MWF_MIXED)])
def test_warm_restart_is_bit_identical(k0, mm, nn, prec):
    """The mixed form stores its AB2 history in FP32: get_onl widens it and
    set_onl rounds it back, so the resume stays bit-identical there too."""
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02, precision=prec)
    a = mw.Velocity(p)
    a.set_state(*mw.ic_gaussian(p, 1, 0.2, 4.0))
    assert a.get_onl()[2], "a new state starts with a forward-Euler step"
    a.vel_TS(30)
    cre, cim = a.get_state()
    hre, him, first = a.get_onl()
    assert not first
    a.vel_TS(20)
    ref = a.get_state()

    b = mw.Velocity(p)
    b.set_state(cre, cim)
    b.set_onl(hre, him)
    b.vel_TS(20)
    got = b.get_state()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])

    # a restart from vel_c alone (the reference's state file) is not the same run
    c = mw.Velocity(p)
    c.set_state(cre, cim)
    c.vel_TS(20)
    cold = c.get_state()
    diff = np.abs(cold[0] - ref[0]).max()
    print(f"k0={k0} mm={mm} nn={nn}: resume with history bit-identical; without it max diff {diff:.2e}")
    assert diff > 0.0


@pytest.mark.parametrize("k0,mm,nn", [(4, 32, 32), (6, 64, 32), (4, 128, 64)])
def test_spectrum_max_matches_oracle(k0, mm, nn):
    p, v, o = setup(k0, mm, nn)
    re, im = v.get_state()
    sre, sim = o.mpt2spec(re, im)                   # (NN, M, KK), no laminar profile
    d = np.sqrt(sre * sre + sim * sim)
    n_ref = d.max(axis=(1, 2))
    M = 2 * mm - 1
    dm = d.max(axis=(0, 2))                         # over n and slots, per stored m
    m_ref = np.array([dm[0]] + [max(dm[m], dm[M - m]) for m in range(1, mm)])
    m_got, n_got = v.spectrum_max()
    np.testing.assert_allclose(m_got, m_ref, rtol=1e-14, atol=0)
    np.testing.assert_allclose(n_got, n_ref, rtol=1e-14, atol=0)


def bases(k0, y):
    beta = math.pi / 2
    h = [1.0] + [math.sin(k * beta * y) if k % 2 else math.cos(k * beta * y) for k in range(1, k0)]
    g = [math.cos(k * beta * y) if k % 2 else math.sin(k * beta * y) for k in range(1, k0)]
    return np.array(h), np.array(g)


@pytest.mark.parametrize("k0,mm,nn", [(4, 32, 32), (6, 64, 32)])
@pytest.mark.parametrize("y", [0.0, 0.3, -0.7])
def test_vel_history_matches_oracle(k0, mm, nn, y):
    p, v, o = setup(k0, mm, nn)
    re, im = v.get_state()
    phys = o.mpt2phys(re, im)                       # (P, Q, KK)
    Q = 3 * nn
    nh = 32
    h, g = bases(k0, y)
    got = v.vel_history(nh, y)
    for j in range(nh):
        ji = (j * Q) // nh + 1
        V = phys[0, ji]
        u = h @ V[:k0]
        vv = g @ V[k0:2 * k0 - 1]
        w = h @ V[2 * k0 - 1:]
        if y == 0.0:  # sin(0) = 0, cos(0) = 1
            assert np.isclose(u, V[0] + V[2:k0:2].sum(), rtol=1e-14, atol=1e-16)
        exp = [ji * (2 * math.pi / p.gamma) / Q, u, vv, w]
        np.testing.assert_allclose(got[j], exp, rtol=1e-12, atol=1e-14 * np.abs(phys).max())


@pytest.mark.parametrize("k0,mm,nn", [(4, 32, 32), (6, 64, 32), (4, 512, 512)])
def test_turb_sample_matches_oracle_and_energy(k0, mm, nn):
    p, v, o = setup(k0, mm, nn)
    # all points: the component energies add up to vel_energy (/(PQ))
    full = v.turb_sample(1, 1)
    e = v.vel_energy()
    P, Q = 3 * mm, 3 * nn
    assert abs(full.sum() / (P * Q) - e) <= 1e-12 * e
    # sampled points against the oracle's physical field
    re, im = v.get_state()
    phys = o.mpt2phys(re, im)
    dm, dn = 6, 6
    got = v.turb_sample(dm, dn)
    S = phys[::dm, ::dn][: P // dm, : Q // dn]
    eu = S[..., 0] ** 2 + 0.5 * (S[..., 1:k0] ** 2).sum(-1)
    ev = 0.5 * (S[..., k0:2 * k0 - 1] ** 2).sum(-1)
    ew = S[..., 2 * k0 - 1] ** 2 + 0.5 * (S[..., 2 * k0:] ** 2).sum(-1)
    ref = np.stack([eu, ev, ew], -1)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * ref.max())
    print(f"k0={k0} mm={mm} nn={nn}: turb sample {got.shape}, sum/PQ vs energy ok")


@pytest.mark.parametrize("factor", [2, 4])
@pytest.mark.parametrize("k0", [4, 6])
def test_extend_x_repeats_the_field(k0, factor):
    """var_doubleX / var_quadX: a state of the box factor*Lx holding the same
    field.  Pin: on the 3/2 grids (P points each) the new field at x_j equals
    the old one at x_{factor*j mod P}, and the energy is unchanged, for a
    state band-limited to |m| < MM/factor."""
    mm, nn = 32, 32
    L = 8 * math.pi
    p = mw.make_params(k0, mm, nn, L, L, 200.0, 0.02)
    q = mw.make_params(k0, mm, nn, factor * L, L, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 5, 0.2, 8.0)
    M = 2 * mm - 1
    signed = np.where(np.arange(M) < mm, np.arange(M), np.arange(M) - M)
    keep = (np.abs(signed) * factor <= mm - 1)[None, :, None]
    re, im = re * keep, im * keep
    src, dst = mw.Velocity(p), mw.Velocity(q)
    src.set_state(re, im)
    src.extend_x_into(dst, factor)
    nre, nim = dst.get_state()
    a = src.vel_mpt2phys(re, im)
    b = dst.vel_mpt2phys(nre, nim)
    P = 3 * mm
    idx = (factor * np.arange(P)) % P
    err = np.abs(b - a[idx]).max() / np.abs(a).max()
    e0, e1 = src.vel_energy(), dst.vel_energy()
    print(f"k0={k0} factor={factor}: tiled field max err {err:.2e}, energy {e0:.15g} -> {e1:.15g}")
    assert err <= 1e-13
    assert abs(e1 - e0) <= 1e-12 * e0
    # dropped modes: nothing outside m = factor*m_s survives
    nz = np.abs(nre).sum(axis=(0, 2)) + np.abs(nim).sum(axis=(0, 2))
    assert np.all(nz[signed % factor != 0] == 0.0)


def test_blend_is_the_interpstate_combination():
    p = mw.make_params(4, 32, 32, 8 * math.pi, 8 * math.pi, 200.0, 0.02)
    s1, s2 = mw.ic_gaussian(p, 1, 0.2, 4.0), mw.ic_gaussian(p, 2, 0.3, 4.0)
    a, b, c = mw.Velocity(p), mw.Velocity(p), mw.Velocity(p)
    a.set_state(*s1)
    b.set_state(*s2)
    d = 0.3
    c.blend_into(a, b, d)
    got = c.get_state()
    for g, x, y in zip(got, s1, s2):
        assert np.array_equal(g, (1.0 - d) * x + d * y)
    assert c.get_onl()[2], "the blended state starts with a forward-Euler step"


def oracle_turb(o, re, im, xs, zs):
    """epos (program/velocity.f90:115-126, the oracle's reference-order C
    restatement) of vel_mpt2phys(c) on the strided grid of io_writeVTK_txz."""
    ph = o.mpt2phys(re, im)
    return np.array([[oracle.epos(o.k0, ph[ix, iz]) for iz in range(0, o.Q, zs)] for ix in range(0, o.P, xs)])


def safe_threshold(ep, q):
    """A threshold near the q-quantile of the sampled epos values, halfway
    between two neighbours, so that rounding cannot move a point across it."""
    v = np.sort(ep.ravel())
    i = int(q * (v.size - 1))
    return 0.5 * (v[i] + v[i + 1])


@pytest.mark.parametrize("k0,mm,nn,xs,zs", [(4, 32, 32, 6, 6), (6, 64, 32, 1, 1), (4, 128, 256, 7, 5),
                                            (4, 512, 512, 6, 6)])
def test_turb_fraction_matches_oracle(k0, mm, nn, xs, zs):
    """io_writeVTK_txz's count (program/io.f90:594-615): epos > ct on the strided
    grid, over (N_+1)(M_+1) with M_ = 3M/xs, N_ = 3N/zs."""
    p, v, o = setup(k0, mm, nn)
    re, im = v.get_state()
    ep = oracle_turb(o, re, im, xs, zs)
    den = (o.Q // zs + 1) * (o.P // xs + 1)
    for q in (0.1, 0.5, 0.9):
        ct = safe_threshold(ep, q)
        want = np.count_nonzero(ep > ct) / den
        got = v.turb_fraction(xs, zs, ct)
        print(f"[turbfrac] k0={k0} mm={mm} nn={nn} q={q}: {got:.6f} vs {want:.6f}", flush=True)
        assert got == want
    assert v.turb_fraction(xs, zs, 1e300) == 0.0
    # the diagnostic does not disturb the stepping
    a = mw.Velocity(p)
    a.set_state(re, im)
    v.set_state(re, im)
    v.turb_fraction(xs, zs, 0.0)
    a.vel_TS(2)
    v.vel_TS(2)
    x, y = a.get_state(), v.get_state()
    assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
    v.close()
    a.close()
