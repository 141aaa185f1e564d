"""GPU tests of the transform entry points (tra_spec2phys / tra_phys2spec,
program/transform.fftw5.f90:64-145) and of two host-boundary behaviours of the
step: the state replacement used by vel_imposesym (program/velocity.f90:68-74),
which keeps AB2, and vel_nonlinear as a pure diagnostic.  All against the
oracle."""
import math

import numpy as np
import pytest

import mwf_amd as mw
import oracle

pytestmark = pytest.mark.gpu

SIZES = [(4, 16, 16), (4, 32, 64), (6, 64, 32), (4, 128, 256), (6, 256, 128), (4, 512, 512), (6, 1024, 64),
         (4, 64, 1024)]


def _pair(k0, mm, nn, precision=mw.MWF_FP64):
    lx, lz = 0.25 * math.pi * 2 * mm, 0.25 * math.pi * 2 * nn
    p = mw.make_params(k0, mm, nn, lx, lz, 200.0, 0.01, precision=precision)
    return mw.Velocity(p), oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, 200.0, 0.01)


@pytest.mark.parametrize("k0,mm,nn", SIZES)
def test_spec2phys_phys2spec_match_oracle(k0, mm, nn):
    v, o = _pair(k0, mm, nn)
    rng = np.random.default_rng(mm * 7 + nn)
    sre, sim = rng.standard_normal((2,) + o.spec_shape)
    print(f"[transform] k0={k0} mm={mm} nn={nn}", flush=True)
    ph = v.tra_spec2phys(sre, sim)
    want = o.spec2phys(sre, sim)
    assert np.abs(ph - want).max() <= 1e-13 * np.abs(want).max()
    br, bi = v.tra_phys2spec(want)
    wr, wi = o.phys2spec(want)
    scale = max(np.abs(wr).max(), np.abs(wi).max())
    assert np.abs(br - wr).max() <= 1e-13 * scale and np.abs(bi - wi).max() <= 1e-13 * scale
    v.close()


def test_transform_round_trip_mixed():
    """In the mixed form the transforms run in FP32: the band-limited round
    trip holds to FP32 rounding."""
    k0, mm, nn = 4, 64, 64
    v, o = _pair(k0, mm, nn, mw.MWF_MIXED)
    rng = np.random.default_rng(4)
    sre, sim = rng.standard_normal((2,) + o.spec_shape)
    M = o.M
    for m in range(1, mm):  # Hermitian n = 0 column, real DC
        sre[0, M - m, :] = sre[0, m, :]
        sim[0, M - m, :] = -sim[0, m, :]
    sim[0, 0, :] = 0.0
    br, bi = v.tra_phys2spec(v.tra_spec2phys(sre, sim))
    assert np.abs(br - sre).max() < 2e-5 and np.abs(bi - sim).max() < 2e-5
    v.close()


def _reflect(re, im):
    """var_reflect (program/variables.f90:113-131) on an M4 mpt (slots 1..7 are
    indices 0..6 here): at (m, n) = (0, 0) zero Re of slots 1, 3, 6; elsewhere
    zero Im of 1, 3, 6 and Re of 2, 4, 5, 7."""
    re, im = re.copy(), im.copy()
    mean_re = re[0, 0, :].copy()
    mean_im = im[0, 0, :].copy()
    im[:, :, [0, 2, 5]] = 0.0
    re[:, :, [1, 3, 4, 6]] = 0.0
    re[0, 0, :], im[0, 0, :] = mean_re, mean_im
    re[0, 0, [0, 2, 5]] = 0.0
    return re, im


def test_symmetry_projection_keeps_ab2():
    """vel_imposesym each step (program/main.f90:22) with s_reflect: the GPU
    state replacement (mwf_put_state) keeps the AB2 history like the
    reference's in-place projection; compared with the oracle doing the same."""
    k0, mm, nn = 4, 32, 32
    p = mw.make_params(k0, mm, nn, 8 * math.pi, 8 * math.pi, 200.0, 0.02)
    re, im = _reflect(*mw.ic_gaussian(p, 1, 0.1, 4.0))
    v = mw.Velocity(p)
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, 200.0, 0.02)
    v.set_state(re, im)
    o.set_state(re, im)
    for s in range(20):
        a = _reflect(*v.get_state())
        b = _reflect(*o.get_state())
        v.put_state(*a)
        o.put_state(*b)
        v.vel_TS(1)
        o.step(1)
    a, b = v.get_state(), o.get_state()
    err = math.sqrt(sum(((a[i] - b[i]) ** 2).sum() for i in range(2)) / sum((b[i] ** 2).sum() for i in range(2)))
    print(f"[symmetry] 20 projected steps: state rel L2 {err:.2e}", flush=True)
    assert err <= 1e-12
    # and the history was really kept: restarting AB2 every step differs
    w = mw.Velocity(p)
    w.set_state(re, im)
    for s in range(20):
        w.set_state(*_reflect(*w.get_state()))
        w.vel_TS(1)
    c = w.get_state()
    assert math.sqrt(sum(((c[i] - b[i]) ** 2).sum() for i in range(2))) > 1e-8
    v.close()
    w.close()


@pytest.mark.parametrize("precision", [mw.MWF_FP64, mw.MWF_MIXED])
def test_nonlinear_diagnostic_does_not_perturb_steps(precision):
    k0, mm, nn = 4, 64, 64
    p = mw.make_params(k0, mm, nn, 16 * math.pi, 16 * math.pi, 200.0, 0.02, precision=precision)
    ic = mw.ic_gaussian(p, 7, 0.2, 8.0)
    a, b = mw.Velocity(p), mw.Velocity(p)
    a.set_state(*ic)
    b.set_state(*ic)
    a.vel_TS(3)
    b.vel_TS(3)
    b.vel_nonlinear()
    a.vel_TS(3)
    b.vel_TS(3)
    x, y = a.get_state(), b.get_state()
    assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
    a.close()
    b.close()
