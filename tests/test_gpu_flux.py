"""GPU: the flux (conservative) form of vel_nonlinear (plan bit 8, k_zc and
the flux x kernels) against the oracle, which follows the reference's
advective form (program/velocity.f90:92-113, modes.M4.f90/modes.M6.f90).
The two forms are equal as mathematics (the identity and its generated tables
are pinned on the CPU in tests/test_oracle_nonlinear.py), so the GPU flux
path must meet the same bar as the advective one: 1e-12 in FP64."""
import math

import numpy as np
import pytest

import mwf_amd as mw
import oracle

pytestmark = pytest.mark.gpu

FLUX = mw.PLAN_FLUX
FLUX_PLANS = [FLUX, FLUX | mw.PLAN_FLUX_W2, FLUX | mw.PLAN_SPLIT_X]


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture
def tuning():
    yield mw.autotune
    mw.autotune(0)


def _pair(k0, mm, nn, precision=mw.MWF_FP64, theta=0.0):
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02, precision=precision,
                       thdeg=theta)
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    return p, o


@pytest.mark.parametrize("k0,mm,nn", [(4, 16, 16), (6, 32, 64), (4, 64, 128), (6, 16, 256), (4, 128, 512),
                                      (6, 64, 512), (4, 1024, 32), (6, 1024, 16)])
@pytest.mark.parametrize("choice", FLUX_PLANS)
def test_flux_plan_steps_match_oracle(tuning, k0, mm, nn, choice):
    tuning(2 + choice)
    p, o = _pair(k0, mm, nn)
    v = mw.Velocity(p)
    assert v.plan_info() == choice
    re, im = mw.ic_gaussian(p, 4, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    v.vel_TS(20)
    o.step(20)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"[flux] plan {choice} k0={k0} mm={mm} nn={nn}: 20 steps rel_l2 {err:.2e}, "
          f"energy {abs(e - eref) / eref:.2e}", flush=True)
    assert err <= 1e-12 and abs(e - eref) <= 1e-12 * eref
    v.close()


@pytest.mark.parametrize("k0,mm,nn,theta", [(4, 32, 32, 0.0), (6, 64, 32, 17.0), (4, 128, 256, -40.0)])
@pytest.mark.parametrize("choice", FLUX_PLANS)
def test_flux_nonlinear_matches_oracle(tuning, k0, mm, nn, theta, choice):
    tuning(2 + choice)
    p, o = _pair(k0, mm, nn, theta=theta)
    v = mw.Velocity(p)
    re, im = mw.ic_gaussian(p, 9, 0.3, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    a, b = v.vel_nonlinear(), o.nonlinear()
    scale = max(np.abs(b[0]).max(), np.abs(b[1]).max())
    err = max(np.abs(a[i] - b[i]).max() for i in range(2)) / scale
    print(f"[flux] nonlinear plan {choice} k0={k0} mm={mm} nn={nn} theta={theta}: {err:.2e}", flush=True)
    assert err <= 1e-12
    v.close()


@pytest.mark.parametrize("k0,mm,nn", [(4, 32, 32), (6, 64, 128), (4, 16, 512), (6, 512, 64), (4, 32, 1024),
                                      (6, 16, 1024)])
@pytest.mark.parametrize("choice", [FLUX, FLUX | mw.PLAN_FLUX_W2])
def test_flux_plan_mixed_form(tuning, k0, mm, nn, choice):
    """Mixed form (FP32 transforms and products) of the flux plan: within the
    mixed bar of the FP64 oracle.  At NN = 1024 the FP32 flux z kernel holds a
    whole 3072-point row per 1024-thread block (one compiled form only)."""
    tuning(2 + choice)
    p, o = _pair(k0, mm, nn, precision=mw.MWF_MIXED)
    v = mw.Velocity(p)
    want = choice if nn < 1024 else FLUX
    assert v.plan_info() == want
    re, im = mw.ic_gaussian(p, 6, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    v.vel_TS(20)
    o.step(20)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"[flux] mixed plan {choice} k0={k0} mm={mm} nn={nn}: rel_l2 {err:.2e}, "
          f"energy {abs(e - eref) / eref:.2e}", flush=True)
    assert err <= 1e-5 and abs(e - eref) <= 1e-5 * eref
    v.close()


def test_flux_m4_1024_shape(tuning):
    """The headline workload's shape (M4, 1024^2 modes, Lx = Lz = 128 pi,
    Re 180, dt 0.02, seed 2) through the flux plan, 20 steps against the
    oracle."""
    tuning(2 + FLUX)
    p = mw.make_params(4, 512, 512, 128 * math.pi, 128 * math.pi, 180.0, 0.02)
    o = oracle.Oracle(4, 512, 512, p.alpha, p.gamma, p.re, p.dt)
    oracle.set_threads(16)
    v = mw.Velocity(p)
    assert v.plan_info() == FLUX
    re, im = mw.ic_gaussian(p, 2, 0.3, 8.0)
    v.set_state(re, im)
    o.set_state(re, im)
    v.vel_TS(20)
    o.step(20)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"[flux] M4 1024^2 20 steps rel_l2 {err:.2e}, energy {abs(e - eref) / eref:.2e}", flush=True)
    assert err <= 1e-12 and abs(e - eref) <= 1e-12 * eref
    v.close()
