"""GPU: the autotuner's candidate plans (mwf_autotune) all keep parity, and
the timed choice is one of them and is reused for a second handle of the
same size.  Plan bits: 1 = phase-sequential z kernel, 2 = separate
x-inverse and x-forward kernels, 4 = two-points-per-thread z kernel, 8 = flux
form (16: its z kernel at two waves per SIMD), 32 = at NN = 1024 the fold
phases of an x-row in three blocks with three NL partials."""
import math

import numpy as np
import pytest

import mwf_amd as mw
import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture
def tuning():
    yield mw.autotune
    mw.autotune(0)


@pytest.mark.parametrize("k0,mm,nn", [(4, 32, 512), (6, 64, 256)])
@pytest.mark.parametrize("choice", [0, 1, 2, 3, 4, 6])
def test_every_plan_keeps_parity(tuning, k0, mm, nn, choice):
    tuning(2 + choice)
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02)
    v = mw.Velocity(p)
    assert v.plan_info() == choice
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    re, im = mw.ic_gaussian(p, 4, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    v.vel_TS(20)
    o.step(20)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"plan {choice} k0={k0} mm={mm} nn={nn}: 20 steps rel_l2 {err:.2e}, energy {abs(e - eref) / eref:.2e}")
    assert err <= 1e-12
    assert abs(e - eref) <= 1e-12 * eref


def test_timed_choice_is_cached(tuning):
    tuning(1)
    p = mw.make_params(6, 64, 512, 16 * math.pi, 128 * math.pi, 200.0, 0.02)
    a = mw.Velocity(p)
    b = mw.Velocity(p)
    assert a.plan_info() in (0, 1, 2, 3, 4, 6, 8, 10, 24, 26, 32, 34)
    assert b.plan_info() == a.plan_info()
    # the tuner leaves a zero state and a forward-Euler first step
    re, im = a.get_state()
    assert not re.any() and not im.any()
    assert a.get_onl()[2]
    print(f"timed plan for M6 64x512: {a.plan_info()}")


@pytest.mark.parametrize("k0,nn", [(4, 16), (4, 64), (6, 128), (4, 256), (6, 512)])
@pytest.mark.parametrize("mixed", [False, True])
def test_two_point_z_kernel_matches_oracle(tuning, k0, nn, mixed):
    """k_z2 (plan bit 4) across row lengths: x-rows per block 512/nn from 32
    down to 1, and ragged last blocks (P = 3*mm not a multiple of it)."""
    tuning(2 + mw.PLAN_Z2)
    mm = 16
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02,
                       precision=mw.MWF_MIXED if mixed else mw.MWF_FP64)
    v = mw.Velocity(p)
    assert v.plan_info() == mw.PLAN_Z2
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    re, im = mw.ic_gaussian(p, 6, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    v.vel_TS(10)
    o.step(10)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"k_z2 k0={k0} mm={mm} nn={nn} mixed={mixed}: 10 steps rel_l2 {err:.2e}, "
          f"energy {abs(e - eref) / eref:.2e}")
    if mixed:
        assert err <= 1e-5 and abs(e - eref) <= 1e-5 * eref
    else:
        assert err <= 1e-12 and abs(e - eref) <= 1e-12 * eref


@pytest.mark.parametrize("k0,mm,nn", [(4, 16, 1024), (6, 32, 1024)])
@pytest.mark.parametrize("mixed", [False, True])
def test_phase_split_z_kernel_matches_oracle(tuning, k0, mm, nn, mixed):
    """Plan bit 32 (with the split x kernels): the three fold phases of each
    1024-long z row in three blocks, their NL partials summed by k_xfwd, for
    the nonlinear term and 20 steps against the oracle."""
    tuning(2 + (mw.PLAN_ZB_SPLIT | mw.PLAN_SPLIT_X))
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02,
                       precision=mw.MWF_MIXED if mixed else mw.MWF_FP64)
    v = mw.Velocity(p)
    assert v.plan_info() == mw.PLAN_ZB_SPLIT | mw.PLAN_SPLIT_X
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    re, im = mw.ic_gaussian(p, 5, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    a, b = v.vel_nonlinear(), o.nonlinear()
    nerr = max(np.abs(a[i] - b[i]).max() for i in range(2)) / max(np.abs(b[0]).max(), np.abs(b[1]).max())
    v.set_state(re, im)
    v.vel_TS(20)
    o.step(20)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    bar = 1e-5 if mixed else 1e-12
    print(f"phase-split z k0={k0} mm={mm} nn={nn} mixed={mixed}: nonlinear {nerr:.2e}, "
          f"20 steps rel_l2 {err:.2e}, energy {abs(e - eref) / eref:.2e}", flush=True)
    assert nerr <= (1e-5 if mixed else 1e-12)
    assert err <= bar and abs(e - eref) <= bar * eref
    v.close()


@pytest.mark.parametrize("k0,mm,nn", [(4, 16, 1024), (6, 32, 1024)])
def test_phase_split_z_kernel_fused_x_matches_oracle(tuning, k0, mm, nn):
    """Plan bit 32 alone in FP64: the phase-split z kernel with the fused x
    kernel summing the three NL partials, 20 steps against the oracle."""
    tuning(2 + mw.PLAN_ZB_SPLIT)
    p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02)
    v = mw.Velocity(p)
    assert v.plan_info() == mw.PLAN_ZB_SPLIT
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    re, im = mw.ic_gaussian(p, 5, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    v.vel_TS(20)
    o.step(20)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"phase-split z + fused x k0={k0} mm={mm} nn={nn}: 20 steps rel_l2 {err:.2e}, "
          f"energy {abs(e - eref) / eref:.2e}", flush=True)
    assert err <= 1e-12 and abs(e - eref) <= 1e-12 * eref
    v.close()
