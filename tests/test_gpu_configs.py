"""GPU parity on the benchmark workloads at their full sizes, the 1024-long
z rows (the phase-sequential z pass) and the 1024-long x rows, and the mixed
FP32/FP64 form.

FP64 bar: relative L2 <= 1e-12 on the spectral state, and energies within
1e-12 relative, against the oracle on the same seeded initial condition.
Mixed bar: energy within 1e-5 relative of the FP64 path.
"""
import math
import os
import sys

import numpy as np
import pytest

import mwf_amd as mw
import oracle

sys.path.insert(0, os.path.dirname(__file__))
import cases  # noqa: E402

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def oracle_of(p):
    return oracle.Oracle(p.k0, p.mm, p.nn, p.alpha, p.gamma, p.re, p.dt, drag=p.drag, vdrag=p.vdrag,
                         hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))


LONG_ROWS = [(4, 16, 1024), (6, 16, 1024), (6, 32, 1024), (4, 1024, 16), (6, 1024, 32)]


def box(mm, nn):
    # the benchmark boxes have L = pi/4 per mode (8pi at 32 modes, 128pi at 512):
    # a box much shorter than that puts kx*dt*|u| beyond the AB2 stability limit
    # and the 100-step comparison then measures that instability, not parity
    return mm * math.pi / 4, nn * math.pi / 4


@pytest.mark.parametrize("k0,mm,nn", LONG_ROWS)
def test_long_rows_nonlinear_energy_phys(k0, mm, nn):
    p = mw.make_params(k0, mm, nn, *box(mm, nn), 200.0, 0.02)
    v, o = mw.Velocity(p), oracle_of(p)
    re, im = mw.ic_gaussian(p, 11, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    e_nl = rel_l2(v.vel_nonlinear(), o.nonlinear())
    e, eref = v.vel_energy(), o.energy()
    ph, phref = v.vel_mpt2phys(re, im), o.mpt2phys(re, im)
    e_ph = np.linalg.norm(ph - phref) / np.linalg.norm(phref)
    print(f"k0={k0} mm={mm} nn={nn}: nonlinear {e_nl:.3e} energy {abs(e - eref) / eref:.3e} phys {e_ph:.3e}")
    assert e_nl <= 1e-12
    assert abs(e - eref) <= 1e-12 * eref
    assert e_ph <= 1e-13


@pytest.mark.parametrize("k0,mm,nn", LONG_ROWS)
def test_long_rows_100_steps(k0, mm, nn):
    p = mw.make_params(k0, mm, nn, *box(mm, nn), 200.0, 0.02)
    v, o = mw.Velocity(p), oracle_of(p)
    re, im = mw.ic_gaussian(p, 1, 0.1, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    worst = 0.0
    for s in range(0, 100, 10):
        e, eref = v.vel_energy(), o.energy()
        worst = max(worst, abs(e - eref) / eref)
        v.vel_TS(10)
        o.step(10)
    err = rel_l2(v.get_state(), o.get_state())
    print(f"100 steps k0={k0} mm={mm} nn={nn}: state {err:.3e} worst energy {worst:.3e}")
    assert err <= 1e-12
    assert worst <= 1e-12


def test_m4_1024_first_100_steps():
    """M4, 1024^2 modes, Lx=Lz=128pi, Re=180, dt=0.02, seed 2: the headline
    drop-in case, 100 steps against the oracle."""
    p = cases.params("m4_1024")
    seed, scl, k0ic = cases.ic_of("m4_1024")
    v, o = mw.Velocity(p), oracle_of(p)
    re, im = mw.ic_gaussian(p, seed, scl, k0ic)
    v.set_state(re, im)
    o.set_state(re, im)
    worst = 0.0
    for s in range(0, 100, 25):
        e, eref = v.vel_energy(), o.energy()
        worst = max(worst, abs(e - eref) / eref)
        v.vel_TS(25)
        o.step(25)
        print(f"  step {s + 25}", flush=True)
    e, eref = v.vel_energy(), o.energy()
    worst = max(worst, abs(e - eref) / eref)
    err = rel_l2(v.get_state(), o.get_state())
    print(f"m4_1024 100 steps: state rel_l2 {err:.3e}, worst energy rel {worst:.3e}")
    assert err <= 1e-12
    assert worst <= 1e-12


def test_m6_2048_two_steps():
    """M6, 2048^2 modes (3072^2 grid), Lx=Lz=256pi, Re=250, dt=0.01, seed 3:
    nonlinear term and two steps against the oracle (about 10 s of oracle
    time per step on 16 cores)."""
    p = cases.params("m6_2048")
    seed, scl, k0ic = cases.ic_of("m6_2048")
    v, o = mw.Velocity(p), oracle_of(p)
    re, im = mw.ic_gaussian(p, seed, scl, k0ic)
    v.set_state(re, im)
    o.set_state(re, im)
    e_nl = rel_l2(v.vel_nonlinear(), o.nonlinear())
    print(f"m6_2048 nonlinear rel_l2 {e_nl:.3e}", flush=True)
    assert e_nl <= 1e-12
    v.vel_TS(2)
    o.step(2)
    err = rel_l2(v.get_state(), o.get_state())
    e, eref = v.vel_energy(), o.energy()
    print(f"m6_2048 2 steps: state rel_l2 {err:.3e}, energy rel {abs(e - eref) / eref:.3e}")
    assert err <= 1e-12
    assert abs(e - eref) <= 1e-12 * eref


@pytest.mark.parametrize("name,steps", [("m4_64", 1000), ("m4_1024", 100), ("m6_2048", 50)])
def test_mixed_energy_within_1e5(name, steps):
    """Mixed form (FP32 transforms and products, FP64 update and energy) against
    the FP64 path from the same initial condition; the per-step energy error
    is checked every step."""
    seed, scl, k0ic = cases.ic_of(name)
    p64, p32 = cases.params(name), cases.params(name, precision=mw.MWF_MIXED)
    v64, v32 = mw.Velocity(p64), mw.Velocity(p32)
    re, im = mw.ic_gaussian(p64, seed, scl, k0ic)
    v64.set_state(re, im)
    v32.set_state(re, im)
    worst = 0.0
    for s in range(steps):
        v64.vel_TS(1)
        v32.vel_TS(1)
        e64, e32 = v64.vel_energy(), v32.vel_energy()
        worst = max(worst, abs(e32 - e64) / e64)
    st = rel_l2(v32.get_state(), v64.get_state())
    print(f"mixed {name}: {steps} steps, worst energy rel {worst:.3e}, final state rel_l2 {st:.3e}")
    assert worst <= 1e-5


def test_m6_2048_phase_split_matches_phase_sequential():
    """M6 2048^2 (the k_zb size), FP64: the phase-split z kernel with the split
    (plan 34) and the fused (plan 32) x kernels against the phase-sequential
    one (pinned to the oracle at this size by test_m6_2048_two_steps):
    nonlinear term and 5 steps, to rounding."""
    p = cases.params("m6_2048")
    seed, scl, k0ic = cases.ic_of("m6_2048")
    re, im = mw.ic_gaussian(p, seed, scl, k0ic)
    out = {}
    try:
        for plan in (2, 34, 32):
            mw.autotune(2 + plan)
            v = mw.Velocity(p)
            assert v.plan_info() == plan
            v.set_state(re, im)
            nl = v.vel_nonlinear()
            v.set_state(re, im)
            v.vel_TS(5)
            out[plan] = (nl, v.get_state(), v.vel_energy())
            v.close()
    finally:
        mw.autotune(0)
    na, sa, ea = out[2]
    for plan in (34, 32):
        nb, sb, eb = out[plan]
        nerr = max(np.abs(na[i] - nb[i]).max() for i in range(2)) / max(np.abs(na[0]).max(), np.abs(na[1]).max())
        serr = rel_l2(sb, sa)
        print(f"m6_2048 phase-split (plan {plan}) vs phase-sequential: nonlinear {nerr:.2e}, 5 steps {serr:.2e}, "
              f"energy {abs(eb - ea) / ea:.2e}", flush=True)
        assert nerr <= 1e-13 and serr <= 1e-13 and abs(eb - ea) <= 1e-13 * ea
