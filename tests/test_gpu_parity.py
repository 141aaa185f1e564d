"""GPU parity: the HIP path through the C-ABI against the oracle.

FP64 bar: relative L2 <= 1e-12 on the spectral state and on energies (the
bar for the reference's FP64 path).  Mixed form: energy within 1e-5
relative of the FP64 path.
"""
import math

import numpy as np
import pytest

import mwf_amd as mw
import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(k0, mm, nn, lx, lz, re, dt, precision=mw.MWF_FP64, **kw):
    p = mw.make_params(k0, mm, nn, lx, lz, re, dt, precision=precision, **kw)
    o = oracle.Oracle(k0, mm, nn, p.alpha, p.gamma, re, dt, drag=p.drag, vdrag=p.vdrag,
                      hypo=p.hypo, theta=p.theta, dragall=bool(p.dragall))
    return p, mw.Velocity(p), o


SIZES = [(4, 16, 16), (4, 32, 32), (4, 64, 32), (4, 32, 64), (4, 128, 128), (4, 256, 128),
         (6, 32, 32), (6, 64, 64)]


@pytest.mark.parametrize("k0,mm,nn", SIZES)
def test_mpt2phys(k0, mm, nn):
    p, v, o = make(k0, mm, nn, 8 * math.pi, 6 * math.pi, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 7, 0.2, 4.0)
    ph = v.vel_mpt2phys(re, im)
    ref = o.mpt2phys(re, im)
    err = np.linalg.norm(ph - ref) / np.linalg.norm(ref)
    print(f"mpt2phys k0={k0} mm={mm} nn={nn} rel_l2={err:.3e}")
    assert err <= 1e-13


@pytest.mark.parametrize("k0,mm,nn", SIZES)
def test_nonlinear(k0, mm, nn):
    p, v, o = make(k0, mm, nn, 8 * math.pi, 6 * math.pi, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 11, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    nl = v.vel_nonlinear()
    ref = o.nonlinear()
    err = rel_l2(nl, ref)
    print(f"nonlinear k0={k0} mm={mm} nn={nn} rel_l2={err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("k0,mm,nn", SIZES)
def test_energy(k0, mm, nn):
    p, v, o = make(k0, mm, nn, 8 * math.pi, 6 * math.pi, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 5, 0.2, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    e, eref = v.vel_energy(), o.energy()
    print(f"energy k0={k0} mm={mm} nn={nn} gpu={e!r} oracle={eref!r} rel={abs(e - eref) / eref:.3e}")
    assert abs(e - eref) <= 1e-12 * abs(eref)


@pytest.mark.parametrize("k0,mm,nn", [(4, 32, 32), (6, 32, 32), (4, 64, 64)])
def test_steps_match_oracle(k0, mm, nn):
    p, v, o = make(k0, mm, nn, 8 * math.pi, 8 * math.pi, 200.0, 0.02)
    re, im = mw.ic_gaussian(p, 1, 0.1, 4.0)
    v.set_state(re, im)
    o.set_state(re, im)
    worst_e = 0.0
    for s in range(100):
        e, eref = v.vel_energy(), o.energy()
        worst_e = max(worst_e, abs(e - eref) / abs(eref))
        v.vel_TS(1)
        o.step(1)
    err = rel_l2(v.get_state(), o.get_state())
    print(f"100 steps k0={k0} mm={mm} nn={nn}: state rel_l2={err:.3e} worst energy rel={worst_e:.3e}")
    assert err <= 1e-12
    assert worst_e <= 1e-12
