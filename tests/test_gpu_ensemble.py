"""GPU tests of the batched ensemble path (mwf_ens_*)."""
import math

import numpy as np
import pytest

import mwf_amd as mw
import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = sum(np.sum((x - y) ** 2) for x, y in zip(a, b))
    den = sum(np.sum(y ** 2) for y in b)
    return math.sqrt(num / den)


def members():
    """Mixed sizes, bases, Re and boxes: four size classes."""
    out = []
    for i, (k0, mm, nn, L, re) in enumerate([(4, 32, 32, 8 * math.pi, 180.0), (4, 32, 32, 8 * math.pi, 210.0),
                                             (4, 64, 32, 12.0, 190.0), (6, 32, 32, 8 * math.pi, 230.0),
                                             (4, 32, 32, 6.0, 160.0), (4, 64, 64, 16 * math.pi, 200.0)]):
        out.append((mw.make_params(k0, mm, nn, L, L, re, 0.02), 100 + i, 0.05 + 0.05 * i, 4.0))
    return out


def test_ensemble_members_match_single_models_and_oracle():
    ms = members()
    ens = mw.Ensemble([m[0] for m in ms])
    assert ens.nclasses() == 4
    ics = [mw.ic_gaussian(p, s, scl, k0) for p, s, scl, k0 in ms]
    for i, ic in enumerate(ics):
        ens.set_state(i, *ic)
    ens.step(25)  # first step eager, the rest from graphs of 10 steps
    for i, (p, s, scl, k0) in enumerate(ms):
        single = mw.Velocity(p)
        single.set_state(*ics[i])
        single.vel_TS(25)
        a, b = ens.get_state(i), single.get_state()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"member {i} differs from single model"
        single.close()
        if i in (0, 3):
            o = oracle.Oracle(p.k0, p.mm, p.nn, p.alpha, p.gamma, p.re, p.dt)
            o.set_state(*ics[i])
            o.step(25)
            err = rel_l2(a, o.get_state())
            print(f"member {i}: rel L2 vs oracle after 25 steps {err:.2e}")
            assert err <= 1e-12
    e = ens.energies()
    assert e.shape == (6,) and np.all(np.isfinite(e))


def test_graph_replay_equals_eager_steps():
    ms = members()[:3]
    a, b = mw.Ensemble([m[0] for m in ms]), mw.Ensemble([m[0] for m in ms])
    for i, (p, s, scl, k0) in enumerate(ms):
        ic = mw.ic_gaussian(p, s, scl, k0)
        a.set_state(i, *ic)
        b.set_state(i, *ic)
    a.step(31)                       # 1 eager + 10 + 10 + 10 from graphs
    for _ in range(31):
        b.step(1)                    # 1 eager + thirty 1-step graphs
    for i in range(len(ms)):
        x, y = a.get_state(i), b.get_state(i)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_relaminarisation_tracking():
    p = mw.make_params(4, 32, 32, 8 * math.pi, 8 * math.pi, 200.0, 0.02)
    quiet = mw.make_params(4, 32, 32, 8 * math.pi, 8 * math.pi, 20.0, 0.02, drag=0.5)
    ens = mw.Ensemble([p, quiet])
    ens.set_state(0, *mw.ic_gaussian(p, 1, 0.3, 4.0))
    ens.set_state(1, *mw.ic_gaussian(quiet, 2, 1e-3, 4.0))   # E = 1e-6, damped by drag 0.5
    ens.track(10, 5e-7)
    ens.step(200)
    e0, r0 = ens.trace(0)
    e1, r1 = ens.trace(1)
    assert len(e0) == 21 and len(e1) == 21
    assert e1[0] == pytest.approx(1e-6, rel=1e-9)
    assert r0 == -1 and 0 < r1 <= 200 and e1[-1] < 5e-7
    assert np.all(e1[1:][np.arange(20) * 10 + 10 >= r1] <= 5e-7 * 1.5)
