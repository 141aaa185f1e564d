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


def test_uneven_mixed_ensemble_512_member_shapes():
    """The shape of the 512-member uneven ensemble: M4 and M6 members of
    64^2..512^2 modes in the mixed form, box 0.4 pi per mode, Re in [150, 300],
    dt 0.01, plus one FP64 member.  Every member equals its single-model run
    in its own precision bit for bit; the FP64 member matches the oracle; each
    mixed member's energy stays within 1e-5 of the same member run in FP64."""
    spec = [  # (k0, modes, Re, precision)
        (4, 64, 180.0, mw.MWF_MIXED), (4, 64, 260.0, mw.MWF_MIXED), (4, 128, 150.0, mw.MWF_MIXED),
        (6, 128, 220.0, mw.MWF_MIXED), (6, 64, 300.0, mw.MWF_MIXED), (4, 256, 240.0, mw.MWF_MIXED),
        (4, 512, 200.0, mw.MWF_MIXED), (6, 256, 170.0, mw.MWF_MIXED), (4, 128, 190.0, mw.MWF_FP64)]
    ps = []
    for k0, modes, re, prec in spec:
        L = 0.4 * math.pi * modes
        ps.append(mw.make_params(k0, modes // 2, modes // 2, L, L, re, 0.01, precision=prec))
    ics = [mw.ic_gaussian(p, 5000 + i, 0.2, 8.0) for i, p in enumerate(ps)]
    ens = mw.Ensemble(ps)
    assert ens.nclasses() == 8
    for i, ic in enumerate(ics):
        ens.set_state(i, *ic)
    nsteps = 24
    ens.step(nsteps)
    e = ens.energies()
    for i, p in enumerate(ps):
        print(f"[uneven] member {i}: k0={p.k0} mm={p.mm} prec={p.precision}", flush=True)
        single = mw.Velocity(p)
        single.set_state(*ics[i])
        single.vel_TS(nsteps)
        a, b = ens.get_state(i), single.get_state()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"member {i} differs from single model"
        assert e[i] == single.vel_energy()
        single.close()
        if p.precision == mw.MWF_MIXED:
            p64 = mw.make_params(p.k0, p.mm, p.nn, 2 * math.pi / p.alpha, 2 * math.pi / p.gamma, p.re, p.dt)
            ref = mw.Velocity(p64)
            ref.set_state(*ics[i])
            ref.vel_TS(nsteps)
            e64 = ref.vel_energy()
            ref.close()
            assert abs(e[i] - e64) <= 1e-5 * e64, (i, e[i], e64)
        else:
            o = oracle.Oracle(p.k0, p.mm, p.nn, p.alpha, p.gamma, p.re, p.dt)
            o.set_state(*ics[i])
            o.step(nsteps)
            assert rel_l2(a, o.get_state()) <= 1e-12
    ens.close()


def test_256_member_class_m4_128():
    """One size class of 256 M4 members of 128^2 modes (the 256-member
    ensemble workload: L = 16 pi, Re in [150, 220], scl in [0.05, 0.4], seeds 1000..1255): after
    10 steps every member equals its single-model run bit for bit."""
    rng = np.random.default_rng(1000)
    res = rng.uniform(150.0, 220.0, 256)
    scls = rng.uniform(0.05, 0.4, 256)
    ps = [mw.make_params(4, 64, 64, 16 * math.pi, 16 * math.pi, float(res[i]), 0.02) for i in range(256)]
    ics = [mw.ic_gaussian(ps[i], 1000 + i, float(scls[i]), 8.0) for i in range(256)]
    ens = mw.Ensemble(ps)
    assert ens.nclasses() == 1
    for i in range(256):
        ens.set_state(i, *ics[i])
    ens.step(10)
    e = ens.energies()
    single = mw.Velocity(ps[0])
    for i in range(256):
        single.close()
        single = mw.Velocity(ps[i])
        single.set_state(*ics[i])
        single.vel_TS(10)
        a, b = ens.get_state(i), single.get_state()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"member {i} differs from single model"
        assert e[i] == single.vel_energy()
    single.close()
    ens.close()


def test_tracking_turbulent_fraction_and_energy_on_device():
    """Per-member energy and turbulent fraction recorded on the device every
    5 steps (utils/turbfrac.f90's count), across three size classes stepped as
    parallel graph branches, equal the single-model diagnostics at the same
    steps, and the device trace buffers survive several drains."""
    ms = members()[:4]
    ps = [m[0] for m in ms]
    ics = [mw.ic_gaussian(p, s, scl, k0) for p, s, scl, k0 in ms]
    ens = mw.Ensemble(ps)
    for i, ic in enumerate(ics):
        ens.set_state(i, *ic)
    ens.track(5, 1e-12)
    ens.track_turb(6, 6, 0.02)
    ens.step(7)
    ens.step(13)  # 20 steps: samples at 0, 5, 10, 15, 20
    for i, p in enumerate(ps):
        e, rel = ens.trace(i)
        f = ens.trace_turb(i)
        assert len(e) == 5 and len(f) == 5 and rel == -1
        v = mw.Velocity(p)
        v.set_state(*ics[i])
        for j in range(5):
            if j:
                v.vel_TS(5)
            assert e[j] == v.vel_energy(), (i, j)
            assert f[j] == v.turb_fraction(6, 6, 0.02), (i, j)
        v.close()
    ens.close()


def test_tracking_buffers_drain_beyond_capacity():
    """More samples than the device buffers hold (256): the host drains them
    between graph launches, and every sample arrives in order."""
    p = mw.make_params(4, 16, 16, 4 * math.pi, 4 * math.pi, 150.0, 0.02)
    ens = mw.Ensemble([p, p])
    ic = mw.ic_gaussian(p, 3, 0.1, 4.0)
    ens.set_state(0, *ic)
    ens.set_state(1, *ic)
    ens.track(1, 0.0)
    ens.step(600)
    e0, _ = ens.trace(0)
    e1, _ = ens.trace(1)
    assert len(e0) == 601 and np.array_equal(e0, e1)
    v = mw.Velocity(p)
    v.set_state(*ic)
    v.vel_TS(600)
    assert e0[-1] == v.vel_energy()
    v.close()
    ens.close()
