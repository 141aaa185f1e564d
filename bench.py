"""Throughput of the MI355X Model Waleffe Flow step on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the M4 basis, 1024 x 1024 Fourier modes (i_MM = i_NN = 512, a
1536 x 1536 padded grid), Lx = Lz = 128 pi, Re = 180, dt = 0.02, seeded
Gaussian-spectrum initial condition (seed 2, scl 0.3 applied as the rms
perturbation velocity, k0 = 8), FP64.  A step
is one vel_TS (program/velocity.f90:21-41): nonlinear term through the
3/2-dealiased transforms, AB2 and the implicit update.  The state is resident
in HBM for the whole timed region.  The same workload is then run in the
mixed form (FP32 transforms and products, FP64 state/update/energy).

Prints one JSON line.  ``value`` is FP64 steps/s; ``roofline`` is the
dominant kernel's algorithmic bytes per launch over its mean event-timed
duration against 8 TB/s; ``cpu_baseline`` times the oracle (the repository's
C restatement of the reference path with its own FFT, FFTW being absent) on
a bounded sample of the same workload on the host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/sec FP64 vs mixed at M4 1024²; ensemble member-steps/sec (1 MI355X)"
PEAK_HBM_GBS = 8000.0
WORKLOAD = dict(k0=4, mm=512, nn=512, lx=128 * np.pi, lz=128 * np.pi, re=180.0, dt=0.02)
IC = dict(seed=2, scl=0.3, k0=8.0)
# turbulent-point threshold on epos for the ensembles' turbulent fractions
# (utils/turbfrac.f90 reads it from input): 1e-2, 2 % of the laminar energy 0.5
TURB_CT = 1e-2


def step_bytes(mw, p):
    """Algorithmic HBM bytes of one step of one member (the two-pass model):
    state c read and written (complex FP64), AB2 history read and written (in
    the transform precision), and one HBM round trip of the x-physical /
    z-spectral intermediates: 3 KK lines of P x NN complex values each way
    (written once, read once)."""
    d = mw.dims(p)
    si = 8 if p.precision == mw.MWF_MIXED else 16
    mpt = d["K"] * d["M"] * p.nn
    return 2 * 16 * mpt + 2 * si * mpt + 2 * 3 * d["KK"] * d["P"] * p.nn * si


def run_gpu(mw, precision, steps, warmup, ic, workload=WORKLOAD, event_steps=200):
    p = mw.make_params(precision=precision, **workload)
    v = mw.Velocity(p)
    plan = v.plan_info()
    v.set_state(*ic)
    # the first step after a state upload also rebuilds U (inverse-only x pass)
    v.vel_TS(1)
    # per-kernel durations first: max(steps, event_steps) steps with HIP events
    # around every kernel on the library's stream (the events cost a few
    # microseconds per launch, so they stay out of the timed steps).  Run
    # before the warm-up, this pass also brings the device to the clocks of a
    # running simulation: after the idle gap of the state upload, a handful of
    # warm-up steps leave the first milliseconds of a short timed run slow.
    v.profile(True)
    v.vel_TS(max(steps, event_steps))
    v.sync()
    kern = {}
    for kid, name in mw.KERNEL_NAMES.items():
        ms, cnt = v.profile_read(kid)
        if cnt:
            kern[name] = dict(kid=kid, total_ms=ms, count=cnt, avg_ms=ms / cnt,
                              bytes=v.kernel_bytes(kid))
    v.profile(False)
    v.vel_TS(warmup)
    v.sync()
    # the timed steps run as the product runs, with no events between kernels
    t0 = time.perf_counter()
    v.vel_TS(steps)
    v.sync()
    t1 = time.perf_counter()
    re, im = v.get_state()
    finite = bool(np.isfinite(re).all() and np.isfinite(im).all())
    e = v.vel_energy()
    v.close()
    return dict(seconds=t1 - t0, steps_per_s=steps / (t1 - t0), kernels=kern, finite=finite,
                energy=e, plan=plan)


def ensemble_m4_128_x256(mw, nsteps, seq_members=16, seq_steps=20, track_every=50):
    """256 M4 members, 128x128 modes, Lx=Lz=16pi, Re ~ U[150,220], rms ~ U[0.05,0.4],
    seeds 1000..1255, dt 0.02, FP64: batched member-steps/s over nsteps (5000 by
    default), with every member's energy traced every track_every steps and its
    relaminarisation step (perturbation energy below 1e-6 of the laminar 0.5)
    recorded on the device; then the same members stepped one after another
    through the single-model path."""
    rng = np.random.default_rng(1000)
    res = rng.uniform(150.0, 220.0, 256)
    scls = rng.uniform(0.05, 0.4, 256)
    ps = [mw.make_params(4, 64, 64, 16 * np.pi, 16 * np.pi, float(res[i]), 0.02) for i in range(256)]
    ens = mw.Ensemble(ps)
    ics = [mw.ic_gaussian(ps[i], 1000 + i, float(scls[i]), 8.0) for i in range(256)]
    for i in range(256):
        ens.set_state(i, *ics[i])
    ens.track(track_every, 1e-6 * 0.5)
    ens.track_turb(6, 6, TURB_CT)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(nsteps)
    ens.sync()
    t1 = time.perf_counter()
    batched = 256 * nsteps / (t1 - t0)
    e = ens.energies()
    relam, tfrac = [], []
    for i in range(256):
        trace, rstep = ens.trace(i)
        relam.append(rstep)
        tfrac.append(ens.trace_turb(i)[-1])
    relam, tfrac = np.array(relam), np.array(tfrac)
    ens.close()
    bpm = step_bytes(mw, ps[0])
    # sequential on the GPU: one model per member, one after another
    vs = []
    for i in range(seq_members):
        v = mw.Velocity(ps[i])
        v.set_state(*ics[i])
        v.vel_TS(2)
        vs.append(v)
    vs[-1].sync()
    t0 = time.perf_counter()
    for v in vs:
        v.vel_TS(seq_steps)
    for v in vs:
        v.sync()
    t1 = time.perf_counter()
    seq = seq_members * seq_steps / (t1 - t0)
    for v in vs:
        v.close()
    done = relam[relam >= 0]
    return {"workload": "256 x M4 128x128 modes (MM=NN=64), Lx=Lz=16pi, Re U[150,220], scl U[0.05,0.4] (as rms velocity), "
                        "seeds 1000..1255, dt 0.02, FP64",
            "member_steps_per_s": batched, "steps": nsteps,
            "sequential_member_steps_per_s": seq, "sequential_sample": f"{seq_members} members x {seq_steps} steps",
            "speedup_vs_sequential": batched / seq, "finite": bool(np.all(np.isfinite(e))),
            "graph": f"hipGraph replay of {track_every}-step chunks with the energy and turbulent-fraction "
                     f"samples inside the graph (device trace buffers)",
            "bytes_per_member_step": bpm,
            "roofline": {"bound": "hbm", "achieved_GBs": batched * bpm / 1e9, "peak": PEAK_HBM_GBS,
                         "frac": batched * bpm / 1e9 / PEAK_HBM_GBS},
            "turbulent_fraction_at_end": {"median": float(np.median(tfrac)), "min": float(tfrac.min()),
                                          "max": float(tfrac.max()), "xs_zs": [6, 6], "ct": TURB_CT},
            "relaminarised": int(done.size), "turbulent_at_end": int(256 - done.size),
            "relam_time_median": float(np.median(done) * 0.02) if done.size else None,
            "relam_time_min": float(done.min() * 0.02) if done.size else None,
            "relam_time_max": float(done.max() * 0.02) if done.size else None}


def config_m6_2048(mw, steps, err_steps, warmup=3):
    """M6 basis, 2048x2048 modes (MM=NN=1024, 3072^2 grid), Lx=Lz=256pi, Re=250,
    dt=0.01, seed 3 (scl 0.3 as rms velocity, k0 8): FP64 and mixed steps/s, then both forms
    stepped side by side from the same state with the per-step energy error
    and the state error of the mixed run against the FP64 run."""
    wl = dict(k0=6, mm=1024, nn=1024, lx=256 * np.pi, lz=256 * np.pi, re=250.0, dt=0.01)
    p = mw.make_params(**wl)
    ic = mw.ic_gaussian(p, 3, 0.3, 8.0)
    f64 = run_gpu(mw, mw.MWF_FP64, steps, warmup, ic, wl)
    mix = run_gpu(mw, mw.MWF_MIXED, steps, warmup, ic, wl)
    v64 = mw.Velocity(mw.make_params(precision=mw.MWF_FP64, **wl))
    v32 = mw.Velocity(mw.make_params(precision=mw.MWF_MIXED, **wl))
    v64.set_state(*ic)
    v32.set_state(*ic)
    e_err, s_err = [], []
    for s in range(1, err_steps + 1):
        v64.vel_TS(1)
        v32.vel_TS(1)
        e64, e32 = v64.vel_energy(), v32.vel_energy()
        e_err.append(abs(e32 - e64) / abs(e64))
        if s % max(1, err_steps // 4) == 0:
            a, b = v64.get_state(), v32.get_state()
            num = np.sqrt(((a[0] - b[0]) ** 2).sum() + ((a[1] - b[1]) ** 2).sum())
            den = np.sqrt((a[0] ** 2).sum() + (a[1] ** 2).sum())
            s_err.append([s, float(num / den)])
    v64.close()
    v32.close()
    e_err = np.array(e_err)
    return {"workload": "M6 basis, 2048x2048 modes (MM=NN=1024, 3072^2 grid), Lx=Lz=256pi, Re=250, dt=0.01, "
                        "seed 3 (scl 0.3 as rms velocity, k0 8)",
            "fp64_steps_per_s": f64["steps_per_s"], "mixed_steps_per_s": mix["steps_per_s"],
            "speedup_mixed": mix["steps_per_s"] / f64["steps_per_s"], "steps": steps,
            "fp64_kernels_ms": {n: round(k["avg_ms"], 4) for n, k in f64["kernels"].items()},
            "mixed_kernels_ms": {n: round(k["avg_ms"], 4) for n, k in mix["kernels"].items()},
            "fp64_plan": f64["plan"], "mixed_plan": mix["plan"],
            "mixed_vs_fp64": {"steps": err_steps, "energy_rel_err_per_step_max": float(e_err.max()),
                              "energy_rel_err_at_steps": {str(k): float(e_err[k - 1]) for k in
                                                          (1, 10, err_steps // 2, err_steps) if k <= err_steps},
                              "state_rel_l2_at_steps": s_err},
            "finite": f64["finite"] and mix["finite"]}


def config_m4_64(mw, steps):
    """M4, 64x64 modes (96^2 grid), Lx=Lz=8pi, Re=200, dt=0.02, seed 1 (scl 0.1 as rms,
    k0 4), FP64: the reference's CPU-sized case, steps/s on the GPU."""
    wl = dict(k0=4, mm=32, nn=32, lx=8 * np.pi, lz=8 * np.pi, re=200.0, dt=0.02)
    ic = mw.ic_gaussian(mw.make_params(**wl), 1, 0.1, 4.0)
    r = run_gpu(mw, mw.MWF_FP64, steps, 10, ic, wl)
    return {"workload": "M4, 64x64 modes (96^2 grid), Lx=Lz=8pi, Re=200, dt=0.02, seed 1 (scl 0.1 as rms velocity, k0 4), FP64",
            "steps_per_s": r["steps_per_s"], "steps": steps, "finite": r["finite"], "energy": r["energy"]}


def ensemble_uneven_x512(mw, nsteps, subset=32, track_every=50):
    """512 members, M4:M6 = 3:1, sizes {64,128,256,512}^2 modes, box 0.4 pi per
    mode, Re ~ U[150,300], seeds 5000..5511, dt 0.01, mixed form; energy error
    of a 32-member FP64 subset after the same steps."""
    rng = np.random.default_rng(5000)
    sizes = rng.choice([64, 128, 256, 512], 512)
    k0s = np.where(rng.uniform(size=512) < 0.25, 6, 4)
    res = rng.uniform(150.0, 300.0, 512)
    ps, ps64, ics = [], [], []
    for i in range(512):
        nmodes = int(sizes[i])
        L = 0.4 * np.pi * nmodes
        kw = dict(k0=int(k0s[i]), mm=nmodes // 2, nn=nmodes // 2, lx=L, lz=L, re=float(res[i]), dt=0.01)
        ps.append(mw.make_params(precision=mw.MWF_MIXED, **kw))
        ps64.append(mw.make_params(precision=mw.MWF_FP64, **kw))
        ics.append(mw.ic_gaussian(ps[i], 5000 + i, 0.2, 8.0))
    ens = mw.Ensemble(ps)
    for i in range(512):
        ens.set_state(i, *ics[i])
    ens.track(track_every, 1e-6 * 0.5)
    ens.track_turb(6, 6, TURB_CT)
    ens.step(1)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(nsteps)
    ens.sync()
    t1 = time.perf_counter()
    e_mixed = ens.energies()
    ncls = ens.nclasses()
    relam = np.array([ens.trace(i)[1] for i in range(512)])
    tfrac = np.array([ens.trace_turb(i)[-1] for i in range(512)])
    ens.close()
    bytes_step = sum(step_bytes(mw, p) for p in ps)
    sub = list(range(subset))
    e64 = mw.Ensemble([ps64[i] for i in sub])
    for j, i in enumerate(sub):
        e64.set_state(j, *ics[i])
    e64.step(1 + nsteps)
    ref = e64.energies()
    e64.close()
    err = np.abs(e_mixed[sub] - ref) / np.abs(ref)
    return {"workload": "512 members, M4:M6 = 3:1, sizes {64,128,256,512}^2 modes, L = 0.4pi per mode, "
                        "Re U[150,300], seeds 5000..5511, dt 0.01, mixed FP32/FP64",
            "member_steps_per_s": 512 * nsteps / (t1 - t0), "steps": nsteps, "size_classes": ncls,
            "bytes_per_member_step_mean": bytes_step / 512,
            "roofline": {"bound": "hbm", "achieved_GBs": bytes_step * nsteps / (t1 - t0) / 1e9,
                         "peak": PEAK_HBM_GBS, "frac": bytes_step * nsteps / (t1 - t0) / 1e9 / PEAK_HBM_GBS},
            "graph": "size classes as parallel branches of one hipGraph per chunk; energy and turbulent-fraction "
                     f"samples every {track_every} steps on the device",
            "relaminarised": int((relam >= 0).sum()),
            "turbulent_fraction_at_end": {"median": float(np.median(tfrac)), "min": float(tfrac.min()),
                                          "max": float(tfrac.max()), "xs_zs": [6, 6], "ct": TURB_CT},
            "fp64_subset_energy_rel_err_max": float(err.max()),
            "fp64_subset_energy_rel_err_median": float(np.median(err)),
            "finite": bool(np.all(np.isfinite(e_mixed)))}


def cpu_baseline(ic, nsteps, single_steps=1):
    """The oracle (the repository's C restatement of the reference path, with
    its own FFT: FFTW is absent offline) timed on the host on a bounded sample
    of the same workload: nsteps steps on all host cores (at most 16 threads,
    OpenMP over x/z lines), then single_steps on one core (the reference
    itself is single-threaded per MPI rank)."""
    import math
    import oracle
    w = WORKLOAD
    o = oracle.Oracle(w["k0"], w["mm"], w["nn"], 2 * math.pi / w["lx"], 2 * math.pi / w["lz"],
                      w["re"], w["dt"])
    threads = min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    o.set_state(*ic)
    o.step(1)  # first step (forward Euler) untimed
    t0 = time.perf_counter()
    o.step(nsteps)
    t1 = time.perf_counter()
    oracle.set_threads(1)
    t2 = time.perf_counter()
    o.step(single_steps)
    t3 = time.perf_counter()
    return dict(value=nsteps / (t1 - t0), unit="steps/s", cores=threads, kind="port",
                single_core_value=single_steps / (t3 - t2),
                sample=f"{nsteps} steps of the same workload (M4 1024^2 modes, FP64) on {threads} threads after "
                       f"one untimed step, then {single_steps} step(s) on 1 thread; oracle = C restatement of "
                       f"the reference path with its own FFT (FFTW 3.3.8 unavailable offline)")


def calib_copy(mib):
    """A plain streaming kernel of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of rocprofv3: torch's elementwise negate over mib MiB
    of float64 reads mib MiB and writes mib MiB per launch (three launches)."""
    import torch
    n = (mib << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    assert float(y[n - 1]) == -1.0
    print(json.dumps({"calib_copy_mib": mib}), flush=True)


PROFILE_CASES = {
    "m4_1024": (WORKLOAD, IC),
    "m6_2048": (dict(k0=6, mm=1024, nn=1024, lx=256 * np.pi, lz=256 * np.pi, re=250.0, dt=0.01),
                dict(seed=3, scl=0.3, k0=8.0)),
}


def profile_case(name, steps, warmup, no_mixed, tune):
    """A short run of one workload for rocprofv3: FP64, then (unless no_mixed)
    the mixed form, each with its own handle."""
    import mwf_amd as mw
    if tune is not None:
        mw.autotune(tune)
    wl, ic = PROFILE_CASES[name]
    p = mw.make_params(**wl)
    state = mw.ic_gaussian(p, ic["seed"], ic["scl"], ic["k0"])
    out = {"profile_case": name}
    for prec in [mw.MWF_FP64] + ([] if no_mixed else [mw.MWF_MIXED]):
        r = run_gpu(mw, prec, steps, warmup, state, wl, event_steps=0)
        out["fp64" if prec == mw.MWF_FP64 else "mixed"] = {
            "steps_per_s": r["steps_per_s"], "plan": r["plan"], "finite": r["finite"],
            "kernels_ms": {n: round(k["avg_ms"], 5) for n, k in r["kernels"].items()}}
    print(json.dumps(out), flush=True)


def traffic_summary(kernel):
    """HBM bytes per launch of one FP64 kernel of the M4 1024^2 step from the
    committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE scaled by the
    copy-kernel calibration), if any."""
    path = os.path.join(ROOT, "profiles", "k_zc_k_xstep_m4_1024_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel, {}).get("fp64")
    if not k:
        return None
    return {"hbm_bytes_per_launch": (k["read_MB"] + k["write_MB"]) * 1e6, "source": d["source"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-steps", type=int, default=5000)
    ap.add_argument("--m6-steps", type=int, default=500)
    ap.add_argument("--m6-err-steps", type=int, default=100)
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="rule-based kernel plans instead of timed ones")
    ap.add_argument("--plan", type=int, default=None,
                    help="force kernel plan bits (1 k_zb, 2 split x, 4 k_z2) instead of timing the plans")
    ap.add_argument("--uneven-steps", type=int, default=2000)
    ap.add_argument("--calib-copy", type=int, default=0, help="only run a MiB-sized device copy")
    ap.add_argument("--profile-case", choices=sorted(PROFILE_CASES), default=None,
                    help="only step one workload (FP64, then mixed) for --steps steps after --warmup, "
                         "for profiling under rocprofv3")
    args = ap.parse_args()
    if args.calib_copy:
        calib_copy(args.calib_copy)
        return
    if args.profile_case:
        profile_case(args.profile_case, args.steps, args.warmup, args.no_mixed,
                     None if args.no_tune else (2 + args.plan if args.plan is not None else 1))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import mwf_amd as mw
    if args.plan is not None:
        mw.autotune(2 + args.plan)
    elif not args.no_tune:
        mw.autotune(1)  # timed plan per size class at handle creation (tra_precompute's role)
    p = mw.make_params(**WORKLOAD)
    ic = mw.ic_gaussian(p, IC["seed"], IC["scl"], IC["k0"])

    fp64 = run_gpu(mw, mw.MWF_FP64, args.steps, args.warmup, ic)
    mixed = None if args.no_mixed else run_gpu(mw, mw.MWF_MIXED, args.steps, args.warmup, ic)

    kern = fp64["kernels"]
    dom_name, dom = max(kern.items(), key=lambda kv: kv[1]["total_ms"])
    achieved = dom["bytes"] / (dom["avg_ms"] * 1e-3) / 1e9
    step_bytes = sum(k["bytes"] for n, k in kern.items() if n != "energy")
    line = {
        "metric": METRIC,
        "value": fp64["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / fp64["steps_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "M4 basis, 1024x1024 Fourier modes (MM=NN=512, 1536^2 padded grid), "
                               "Lx=Lz=128pi, Re=180, dt=0.02, seeded Gaussian IC (seed 2, scl 0.3 as rms velocity, k0 8)",
                   "form": "FP64"},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None,
                     "bytes_per_launch": dom["bytes"], "avg_ms": dom["avg_ms"]},
        "step_roofline": {"bytes_per_step": step_bytes,
                          "achieved_GBs": step_bytes * fp64["steps_per_s"] / 1e9,
                          "frac": step_bytes * fp64["steps_per_s"] / 1e9 / PEAK_HBM_GBS},
        "kernels_ms": {n: round(k["avg_ms"], 5) for n, k in kern.items()},
        "plan": fp64["plan"],
        "finite": fp64["finite"],
    }
    tr = traffic_summary(dom_name)
    if tr:
        line["roofline"]["traffic"] = tr["hbm_bytes_per_launch"]
        line["roofline"]["traffic_source"] = tr["source"]
    if mixed:
        line["mixed"] = {"steps_per_s": mixed["steps_per_s"],
                         "speedup_vs_fp64": mixed["steps_per_s"] / fp64["steps_per_s"],
                         "plan": mixed["plan"],
                         "kernels_ms": {n: round(k["avg_ms"], 5) for n, k in mixed["kernels"].items()},
                         "energy_rel_err_vs_fp64": abs(mixed["energy"] - fp64["energy"]) / abs(fp64["energy"]),
                         "finite": mixed["finite"]}
    if not args.no_configs:
        line["configs"] = {"m4_64": config_m4_64(mw, 1000),
                           "m6_2048": config_m6_2048(mw, args.m6_steps, args.m6_err_steps)}
    if not args.no_ensemble:
        line["ensemble"] = {"m4_128_x256": ensemble_m4_128_x256(mw, args.ens_steps),
                            "uneven_x512_mixed": ensemble_uneven_x512(mw, args.uneven_steps)}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ic, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
