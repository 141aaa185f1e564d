"""Model configurations used by the tests and the benchmark.

The five workloads are the ones the benchmark is defined on: Model Waleffe
Flow with the M4/M6 wall-normal bases, periodic box Lx = Lz, seeded
Gaussian-spectrum initial conditions (mwf_ic_gaussian), parameters.f90
defaults elsewhere (drag 1e-2, dragall, no hypo-viscosity, theta 0).
A workload of "N x N Fourier modes" means i_MM = i_NN = N/2.
"""
import math

import mwf_amd as mw

PI = math.pi

# name: (k0, mm, nn, L, Re, dt, seed, scl, k0_ic)
CASES = {
    "m4_64": (4, 32, 32, 8 * PI, 200.0, 0.02, 1, 0.1, 4.0),        # 64^2 modes, 96^2 grid
    "m4_1024": (4, 512, 512, 128 * PI, 180.0, 0.02, 2, 0.3, 8.0),  # 1024^2 modes, 1536^2 grid
    "m6_2048": (6, 1024, 1024, 256 * PI, 250.0, 0.01, 3, 0.3, 8.0),
    "m4_128_member": (4, 64, 64, 16 * PI, 185.0, 0.02, 1000, 0.2, 8.0),
}


def params(name, precision=mw.MWF_FP64, **over):
    k0, mm, nn, L, re, dt, seed, scl, k0ic = CASES[name]
    d = dict(k0=k0, mm=mm, nn=nn, lx=L, lz=L, re=re, dt=dt, precision=precision)
    d.update(over)
    return mw.make_params(**d)


def ic_of(name):
    k0, mm, nn, L, re, dt, seed, scl, k0ic = CASES[name]
    return seed, scl, k0ic
