"""CPU restatement of the reference per-step path -- the oracle.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP library in
``mwf_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product never does.

Arrays use the reference's memory layout (Fortran column-major).  In numpy
they are C-ordered with the axes reversed:

* mpt  ``(NN, M, K)``    == Fortran ``mpt%Re(K, 0:M-1, 0:NN-1)``
* spec ``(NN, M, KK)``   == Fortran ``spec%Re(KK, 0:M-1, 0:NN-1)``
* phys ``(P, Q, KK)``    == Fortran ``phys%Re(KK, 0:Q-1, 0:P-1)``

(program/variables.f90:8-20).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [("k0", ctypes.c_int), ("mm", ctypes.c_int), ("nn", ctypes.c_int),
                ("dragall", ctypes.c_int), ("alpha", ctypes.c_double),
                ("gamma", ctypes.c_double), ("re", ctypes.c_double),
                ("dt", ctypes.c_double), ("drag", ctypes.c_double),
                ("vdrag", ctypes.c_double), ("hypo", ctypes.c_double),
                ("theta", ctypes.c_double)]


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_ref():
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        for name in ("orc_set_state", "orc_put_state", "orc_get_state", "orc_nonlinear", "orc_lhsrhs",
                     "orc_addlam"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _dp, _dp][: 2 if name == "orc_addlam" else 3]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_energy.restype = ctypes.c_double
        L.orc_energy.argtypes = [ctypes.c_void_p]
        L.orc_mpt2phys.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_mpt2spec.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_spec2mpt.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_spec_grad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_spec2phys.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_phys2spec.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_udotgradu.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_nluw.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_nlvv.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_nluv.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_nluyv.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_nlvvy.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.orc_epos.restype = ctypes.c_double
        L.orc_epos.argtypes = [ctypes.c_int, _dp]
        L.orc_dft.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


BETA = 3.1415926535897931 / 2.0


def set_threads(n):
    lib().orc_set_threads(int(n))


class Oracle:
    """One model instance: parameters as in program/parameters.f90."""

    def __init__(self, k0, mm, nn, alpha, gamma, re, dt, drag=1e-2, vdrag=0.0,
                 hypo=0.0, theta=0.0, dragall=True):
        self.k0, self.mm, self.nn = int(k0), int(mm), int(nn)
        self.K, self.KK = 2 * k0 - 1, 3 * k0 - 1
        self.M, self.P, self.Q = 2 * mm - 1, 3 * mm, 3 * nn
        self.params = OrcParams(self.k0, self.mm, self.nn, int(bool(dragall)), alpha, gamma,
                                re, dt, drag, vdrag, hypo, theta)
        self._h = lib().orc_create(ctypes.byref(self.params))
        if not self._h:
            raise ValueError("oracle: unsupported parameters")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    # shapes
    @property
    def mpt_shape(self):
        return (self.nn, self.M, self.K)

    @property
    def spec_shape(self):
        return (self.nn, self.M, self.KK)

    @property
    def phys_shape(self):
        return (self.P, self.Q, self.KK)

    def set_state(self, re, im):
        re, im = _f64(re).reshape(self.mpt_shape), _f64(im).reshape(self.mpt_shape)
        lib().orc_set_state(self._h, _p(re), _p(im))

    def put_state(self, re, im):
        """Replace the state keeping the AB2 history (vel_imposesym)."""
        re, im = _f64(re).reshape(self.mpt_shape), _f64(im).reshape(self.mpt_shape)
        lib().orc_put_state(self._h, _p(re), _p(im))

    def get_state(self):
        re, im = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        lib().orc_get_state(self._h, _p(re), _p(im))
        return re, im

    def step(self, n=1):
        lib().orc_step(self._h, int(n))

    def energy(self):
        return lib().orc_energy(self._h)

    def nonlinear(self):
        re, im = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        lib().orc_nonlinear(self._h, _p(re), _p(im))
        return re, im

    def lhsrhs(self):
        l, r = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        lib().orc_lhsrhs(self._h, _p(l), _p(r))
        return l, r

    def mpt2spec(self, re, im):
        re, im = _f64(re), _f64(im)
        sr, si = np.empty(self.spec_shape), np.empty(self.spec_shape)
        lib().orc_mpt2spec(self._h, _p(re), _p(im), _p(sr), _p(si))
        return sr, si

    def addlam(self, sre):
        sre = _f64(sre).copy()
        lib().orc_addlam(self._h, _p(sre))
        return sre

    def spec2mpt(self, sre, sim):
        sre, sim = _f64(sre), _f64(sim)
        mr, mi = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        lib().orc_spec2mpt(self._h, _p(sre), _p(sim), _p(mr), _p(mi))
        return mr, mi

    def spec_grad(self, sre, sim):
        sre, sim = _f64(sre), _f64(sim)
        out = [np.empty(self.spec_shape) for _ in range(4)]
        lib().orc_spec_grad(self._h, _p(sre), _p(sim), *[_p(a) for a in out])
        return out

    def spec2phys(self, sre, sim):
        sre, sim = _f64(sre), _f64(sim)
        ph = np.empty(self.phys_shape)
        lib().orc_spec2phys(self._h, _p(sre), _p(sim), _p(ph))
        return ph

    def phys2spec(self, ph):
        ph = _f64(ph)
        sr, si = np.empty(self.spec_shape), np.empty(self.spec_shape)
        lib().orc_phys2spec(self._h, _p(ph), _p(sr), _p(si))
        return sr, si

    def mpt2phys(self, re, im):
        re, im = _f64(re), _f64(im)
        ph = np.empty(self.phys_shape)
        lib().orc_mpt2phys(self._h, _p(re), _p(im), _p(ph))
        return ph


# ---- Galerkin closures (program/modes.M4.f90, program/modes.M6.f90) ----

def udotgradu(k0, v, vx, vz, beta=BETA):
    v, vx, vz = _f64(v), _f64(vx), _f64(vz)
    ans = np.empty(3 * k0 - 1)
    lib().orc_udotgradu(k0, beta, _p(v), _p(vx), _p(vz), _p(ans))
    return ans


def nluw(k0, u, w):
    u, w = _f64(u), _f64(w)
    ans = np.empty(k0)
    lib().orc_nluw(k0, _p(u), _p(w), _p(ans))
    return ans


def nlvv(k0, v):
    v = _f64(v)
    ans = np.empty(k0)
    lib().orc_nlvv(k0, _p(v), _p(ans))
    return ans


def nluv(k0, u, v):
    u, v = _f64(u), _f64(v)
    ans = np.empty(k0 - 1)
    lib().orc_nluv(k0, _p(u), _p(v), _p(ans))
    return ans


def nluyv(k0, u, v, beta=BETA):
    u, v = _f64(u), _f64(v)
    ans = np.empty(k0)
    lib().orc_nluyv(k0, beta, _p(u), _p(v), _p(ans))
    return ans


def nlvvy(k0, v, beta=BETA):
    v = _f64(v)
    ans = np.empty(k0 - 1)
    lib().orc_nlvvy(k0, beta, _p(v), _p(ans))
    return ans


def epos(k0, u):
    u = _f64(u)
    return lib().orc_epos(k0, _p(u))


def dft(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    n = x.size
    inp = np.empty(2 * n)
    inp[0::2], inp[1::2] = x.real, x.imag
    out = np.empty(2 * n)
    lib().orc_dft(n, int(sign), _p(inp), _p(out))
    return out[0::2] + 1j * out[1::2]
