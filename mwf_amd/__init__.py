"""MI355X-native Model Waleffe Flow timestepper (Python host mirror).

The compute runs in ``libmwf.so`` (hand-written HIP for gfx950, C-ABI in
``include/mwf.h``).  This module binds it with ctypes and mirrors the
reference's ``velocity`` module interface (program/velocity.f90):

=====================  ================================================
reference              here
=====================  ================================================
``vel_precompute``     ``Velocity(params)``
``vel_c`` (mpt)        ``Velocity.vel_c`` (get) / ``set_state``
``vel_TS``             ``Velocity.vel_TS(nsteps=1)``
``vel_nonlinear``      ``Velocity.vel_nonlinear()`` (returns vel_nl)
``vel_mpt2phys``       ``Velocity.vel_mpt2phys(re, im)``
``vel_imposesym``      ``Velocity.put_state`` (replace vel_c, keep AB2)
``tra_spec2phys``      ``Velocity.tra_spec2phys(re, im)``
``tra_phys2spec``      ``Velocity.tra_phys2spec(phys)``
``vel_energy``         ``Velocity.vel_energy()`` (of the current vel_c)
``vel_onl`` (restart)  ``Velocity.get_onl()`` / ``set_onl``
``io_save_Uspec``      ``Velocity.spectrum_max()``
``vel_history``        ``Velocity.vel_history(nh, y)``
``turb_sample``        ``Velocity.turb_sample(dm, dn)``
``var_doubleX/quadX``  ``Velocity.extend_x_into(dst, factor)``
``interpstate``        ``Velocity.blend_into(a, b, d)``
=====================  ================================================

mpt arrays are numpy float64 of shape ``(NN, M, K)`` -- the memory layout of
the reference's ``mpt%Re(K, 0:M-1, 0:NN-1)``; phys arrays ``(P, Q, KK)``.
There is no CPU fallback: if the library is missing or the GPU call fails,
an ``MWFError`` is raised.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libmwf.so")

MWF_FP64 = 0
MWF_MIXED = 1

KID_XINV, KID_Z, KID_XFWD, KID_ENERGY, KID_XSTEP = 0, 1, 2, 3, 4
KERNEL_NAMES = {KID_XINV: "k_xinv", KID_Z: "k_z", KID_XFWD: "k_xfwd", KID_ENERGY: "energy",
                KID_XSTEP: "k_xstep"}


class MWFError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """mwf_params of include/mwf.h (program/parameters.f90 as run-time values)."""
    _fields_ = [("k0", ctypes.c_int32), ("mm", ctypes.c_int32), ("nn", ctypes.c_int32),
                ("dragall", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("alpha", ctypes.c_double),
                ("gamma", ctypes.c_double), ("re", ctypes.c_double), ("dt", ctypes.c_double),
                ("drag", ctypes.c_double), ("vdrag", ctypes.c_double), ("hypo", ctypes.c_double),
                ("theta", ctypes.c_double)]


def make_params(k0, mm, nn, lx, lz, re, dt, drag=1e-2, vdrag=0.0, hypo=0.0, thdeg=0.0,
                dragall=True, precision=MWF_FP64):
    """Parameters with program/parameters.f90 defaults (drag 1e-2, dragall, no
    hypo-viscosity, theta = 0)."""
    return Params(int(k0), int(mm), int(nn), int(bool(dragall)), int(precision), 0,
                  2.0 * math.pi / lx, 2.0 * math.pi / lz, float(re), float(dt), float(drag),
                  float(vdrag), float(hypo), thdeg / 180.0 * math.pi)


_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Load libmwf.so (built by mwf_amd.build) -- raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise MWFError(f"{LIBPATH} is not built; run `python -m mwf_amd.build`")
        L = ctypes.CDLL(LIBPATH)
        vp = ctypes.c_void_p
        L.mwf_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.mwf_destroy.argtypes = [vp]
        L.mwf_set_state.argtypes = [vp, _dp, _dp]
        L.mwf_get_state.argtypes = [vp, _dp, _dp]
        L.mwf_put_state.argtypes = [vp, _dp, _dp]
        L.mwf_spec2phys.argtypes = [vp, _dp, _dp, _dp]
        L.mwf_phys2spec.argtypes = [vp, _dp, _dp, _dp]
        L.mwf_step.argtypes = [vp, ctypes.c_int32]
        L.mwf_energy.argtypes = [vp, _dp]
        L.mwf_nonlinear.argtypes = [vp, _dp, _dp]
        L.mwf_mpt2phys.argtypes = [vp, _dp, _dp, _dp]
        L.mwf_sync.argtypes = [vp]
        L.mwf_last_error.restype = ctypes.c_char_p
        L.mwf_ic_gaussian.argtypes = [ctypes.POINTER(Params), ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_double, _dp, _dp]
        L.mwf_profile_enable.argtypes = [vp, ctypes.c_int32]
        L.mwf_profile_read.argtypes = [vp, ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.mwf_kernel_bytes.argtypes = [vp, ctypes.c_int32, _dp]
        L.mwf_get_onl.argtypes = [vp, _dp, _dp, ctypes.POINTER(ctypes.c_int32)]
        L.mwf_set_onl.argtypes = [vp, _dp, _dp]
        L.mwf_spectrum_max.argtypes = [vp, _dp, _dp]
        L.mwf_vel_history.argtypes = [vp, ctypes.c_int32, ctypes.c_double, _dp]
        L.mwf_turb_sample.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _dp]
        L.mwf_turb_fraction.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _dp]
        L.mwf_extend_x.argtypes = [vp, ctypes.c_int32, vp]
        L.mwf_blend.argtypes = [vp, vp, ctypes.c_double, vp]
        L.mwf_autotune.argtypes = [ctypes.c_int32]
        L.mwf_plan_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise MWFError(f"libmwf error {rc}: {lib().mwf_last_error().decode()}")


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def autotune(mode=1):
    """Process-wide plan selection for handles created afterwards (mwf_autotune),
    the role of tra_precompute's FFTW_PATIENT planning: 0 = rule-based plans,
    1 = timed choice per size, 2..65 = force plan bits mode - 2."""
    _check(lib().mwf_autotune(int(mode)))


PLAN_ZB, PLAN_SPLIT_X, PLAN_Z2, PLAN_FLUX, PLAN_FLUX_W2, PLAN_ZB_SPLIT = 1, 2, 4, 8, 16, 32


def dims(p):
    k0, mm, nn = p.k0, p.mm, p.nn
    return dict(K=2 * k0 - 1, KK=3 * k0 - 1, M=2 * mm - 1, P=3 * mm, Q=3 * nn)


def ic_gaussian(p, seed, scl, k0):
    """Seeded Gaussian-spectrum divergence-free initial condition (mwf_ic_gaussian)."""
    d = dims(p)
    re = np.empty((p.nn, d["M"], d["K"]))
    im = np.empty_like(re)
    _check(lib().mwf_ic_gaussian(ctypes.byref(p), ctypes.c_uint64(seed), float(scl), float(k0),
                                 _p(re), _p(im)))
    return re, im


class Velocity:
    """One model instance on the GPU: the reference's module velocity."""

    def __init__(self, params):
        self.params = params
        d = dims(params)
        self.K, self.KK, self.M, self.P, self.Q = d["K"], d["KK"], d["M"], d["P"], d["Q"]
        self.NN = params.nn
        h = ctypes.c_void_p()
        _check(lib().mwf_create(ctypes.byref(params), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mwf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def mpt_shape(self):
        return (self.NN, self.M, self.K)

    def set_state(self, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64).reshape(self.mpt_shape)
        im = np.ascontiguousarray(im, dtype=np.float64).reshape(self.mpt_shape)
        _check(lib().mwf_set_state(self._h, _p(re), _p(im)))

    def get_state(self):
        re, im = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        _check(lib().mwf_get_state(self._h, _p(re), _p(im)))
        return re, im

    vel_c = property(get_state)

    def put_state(self, re, im):
        """Replace vel_c keeping the AB2 history (vel_imposesym's projections,
        program/velocity.f90:68-74)."""
        re = np.ascontiguousarray(re, dtype=np.float64).reshape(self.mpt_shape)
        im = np.ascontiguousarray(im, dtype=np.float64).reshape(self.mpt_shape)
        _check(lib().mwf_put_state(self._h, _p(re), _p(im)))

    @property
    def spec_shape(self):
        return (self.NN, self.M, self.KK)

    def tra_spec2phys(self, re, im):
        """tra_spec2phys (program/transform.fftw5.f90:64-87): spec (NN, M, KK)
        planes -> phys (P, Q, KK)."""
        re = np.ascontiguousarray(re, dtype=np.float64).reshape(self.spec_shape)
        im = np.ascontiguousarray(im, dtype=np.float64).reshape(self.spec_shape)
        ph = np.empty((self.P, self.Q, self.KK))
        _check(lib().mwf_spec2phys(self._h, _p(re), _p(im), _p(ph)))
        return ph

    def tra_phys2spec(self, ph):
        """tra_phys2spec (program/transform.fftw5.f90:122-145): phys -> spec planes."""
        ph = np.ascontiguousarray(ph, dtype=np.float64).reshape((self.P, self.Q, self.KK))
        re, im = np.empty(self.spec_shape), np.empty(self.spec_shape)
        _check(lib().mwf_phys2spec(self._h, _p(ph), _p(re), _p(im)))
        return re, im

    def vel_TS(self, nsteps=1):
        _check(lib().mwf_step(self._h, int(nsteps)))

    def vel_energy(self):
        e = ctypes.c_double()
        _check(lib().mwf_energy(self._h, ctypes.byref(e)))
        return e.value

    def vel_nonlinear(self):
        re, im = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        _check(lib().mwf_nonlinear(self._h, _p(re), _p(im)))
        return re, im

    def vel_mpt2phys(self, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64).reshape(self.mpt_shape)
        im = np.ascontiguousarray(im, dtype=np.float64).reshape(self.mpt_shape)
        ph = np.empty((self.P, self.Q, self.KK))
        _check(lib().mwf_mpt2phys(self._h, _p(re), _p(im), _p(ph)))
        return ph

    def sync(self):
        _check(lib().mwf_sync(self._h))

    # AB2 history for a bit-identical warm restart (program/velocity.f90:16,24-27)
    def get_onl(self):
        re, im = np.empty(self.mpt_shape), np.empty(self.mpt_shape)
        first = ctypes.c_int32()
        _check(lib().mwf_get_onl(self._h, _p(re), _p(im), ctypes.byref(first)))
        return re, im, bool(first.value)

    def set_onl(self, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64).reshape(self.mpt_shape)
        im = np.ascontiguousarray(im, dtype=np.float64).reshape(self.mpt_shape)
        _check(lib().mwf_set_onl(self._h, _p(re), _p(im)))

    # on-device diagnostics of the kept host I/O
    def spectrum_max(self):
        """io_save_Uspec (program/io.f90:345-395): (m_max[MM], n_max[NN])."""
        m, n = np.empty(self.params.mm), np.empty(self.NN)
        _check(lib().mwf_spectrum_max(self._h, _p(m), _p(n)))
        return m, n

    def vel_history(self, nh=32, y=0.0):
        """vel_history (program/velocity.f90:164-180): array (nh, 4) of
        z, u, v, w at x index 0 and height y."""
        ans = np.empty((nh, 4))
        _check(lib().mwf_vel_history(self._h, int(nh), float(y), _p(ans)))
        return ans

    def turb_sample(self, dm, dn):
        """turb_sample (program/turb.f90:37-47): array (SM, SN, 3) of eposC at
        x = m*dm, z = n*dn (the Fortran (3, SN, SM) layout)."""
        sm, sn = self.P // dm, self.Q // dn
        out = np.empty((sm, sn, 3))
        _check(lib().mwf_turb_sample(self._h, int(dm), int(dn), _p(out)))
        return out

    def turb_fraction(self, xs, zs, ct):
        """Turbulent fraction of io_writeVTK_txz (program/io.f90:550-617,
        utils/turbfrac.f90): points with epos > ct on the (xs, zs)-strided grid
        over (Q/zs + 1)(P/xs + 1)."""
        f = ctypes.c_double()
        _check(lib().mwf_turb_fraction(self._h, int(xs), int(zs), float(ct), ctypes.byref(f)))
        return f.value

    def plan_info(self):
        """Plan bits: PLAN_ZB (phase-sequential z kernel), PLAN_SPLIT_X (separate x kernels),
        PLAN_Z2 (two-points-per-thread z kernel), PLAN_FLUX (flux form), PLAN_FLUX_W2 (its
        two-wave z kernel), PLAN_ZB_SPLIT (the fold phases of a row in three blocks)."""
        ch = ctypes.c_int32()
        _check(lib().mwf_plan_info(self._h, ctypes.byref(ch)))
        return ch.value

    # large-box utilities (utils/doubleX.f90, quadX.f90, interpstate.f90)
    def extend_x_into(self, dst, factor):
        """dst (a Velocity with alpha / factor) gets this state with m -> factor*m."""
        _check(lib().mwf_extend_x(self._h, int(factor), dst._h))

    def blend_into(self, a, b, d):
        """This handle's state becomes (1 - d) a + d b."""
        _check(lib().mwf_blend(a._h, b._h, float(d), self._h))

    # per-kernel timing
    def profile(self, on=True):
        _check(lib().mwf_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, kid):
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        _check(lib().mwf_profile_read(self._h, kid, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def kernel_bytes(self, kid):
        b = ctypes.c_double()
        _check(lib().mwf_kernel_bytes(self._h, kid, ctypes.byref(b)))
        return b.value


class Ensemble:
    """Many independent models stepped together (mwf_ens_* in include/mwf.h).

    Members may differ in every parameter; members sharing (k0, mm, nn,
    precision) form a size class stepped by one launch per kernel, and the
    step sequence is replayed from a hipGraph."""

    def __init__(self, params_list):
        L = lib()
        vp = ctypes.c_void_p
        if not getattr(L, "_ens_typed", False):
            L.mwf_ens_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int32, ctypes.POINTER(vp)]
            L.mwf_ens_destroy.argtypes = [vp]
            L.mwf_ens_set_state.argtypes = [vp, ctypes.c_int32, _dp, _dp]
            L.mwf_ens_get_state.argtypes = [vp, ctypes.c_int32, _dp, _dp]
            L.mwf_ens_step.argtypes = [vp, ctypes.c_int32]
            L.mwf_ens_energy.argtypes = [vp, _dp]
            L.mwf_ens_sync.argtypes = [vp]
            L.mwf_ens_track.argtypes = [vp, ctypes.c_int32, ctypes.c_double]
            L.mwf_ens_trace.argtypes = [vp, ctypes.c_int32, _dp, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]
            L.mwf_ens_classes.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
            L.mwf_ens_track_turb.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]
            L.mwf_ens_trace_turb.argtypes = [vp, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
            L._ens_typed = True
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        _check(L.mwf_ens_create(arr, len(self.params), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().mwf_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.params)

    def mpt_shape(self, i):
        d = dims(self.params[i])
        return (self.params[i].nn, d["M"], d["K"])

    def set_state(self, i, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64).reshape(self.mpt_shape(i))
        im = np.ascontiguousarray(im, dtype=np.float64).reshape(self.mpt_shape(i))
        _check(lib().mwf_ens_set_state(self._h, int(i), _p(re), _p(im)))

    def get_state(self, i):
        re, im = np.empty(self.mpt_shape(i)), np.empty(self.mpt_shape(i))
        _check(lib().mwf_ens_get_state(self._h, int(i), _p(re), _p(im)))
        return re, im

    def step(self, nsteps=1):
        _check(lib().mwf_ens_step(self._h, int(nsteps)))

    def energies(self):
        e = np.empty(len(self.params))
        _check(lib().mwf_ens_energy(self._h, _p(e)))
        return e

    def sync(self):
        _check(lib().mwf_ens_sync(self._h))

    def track(self, every, threshold):
        _check(lib().mwf_ens_track(self._h, int(every), float(threshold)))

    def trace(self, i, cap=1 << 20):
        cnt, rel = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().mwf_ens_trace(self._h, int(i), None, 0, ctypes.byref(cnt), ctypes.byref(rel)))
        buf = np.empty(max(cnt.value, 1))
        _check(lib().mwf_ens_trace(self._h, int(i), _p(buf), cnt.value, ctypes.byref(cnt), ctypes.byref(rel)))
        return buf[:cnt.value], rel.value

    def track_turb(self, xs, zs, ct):
        """Add each member's turbulent fraction to the tracking samples."""
        _check(lib().mwf_ens_track_turb(self._h, int(xs), int(zs), float(ct)))

    def trace_turb(self, i):
        cnt = ctypes.c_int32()
        _check(lib().mwf_ens_trace_turb(self._h, int(i), None, 0, ctypes.byref(cnt)))
        buf = np.empty(max(cnt.value, 1))
        _check(lib().mwf_ens_trace_turb(self._h, int(i), _p(buf), cnt.value, ctypes.byref(cnt)))
        return buf[:cnt.value]

    def nclasses(self):
        n = ctypes.c_int32()
        _check(lib().mwf_ens_classes(self._h, ctypes.byref(n)))
        return n.value
