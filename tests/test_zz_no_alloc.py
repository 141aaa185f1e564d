"""GPU: the entry point that allocates nothing (mwf_energy_noalloc) against
the one it stands in for (mwf_energy), through the C-ABI: calls in a row,
each on a smaller handle than the one before, each bit-identical to
mwf_energy.  The device output lives in memory the test allocates with the
HIP runtime that libmwf itself uses (torch ships its own runtime, which
cannot share the device with it in one process)."""
import ctypes
import math

import pytest

import mwf_amd as mw

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libmwf.so is linked against (loaded with it)."""
    for path in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
        try:
            H = ctypes.CDLL(path)
            break
        except OSError:
            continue
    else:
        raise RuntimeError("libamdhip64 not found")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return H


def test_energy_noalloc_matches_energy():
    L = mw.lib()
    L.mwf_energy_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    H = _hip()
    dev = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dev), 8 * 4) == 0
    host = (ctypes.c_double * 4)()
    D2H = 2  # hipMemcpyDeviceToHost
    try:
        # (k0, mm, nn, precision): each handle is smaller than the one before it
        for i, (k0, mm, nn, prec) in enumerate([(6, 1024, 1024, mw.MWF_FP64), (6, 512, 256, mw.MWF_FP64),
                                                (4, 256, 256, mw.MWF_MIXED), (4, 16, 16, mw.MWF_MIXED)]):
            p = mw.make_params(k0, mm, nn, mm * math.pi / 4, nn * math.pi / 4, 200.0, 0.02, precision=prec)
            v = mw.Velocity(p)
            v.set_state(*mw.ic_gaussian(p, 11 + i, 0.2, 4.0))
            v.vel_TS(3)
            print(f"[noalloc] call {i}: k0={k0} mm={mm} nn={nn} precision={prec}", flush=True)
            for rep in range(2):  # the shared scratch keeps nothing between calls
                if rep:
                    v.vel_TS(2)
                want = v.vel_energy()
                rc = L.mwf_energy_noalloc(v._h, ctypes.c_void_p(dev.value + 8 * i))
                assert rc == 0, L.mwf_last_error().decode()
                v.sync()
                assert H.hipMemcpy(ctypes.cast(host, ctypes.c_void_p), dev, 8 * 4, D2H) == 0
                assert host[i] == want, (i, rep, host[i], want)
            v.close()
    finally:
        H.hipFree(dev)
