"""ctypes binding of the CPU checker (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker (or the timed CPU baseline).  The
product package convfpga_amd never imports this module.  See oracle.h for the
reference file:line each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile oracle/liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        f64p = ctypes.POINTER(ctypes.c_double)
        sz, i, u64, d = ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        sig = {
            "orc_fill_uniform": (None, [f32p, sz, u64]),
            "orc_fill_loguniform": (None, [f32p, sz, u64]),
            "orc_fill_ref01": (None, [f32p, sz, ctypes.c_uint]),
            "orc_fft3d_f64": (i, [i, f64p, i]),
            "orc_conv3d_f64": (i, [i, f32p, f32p, f64p, d]),
            "orc_conv3d_brute_f64": (i, [i, f32p, f32p, f64p, d]),
            "orc_conv3d_f32": (i, [i, i, f32p, f32p, f32p, i]),
            "orc_fft3d_f32": (i, [i, f32p, i, i]),
            "orc_conv3d_f32_spec": (i, [i, f32p, f32p, f32p, i]),
            "orc_snr_db": (d, [f32p, f32p, sz]),
            "orc_rel_l2_f32_f64": (d, [f32p, f64p, sz]),
            "orc_rel_l2_f32_f32": (d, [f32p, f32p, sz]),
            "orc_fftw_available": (i, []),
            "orc_max_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f32(a):
    assert a.dtype == np.complex64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f64(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def uniform(npts, seed):
    """npts complex64 values, re and im i.i.d. U[-1,1) (portable splitmix64)."""
    out = np.empty(npts, np.complex64)
    lib().orc_fill_uniform(_f32(out), 2 * npts, seed)
    return out


def loguniform(npts, seed):
    """npts complex64 values, |v| = 10^U(-3,3), arg U[0,2pi)."""
    out = np.empty(npts, np.complex64)
    lib().orc_fill_loguniform(_f32(out), npts, seed)
    return out


def ref01(npts, batch=1):
    """The reference's create_data (helper.cpp:148-172)."""
    out = np.empty(npts * batch, np.complex64)
    lib().orc_fill_ref01(_f32(out), npts, batch)
    return out


def fft3d_f64(x, sign=-1):
    n = round(len(x) ** (1 / 3))
    a = np.ascontiguousarray(x, np.complex128).copy()
    assert lib().orc_fft3d_f64(n, _f64(a), sign) == 0
    return a


def conv3d_f64(n, sig, filt, scale=1.0):
    out = np.empty(n ** 3, np.complex128)
    assert lib().orc_conv3d_f64(n, _f32(sig), _f32(filt), _f64(out), scale) == 0
    return out


def conv3d_brute_f64(n, sig, filt, scale=1.0):
    out = np.empty(n ** 3, np.complex128)
    assert lib().orc_conv3d_brute_f64(n, _f32(sig), _f32(filt), _f64(out), scale) == 0
    return out


def conv3d_f32(n, sig, filt, batch=1, nthreads=0):
    out = np.empty(n ** 3 * batch, np.complex64)
    assert lib().orc_conv3d_f32(n, batch, _f32(sig), _f32(filt), _f32(out), nthreads) == 0
    return out


def fft3d_f32(x, sign=-1, nthreads=0):
    n = round(len(x) ** (1 / 3))
    a = np.ascontiguousarray(x, np.complex64).copy()
    assert lib().orc_fft3d_f32(n, _f32(a), sign, nthreads) == 0
    return a


def conv3d_f32_spec(n, sig, H, out, nthreads=0):
    """CPU baseline step: out = IFFT(FFT(sig) * H) in fp32 with H precomputed."""
    assert lib().orc_conv3d_f32_spec(n, _f32(sig), _f32(H), _f32(out), nthreads) == 0
    return out


def snr_db(ref, test):
    return lib().orc_snr_db(_f32(ref), _f32(test), len(ref))


def rel_l2(test, ref):
    if ref.dtype == np.complex128:
        return lib().orc_rel_l2_f32_f64(_f32(test), _f64(ref), len(ref))
    return lib().orc_rel_l2_f32_f32(_f32(test), _f32(ref), len(ref))


def fftw_available():
    return bool(lib().orc_fftw_available())


def max_threads():
    return lib().orc_max_threads()
