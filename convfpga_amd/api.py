"""Python mirror of the convfpga C API (ctypes over lib/libconvfpga.so).

The drop-in functions keep the reference's names, arguments and error
behaviour (host/include/convfpga/convfpga.h:36-46): fpga_initialize returns
0 / -1..-5, the conv calls return an fpga_t whose .valid is False on bad
arguments.  Host buffers are numpy complex64 arrays in natural order
(index z*N*N + y*N + x).  The extension calls (include/convfpga/convfpga_ext.h)
take device pointers: anything with a data_ptr() method (a torch tensor on the
GPU) or a plain integer address.

This module never falls back to a CPU path: if the native library or the code
object is missing it raises.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libconvfpga.so")
CODE_OBJECT = os.path.join(_PKG, "lib", "convfpga_gfx950.hsaco")

CF_OK = 0
CF_ERR_ARG, CF_ERR_SIZE, CF_ERR_STATE, CF_ERR_RUNTIME, CF_ERR_NOMEM = -1, -2, -3, -4, -5
CF_SCALE_NONE, CF_SCALE_INV_N3 = 0, 1
CF_PREC_FP32, CF_PREC_SPLIT_FP16 = 0, 1
CF_MAX_PASSES = 8
CF_NOALLOC_SCRATCH_BYTES = 128 << 20
CF_PLAN_AUTO, CF_PLAN_AXIS5, CF_PLAN_PLANE3 = 0, 1, 2
CF_PLAN_SPLIT3_G22, CF_PLAN_SPLIT3_G24, CF_PLAN_SPLIT3_G42, CF_PLAN_SPLIT3_G44 = 3, 4, 5, 6
CF_PLAN_SPLIT3_G24M2, CF_PLAN_SPLIT3_G24M1, CF_PLAN_SPLIT3_G28M4, CF_PLAN_SPLIT3_G28M2, CF_PLAN_SPLIT3_G28M1 = 7, 8, 9, 10, 11
CF_PLAN_AXIS5Z = 12
CF_PLAN_SPLIT3_G24M2O, CF_PLAN_SPLIT3_G28M2O, CF_PLAN_SPLIT3_G24M2H, CF_PLAN_SPLIT3_G28M2H = 13, 14, 15, 16
CF_PLAN_SPLIT3_G28M2H4 = 17
CF_PLAN_SPLIT3_G28M2OC4 = 18
CF_PLAN_SPLIT3_G18M2OC4 = 19
CF_NUM_PLANS = 20
CF_BATCH_PERSISTENT, CF_BATCH_GRAPH = 0, 1
PLAN_NAMES = {1: "axis5", 2: "plane3", 3: "split3_g22", 4: "split3_g24", 5: "split3_g42", 6: "split3_g44",
              7: "split3_g24m2", 8: "split3_g24m1", 9: "split3_g28m4", 10: "split3_g28m2", 11: "split3_g28m1",
              12: "axis5z", 13: "split3_g24m2o", 14: "split3_g28m2o",
              15: "split3_g24m2h", 16: "split3_g28m2h", 17: "split3_g28m2h4", 18: "split3_g28m2oc4",
              19: "split3_g18m2oc4"}
SUPPORTED_SIZES = (8, 16, 32, 64, 128, 256, 512)

# names of every symbol include/convfpga/*.h declares
DROPIN_SYMBOLS = ("fpga_initialize", "fpga_final", "fpgaf_complex_malloc", "fpgaf_conv3D",
                  "fpgaf_conv3D_svm", "fpgaf_conv3D_svm_batch")
EXT_SYMBOLS = ("cf_last_error", "cf_supported_size", "cf_spectrum_create", "cf_spectrum_destroy",
               "cf_spectrum_data", "cf_spectrum_edge", "cf_spectrum_count", "cf_spectrum_view", "cf_conv3d_device", "cf_conv3d_device_noalloc", "cf_conv3d_host",
               "cf_conv3d_device_timed", "cf_conv3d_share_timed", "cf_fft3d_device", "cf_rel_l2_device", "cf_profile_begin",
               "cf_profile_end", "cf_fill_uniform", "cf_fill_loguniform", "cf_plan_supported", "cf_set_plan", "cf_get_plan",
               "cf_autotune", "cf_spectrum_plan", "cf_batch_create", "cf_batch_launch", "cf_batch_check",
               "cf_batch_destroy", "cf_batch_create_stall_test")


class fpga_t(ctypes.Structure):
    """Layout of the reference's fpga_t (convfpga.h:24-34)."""
    _fields_ = [("filter_pcie_wr_t", ctypes.c_double), ("filter_exec_t", ctypes.c_double),
                ("sig_pcie_wr_t", ctypes.c_double), ("sig_pcie_rd_t", ctypes.c_double),
                ("sig_exec_t", ctypes.c_double), ("siginv_exec_t", ctypes.c_double),
                ("valid", ctypes.c_bool)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class cf_pass_times_t(ctypes.Structure):
    _fields_ = [("npass", ctypes.c_uint), ("ms", ctypes.c_float * CF_MAX_PASSES),
                ("name", (ctypes.c_char * 24) * CF_MAX_PASSES),
                ("bytes", ctypes.c_double * CF_MAX_PASSES)]

    def as_list(self):
        return [{"name": self.name[i].value.decode(), "ms": float(self.ms[i]), "bytes": float(self.bytes[i])}
                for i in range(self.npass)]


class cf_host_times_t(ctypes.Structure):
    _fields_ = [("pass_ms", ctypes.c_double), ("wall_ms", ctypes.c_double)]


class cf_instance_t(ctypes.Structure):
    _fields_ = [("d_sig", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("h", ctypes.c_void_p)]


class ConvError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libconvfpga.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ConvError("libconvfpga.so not built: run `python -m convfpga_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, u, i, sz, d = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_size_t, ctypes.c_double
        sig = {
            "fpga_initialize": (i, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_bool]),
            "fpga_final": (None, []),
            "fpgaf_complex_malloc": (vp, [sz]),
            "fpgaf_conv3D": (fpga_t, [u, vp, vp, vp]),
            "fpgaf_conv3D_svm": (fpga_t, [u, vp, vp, vp]),
            "fpgaf_conv3D_svm_batch": (fpga_t, [u, vp, vp, vp, u]),
            "cf_last_error": (ctypes.c_char_p, []),
            "cf_supported_size": (i, [u]),
            "cf_spectrum_create": (i, [u, u, vp, i, vp, ctypes.POINTER(vp)]),
            "cf_spectrum_destroy": (None, [vp]),
            "cf_spectrum_data": (vp, [vp]),
            "cf_spectrum_edge": (u, [vp]),
            "cf_spectrum_count": (u, [vp]),
            "cf_spectrum_view": (i, [vp, u, ctypes.POINTER(vp)]),
            "cf_conv3d_device": (i, [vp, vp, vp, u, i, i, vp]),
            "cf_conv3d_device_noalloc": (i, [vp, vp, vp, u, i, i, vp]),
            "cf_conv3d_host": (i, [vp, vp, vp, u, i, i, ctypes.POINTER(cf_host_times_t)]),
            "cf_conv3d_device_timed": (i, [vp, vp, vp, u, i, i, vp, ctypes.POINTER(cf_pass_times_t)]),
            "cf_conv3d_share_timed": (i, [vp, vp, vp, u, i, vp, ctypes.POINTER(cf_pass_times_t)]),
            "cf_fft3d_device": (i, [u, vp, vp, u, i, vp]),
            "cf_rel_l2_device": (i, [vp, vp, i, sz, vp, ctypes.POINTER(d)]),
            "cf_profile_begin": (i, [u, u]),
            "cf_fill_uniform": (i, [vp, sz, ctypes.c_ulonglong, vp]),
            "cf_fill_loguniform": (i, [vp, sz, ctypes.c_ulonglong, vp]),
            "cf_plan_supported": (i, [u, i]),
            "cf_set_plan": (i, [u, i]),
            "cf_get_plan": (i, [u]),
            "cf_autotune": (i, [u, ctypes.POINTER(ctypes.c_float)]),
            "cf_spectrum_plan": (i, [vp]),
            "cf_batch_create": (i, [ctypes.POINTER(cf_instance_t), u, i, i, ctypes.POINTER(vp)]),
            "cf_batch_launch": (i, [vp, vp]),
            "cf_batch_check": (i, [vp]),
            "cf_batch_destroy": (None, [vp]),
            "cf_batch_create_stall_test": (i, [ctypes.POINTER(cf_instance_t), u, i, ctypes.c_ulonglong,
                                               ctypes.POINTER(vp)]),
            "cf_profile_end": (i, [ctypes.POINTER(cf_pass_times_t), ctypes.POINTER(u)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().cf_last_error().decode()


def _check(rc, what):
    if rc != CF_OK:
        raise ConvError("%s failed (%d): %s" % (what, rc, last_error()))


def _host(a, n_elems=None):
    if not isinstance(a, np.ndarray) or a.dtype != np.complex64 or not a.flags.c_contiguous:
        raise TypeError("host buffers must be C-contiguous numpy complex64 arrays")
    if n_elems is not None and a.size < n_elems:
        raise ValueError("buffer holds %d values, need %d" % (a.size, n_elems))
    return a.ctypes.data


_DEV_DTYPES = {"c64": ("complex64",), "c128": ("complex128",), "f32": ("float32", "complex64")}


def _dev(x, nbytes, kind="c64"):
    """Device address of x for a call that reads or writes `nbytes` bytes of
    it.  A torch tensor must live on the GPU, be contiguous, hold elements of
    the expected type and at least `nbytes` bytes; anything else raises before
    the library is called (a kernel handed a host pointer or a short buffer
    would fault the GPU).  A plain integer is taken as a raw device address
    whose extent the caller vouches for."""
    if x is None:
        return None
    if isinstance(x, bool):
        raise TypeError("a device buffer is a torch tensor on the GPU or an integer address, not bool")
    if isinstance(x, int):
        return x
    for attr in ("is_cuda", "is_contiguous", "element_size", "numel", "dtype", "data_ptr"):
        if not hasattr(x, attr):
            raise TypeError("a device buffer is a torch tensor on the GPU or an integer address, not %s"
                            % type(x).__name__)
    if not x.is_cuda:
        raise TypeError("device buffer is not on the GPU (device %s)" % (x.device,))
    if not x.is_contiguous():
        raise ValueError("device buffer must be contiguous")
    dt = str(x.dtype).replace("torch.", "")
    if dt not in _DEV_DTYPES[kind]:
        raise TypeError("device buffer of dtype %s where %s is expected" % (dt, " or ".join(_DEV_DTYPES[kind])))
    have = x.numel() * x.element_size()
    if have < nbytes:
        raise ValueError("device buffer holds %d bytes, the call needs %d" % (have, nbytes))
    return x.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream or None
    return stream.cuda_stream or None


# ---------------------------------------------------------------- drop-in API
def fpga_initialize(platform_name="gfx950", path=CODE_OBJECT, use_svm=False):
    pn = None if platform_name is None else platform_name.encode()
    pa = None if path is None else path.encode()
    return lib().fpga_initialize(pn, pa, bool(use_svm))


def fpga_final():
    lib().fpga_final()


def fpgaf_complex_malloc(sz):
    return lib().fpgaf_complex_malloc(sz)


def _conv_args(N, sig, filt, out, how_many=1):
    # sizes are only checked for a power-of-two N; any other N is passed
    # through so that the library reports it as the reference does (valid=0)
    n3 = N ** 3 if N > 0 and (N & (N - 1)) == 0 else 0
    p = []
    for a, need in ((sig, n3 * how_many), (filt, n3), (out, n3 * how_many)):
        p.append(None if a is None else _host(a, need))
    return p


def fpgaf_conv3D(N, sig, filt, out):
    s, f, o = _conv_args(N, sig, filt, out)
    return lib().fpgaf_conv3D(N, s, f, o)


def fpgaf_conv3D_svm(N, sig, filt, out):
    s, f, o = _conv_args(N, sig, filt, out)
    return lib().fpgaf_conv3D_svm(N, s, f, o)


def fpgaf_conv3D_svm_batch(N, sig, filt, out, how_many):
    s, f, o = _conv_args(N, sig, filt, out, max(how_many, 1))
    return lib().fpgaf_conv3D_svm_batch(N, s, f, o, how_many)


# -------------------------------------------------------------- extensions
def supported_size(N):
    return bool(lib().cf_supported_size(N))


class Spectrum:
    """Precomputed filter spectra (cf_spectrum_t): count filters of edge N."""

    def __init__(self, N, filters, count=1, on_device=True, stream=None):
        self._parent = None
        self._users, self._close_pending = 0, False  # open views and batches using it
        h = ctypes.c_void_p()
        if on_device:
            ptr = _dev(filters, 8 * N ** 3 * count)
        else:
            ptr = _host(filters, N ** 3 * count)
        _check(lib().cf_spectrum_create(N, count, ptr, 1 if on_device else 0, _stream(stream), ctypes.byref(h)),
               "cf_spectrum_create")
        self._h = h
        self.N, self.count = N, count

    @classmethod
    def _wrap(cls, h, N, parent):
        obj = cls.__new__(cls)
        obj._h, obj.N, obj.count, obj._parent = h, N, 1, parent
        obj._users, obj._close_pending = 0, False
        return obj

    def _acquire(self):
        if not self._h or self._close_pending:
            raise ConvError("the spectrum is closed")
        self._users += 1

    def _release(self):
        self._users -= 1
        if self._close_pending and self._users == 0:
            self.close()

    def view(self, index):
        """Spectrum handle for filter `index` alone, sharing this one's memory
        (cf_spectrum_view); the C handle of this one is kept until its last
        view is closed."""
        self._acquire()
        h = ctypes.c_void_p()
        rc = lib().cf_spectrum_view(self._h, index, ctypes.byref(h))
        if rc < 0:
            self._release()
            _check(rc, "cf_spectrum_view")
        return Spectrum._wrap(h, self.N, self)

    @property
    def handle(self):
        return self._h

    def data_ptr(self):
        return lib().cf_spectrum_data(self._h)

    @property
    def plan(self):
        return lib().cf_spectrum_plan(self._h)

    def close(self):
        """Destroys the handle.  A view is valid only while its parent lives
        (cf_spectrum_view) and a batch reads its spectra at every launch, so a
        spectrum with open views or batches is destroyed when the last of them
        is closed."""
        if not self._h:
            return
        if self._users:
            self._close_pending = True
            return
        lib().cf_spectrum_destroy(self._h)
        self._h = ctypes.c_void_p()
        parent, self._parent = self._parent, None
        if parent is not None:
            parent._release()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def conv3d_device(spec, sig, out, how_many=1, scale=CF_SCALE_INV_N3, precision=CF_PREC_FP32, stream=None):
    nb = 8 * spec.N ** 3 * how_many
    _check(lib().cf_conv3d_device(spec.handle, _dev(sig, nb), _dev(out, nb), how_many, scale, precision,
                                  _stream(stream)),
           "cf_conv3d_device")


def conv3d_device_noalloc(spec, sig, out, how_many=1, scale=CF_SCALE_INV_N3, precision=CF_PREC_FP32, stream=None):
    """cf_conv3d_device without any allocation: the intermediate of the
    four-step and z-blocked plans lives in the code object's fixed scratch
    (CF_NOALLOC_SCRATCH_BYTES); calls overlapping on the GPU race."""
    nb = 8 * spec.N ** 3 * how_many
    _check(lib().cf_conv3d_device_noalloc(spec.handle, _dev(sig, nb), _dev(out, nb), how_many, scale, precision,
                                          _stream(stream)), "cf_conv3d_device_noalloc")


def conv3d_host(spec, sig, out, how_many=1, scale=CF_SCALE_INV_N3, precision=CF_PREC_FP32):
    """Host-resident batch through the overlapped upload / passes / download
    pipeline (cf_conv3d_host); sig and out are numpy complex64 arrays of
    how_many cubes (may be the same array).  Returns (pass_ms, wall_ms)."""
    P = spec.N ** 3 * how_many
    t = cf_host_times_t()
    _check(lib().cf_conv3d_host(spec.handle, _host(sig, P), _host(out, P), how_many, scale, precision,
                                ctypes.byref(t)), "cf_conv3d_host")
    return t.pass_ms, t.wall_ms


def conv3d_device_timed(spec, sig, out, how_many=1, scale=CF_SCALE_INV_N3, precision=CF_PREC_FP32, stream=None):
    t = cf_pass_times_t()
    nb = 8 * spec.N ** 3 * how_many
    _check(lib().cf_conv3d_device_timed(spec.handle, _dev(sig, nb), _dev(out, nb), how_many, scale, precision,
                                        _stream(stream), ctypes.byref(t)), "cf_conv3d_device_timed")
    return t.as_list()


def conv3d_share_timed(spec, sig, out, shares, precision=CF_PREC_FP32, stream=None):
    """Per-pass times of one GPU's share (1/shares of every pass's tiles) of a
    slab-decomposed convolution of one cube (cf_conv3d_share_timed)."""
    t = cf_pass_times_t()
    nb = 8 * spec.N ** 3
    _check(lib().cf_conv3d_share_timed(spec.handle, _dev(sig, nb), _dev(out, nb), shares, precision, _stream(stream),
                                       ctypes.byref(t)), "cf_conv3d_share_timed")
    return t.as_list()


def fft3d_device(N, x, out, how_many=1, inverse=False, stream=None):
    nb = 8 * N ** 3 * how_many
    _check(lib().cf_fft3d_device(N, _dev(x, nb), _dev(out, nb), how_many, 1 if inverse else 0, _stream(stream)),
           "cf_fft3d_device")


def rel_l2_device(test, ref, npts, ref_is_f64, stream=None):
    r = ctypes.c_double()
    refp = _dev(ref, 16 * npts, "c128") if ref_is_f64 else _dev(ref, 8 * npts)
    _check(lib().cf_rel_l2_device(_dev(test, 8 * npts), refp, 1 if ref_is_f64 else 0, npts, _stream(stream),
                                  ctypes.byref(r)), "cf_rel_l2_device")
    return r.value


def profile_begin(max_calls, every=1):
    """Per-pass events on the last of every group of `every` of the next
    cf_conv3d_device calls, up to max_calls recorded calls."""
    _check(lib().cf_profile_begin(max_calls, every), "cf_profile_begin")


def profile_end():
    """-> (list of {name, ms (mean per call), bytes}, number of calls recorded)"""
    t = cf_pass_times_t()
    n = ctypes.c_uint()
    _check(lib().cf_profile_end(ctypes.byref(t), ctypes.byref(n)), "cf_profile_end")
    return t.as_list(), n.value


def fill_uniform(x, nfloats, seed, stream=None):
    """Device-side U[-1,1) fill, bit-identical to the host splitmix64 generator."""
    _check(lib().cf_fill_uniform(_dev(x, 4 * nfloats, "f32"), nfloats, seed, _stream(stream)), "cf_fill_uniform")


def fill_loguniform(x, npts, seed, stream=None):
    _check(lib().cf_fill_loguniform(_dev(x, 8 * npts), npts, seed, _stream(stream)), "cf_fill_loguniform")


def plan_supported(N, plan):
    return bool(lib().cf_plan_supported(N, plan))


def set_plan(N, plan):
    _check(lib().cf_set_plan(N, plan), "cf_set_plan")


def get_plan(N):
    rc = lib().cf_get_plan(N)
    if rc < 0:
        _check(rc, "cf_get_plan")
    return rc


def autotune(N):
    """-> (chosen plan, {plan: mean ms per convolution})"""
    ms = (ctypes.c_float * CF_NUM_PLANS)()
    rc = lib().cf_autotune(N, ms)
    if rc < 0:
        _check(rc, "cf_autotune")
    return rc, {p: ms[p] for p in range(1, CF_NUM_PLANS) if ms[p] > 0}


class Batch:
    """A fixed list of (signal, output, Spectrum) convolutions replayed with one
    graph launch (cf_batch_t).  Keeps its signals, outputs and spectra alive
    until it is closed (a spectrum closed meanwhile is destroyed then)."""

    def __init__(self, instances, scale=CF_SCALE_INV_N3, mode=CF_BATCH_PERSISTENT, stall_test_spin_limit=None):
        """stall_test_spin_limit: build the test-only batch of
        cf_batch_create_stall_test instead (persistent mode; its last stage
        waits for a tile that never comes, at most that many polls)."""
        self._h, self._keep, self._specs, self.mode = ctypes.c_void_p(), list(instances), [], mode
        arr = (cf_instance_t * len(instances))()
        for k, (sig, out, spec) in enumerate(instances):
            nb = 8 * spec.N ** 3
            arr[k] = cf_instance_t(_dev(sig, nb), _dev(out, nb), spec.handle)
        h = ctypes.c_void_p()
        if stall_test_spin_limit is None:
            _check(lib().cf_batch_create(arr, len(instances), scale, mode, ctypes.byref(h)), "cf_batch_create")
        else:
            _check(lib().cf_batch_create_stall_test(arr, len(instances), scale, stall_test_spin_limit,
                                                    ctypes.byref(h)), "cf_batch_create_stall_test")
        self._h = h
        for spec in {id(i[2]): i[2] for i in instances}.values():
            spec._acquire()
            self._specs.append(spec)

    def launch(self, stream=None):
        _check(lib().cf_batch_launch(self._h, _stream(stream)), "cf_batch_launch")

    def check(self):
        _check(lib().cf_batch_check(self._h), "cf_batch_check")

    def close(self):
        if self._h:
            lib().cf_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        specs, self._specs = self._specs, []
        for spec in specs:
            spec._release()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
