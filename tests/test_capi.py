"""CPU-only checks of the C-ABI boundary: the library loads, exports exactly
what include/convfpga/*.h declares, keeps the reference's type layouts, and
validates arguments before touching the GPU (host/src/convfpga.c:33-37,
host/src/fpga_init.c:37-43, 57-122).  No kernel is launched here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import convfpga_amd as C
from convfpga_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", "convfpga", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b((?:fpga|fpgaf|cf)_[a-z0-9_A-Z]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    declared = _declared("convfpga.h") | _declared("convfpga_ext.h")
    out = subprocess.check_output(["nm", "-D", "--defined-only", api.LIB_PATH], text=True)
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared == exported, (declared ^ exported)
    assert set(api.DROPIN_SYMBOLS) | set(api.EXT_SYMBOLS) == declared
    for name in declared:
        assert hasattr(api.lib(), name)


def test_code_object_built_for_gfx950():
    assert os.path.exists(api.CODE_OBJECT)
    blob = open(api.CODE_OBJECT, "rb").read()
    assert b"gfx950" in blob
    for n in C.SUPPORTED_SIZES:
        for k in ("row_fwd", "row_inv", "coly_fwd", "coly_inv", "colz_fwd", "colz_inv", "colz_conv"):
            assert ("cf_%s_%d" % (k, n)).encode() in blob


def test_fpga_t_layout():
    assert ctypes.sizeof(api.fpga_t) == 56
    assert api.fpga_t.valid.offset == 48


def test_initialize_error_codes_without_device():
    assert C.fpga_initialize("gfx950", None) == -1
    assert C.fpga_initialize("gfx950", "") == -1
    assert C.fpga_initialize(None, "x.hsaco") == -2
    assert C.fpga_initialize("", "x.hsaco") == -2
    assert C.fpga_initialize("Intel(R) FPGA SDK for OpenCL(TM)", "x.hsaco") == -2


def test_final_is_idempotent():
    C.fpga_final()
    C.fpga_final()


def test_complex_malloc():
    assert C.fpgaf_complex_malloc(0) is None
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [ctypes.c_void_p]
    for sz in (1, 8, 4096, 1 << 20):
        p = C.fpgaf_complex_malloc(sz)
        assert p is not None and p % 64 == 0
        libc.free(p)  # the caller frees with free() (src/main.cpp:90-92)


@pytest.mark.parametrize("fn", ["fpgaf_conv3D", "fpgaf_conv3D_svm"])
def test_conv_rejects_bad_arguments(fn):
    f = getattr(C, fn)
    a = np.zeros(8 ** 3, np.complex64)
    for args in ((8, None, a, a), (8, a, None, a), (8, a, a, None), (6, a, a, a), (0, a, a, a), (12, a, a, a)):
        r = f(*args)
        assert r.valid is False
        assert all(v == 0.0 for k, v in r.as_dict().items() if k != "valid")
    # power of two but not built (and library not initialised) -> invalid, no crash
    big = np.zeros(4 ** 3, np.complex64)
    assert f(4, big, big, big).valid is False
    assert f(8, a, a, a).valid is False  # not initialised


def test_batch_rejects_zero_count():
    a = np.zeros(8 ** 3, np.complex64)
    assert C.fpgaf_conv3D_svm_batch(8, a, a, a, 0).valid is False
    assert C.fpgaf_conv3D_svm_batch(8, a, a, a, 1).valid is False  # not initialised


def test_ext_api_reports_state_errors():
    L = api.lib()
    assert L.cf_supported_size(256) == 1 and L.cf_supported_size(1024) == 0 and L.cf_supported_size(4) == 0
    h = ctypes.c_void_p()
    buf = np.zeros(8 ** 3, np.complex64)
    rc = L.cf_spectrum_create(8, 1, buf.ctypes.data, 0, None, ctypes.byref(h))
    assert rc == api.CF_ERR_STATE and "fpga_initialize" in api.last_error()
    assert L.cf_spectrum_create(8, 0, buf.ctypes.data, 0, None, ctypes.byref(h)) == api.CF_ERR_ARG
    assert L.cf_conv3d_device(None, None, None, 1, 0, 0, None) == api.CF_ERR_ARG
    assert L.cf_conv3d_host(None, buf.ctypes.data, buf.ctypes.data, 1, 0, 0, None) == api.CF_ERR_ARG
    assert L.cf_spectrum_view(None, 0, ctypes.byref(h)) == api.CF_ERR_ARG
    r = ctypes.c_double()
    assert L.cf_rel_l2_device(None, None, 0, 10, None, ctypes.byref(r)) == api.CF_ERR_ARG
    assert L.cf_fft3d_device(8, buf.ctypes.data, buf.ctypes.data, 1, 0, None) == api.CF_ERR_STATE


def test_host_buffer_type_checks():
    with pytest.raises(TypeError):
        C.fpgaf_conv3D(8, np.zeros(512, np.complex128), np.zeros(512, np.complex64), np.zeros(512, np.complex64))
    with pytest.raises(ValueError):
        C.fpgaf_conv3D(8, np.zeros(100, np.complex64), np.zeros(512, np.complex64), np.zeros(512, np.complex64))


def test_plan_support_table():
    L = api.lib()
    for n in C.SUPPORTED_SIZES:
        assert L.cf_plan_supported(n, api.CF_PLAN_AXIS5) == 1
        assert L.cf_plan_supported(n, api.CF_PLAN_PLANE3) == (1 if n in (32, 64, 128) else 0)
        four_step = list(range(api.CF_PLAN_SPLIT3_G22, api.CF_PLAN_SPLIT3_G28M1 + 1)) + \
            list(range(api.CF_PLAN_SPLIT3_G24M2O, api.CF_PLAN_SPLIT3_G18M2OC4 + 1))
        for p in four_step:
            want = n in (128, 256)
            assert L.cf_plan_supported(n, p) == (1 if want else 0)
        assert L.cf_plan_supported(n, api.CF_PLAN_AXIS5Z) == (1 if n >= 64 else 0)
    assert api.CF_NUM_PLANS == api.CF_PLAN_SPLIT3_G18M2OC4 + 1 == len(api.PLAN_NAMES) + 1
    assert L.cf_plan_supported(1024, api.CF_PLAN_AXIS5) == 0
    assert L.cf_plan_supported(256, 99) == 0
    assert L.cf_set_plan(256, 99) == api.CF_ERR_ARG
    assert L.cf_set_plan(256, api.CF_PLAN_SPLIT3_G44) == api.CF_ERR_STATE  # not initialised


def test_batch_api_argument_checks():
    L = api.lib()
    h = ctypes.c_void_p()
    assert L.cf_batch_create(None, 1, 0, 0, ctypes.byref(h)) == api.CF_ERR_ARG
    inst = (api.cf_instance_t * 1)()
    assert L.cf_batch_create(inst, 0, 0, 0, ctypes.byref(h)) == api.CF_ERR_ARG
    assert L.cf_batch_create(inst, 1, 7, 0, ctypes.byref(h)) == api.CF_ERR_ARG
    assert L.cf_batch_create(inst, 1, 0, 9, ctypes.byref(h)) == api.CF_ERR_ARG
    assert L.cf_batch_create(inst, 1, 0, 0, ctypes.byref(h)) == api.CF_ERR_STATE  # not initialised
    assert L.cf_batch_launch(None, None) == api.CF_ERR_ARG
    assert L.cf_batch_check(None) == api.CF_ERR_ARG
    L.cf_batch_destroy(None)


def test_batch_descriptor_is_64_bytes():
    # CfBatchSeg (convfpga_amd/csrc/batch_desc.h) is shared by host and device code
    src = open(os.path.join(ROOT, "convfpga_amd", "csrc", "batch_desc.h")).read()
    assert "sizeof(CfBatchSeg) == 64" in src


def test_spectrum_close_waits_for_its_views(monkeypatch):
    """A view is valid only while its parent lives (cf_spectrum_view): the
    Python mirror destroys a parent closed with open views when the last view
    is closed, and refuses new views of it meanwhile."""
    destroyed = []

    class FakeLib:
        def cf_spectrum_view(self, parent, index, out):
            out._obj.value = 1000 + index
            return 0

        def cf_spectrum_destroy(self, h):
            destroyed.append(h.value)

    monkeypatch.setattr(api, "lib", lambda: FakeLib())
    parent = api.Spectrum._wrap(ctypes.c_void_p(7), 8, None)
    v0, v1 = parent.view(0), parent.view(1)
    parent.close()
    assert destroyed == [] and parent.handle
    with pytest.raises(api.ConvError):
        parent.view(0)
    v0.close()
    assert destroyed == [1000]
    v1.close()
    assert destroyed == [1000, 1001, 7] and not parent.handle
    parent.close()  # idempotent
    assert destroyed == [1000, 1001, 7]


def test_spectrum_close_waits_for_its_batches(monkeypatch):
    """A batch reads its spectra at every launch: a spectrum closed while a
    batch uses it is destroyed when the batch is closed."""
    calls = []

    class FakeLib:
        def cf_batch_create(self, arr, count, scale, mode, out):
            out._obj.value = 55
            return 0

        def cf_batch_destroy(self, h):
            calls.append(("batch", h.value))

        def cf_spectrum_destroy(self, h):
            calls.append(("spectrum", h.value))

    monkeypatch.setattr(api, "lib", lambda: FakeLib())
    monkeypatch.setattr(api, "_dev", lambda t, *a, **k: 0)
    spec = api.Spectrum._wrap(ctypes.c_void_p(9), 8, None)
    b = api.Batch([(None, None, spec), (None, None, spec)])
    spec.close()
    assert calls == [] and spec.handle
    b.close()
    assert calls == [("batch", 55), ("spectrum", 9)] and not spec.handle


class _FakeDeviceTensor:
    """Stands in for a GPU tensor in the argument checks (no GPU here)."""

    def __init__(self, numel, dtype="torch.complex64", esize=8, cuda=True, contiguous=True):
        self._n, self.dtype, self._e, self.is_cuda, self._c = numel, dtype, esize, cuda, contiguous
        self.device = "cuda:0" if cuda else "cpu"

    def is_contiguous(self):
        return self._c

    def numel(self):
        return self._n

    def element_size(self):
        return self._e

    def data_ptr(self):
        return 0x1000


def test_device_buffers_are_checked_before_any_launch():
    """The Python mirror refuses host tensors, non-contiguous or mistyped
    tensors and tensors too small for the call, before the library runs a
    kernel on them."""
    import torch
    from convfpga_amd import api
    n = 8
    spec = api.Spectrum._wrap(ctypes.c_void_p(0), n, None)  # never reaches the library
    good = _FakeDeviceTensor(n ** 3)
    cases = [
        (torch.zeros(n ** 3, dtype=torch.complex64), TypeError),         # host memory
        (_FakeDeviceTensor(n ** 3 - 1), ValueError),                     # one point short
        (_FakeDeviceTensor(2 * n ** 3, "torch.float32", 4), TypeError),  # float32 cube
        (_FakeDeviceTensor(n ** 3, contiguous=False), ValueError),
        ("0x1000", TypeError),
    ]
    for bad, exc in cases:
        with pytest.raises(exc):
            api.conv3d_device(spec, bad, good)
        with pytest.raises(exc):
            api.conv3d_device(spec, good, bad)
    with pytest.raises(ValueError):  # two cubes need twice the bytes
        api.conv3d_device(spec, good, good, how_many=2)
    with pytest.raises(ValueError):
        api.fft3d_device(n, good, _FakeDeviceTensor(n ** 3 // 2))
    with pytest.raises(ValueError):
        api.Spectrum(n, _FakeDeviceTensor(n ** 3), count=2)
    with pytest.raises(TypeError):  # an FP64 reference must be complex128
        api.rel_l2_device(good, good, n ** 3, ref_is_f64=True)
    with pytest.raises(ValueError):
        api.fill_uniform(_FakeDeviceTensor(10), 21, 1)
    with pytest.raises(ValueError):
        api.fill_loguniform(_FakeDeviceTensor(10), 11, 1)
    with pytest.raises(ValueError):
        api.Batch([(good, _FakeDeviceTensor(3), spec)])
    # float32 storage is fine for the fills, which count floats
    assert api._dev(_FakeDeviceTensor(10, "torch.float32", 4), 40, "f32") == 0x1000
    assert api._dev(12345, 1 << 30) == 12345  # a raw address is the caller's responsibility
