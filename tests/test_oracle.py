"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle fails it:
numpy pocketfft golden vectors (committed fixtures and live), the O(P^2)
definition of circular convolution, closed forms (delta, shifted delta,
single tone, constant), Parseval, linearity, the round trip, and glibc's own
rand() for the reference's input generator.
"""
import ctypes
import hashlib
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "conv_golden.npz")


def _case(z, name):
    n, s1, s2, dist = z[name + "_meta"]
    return int(n), int(s1), int(s2), int(dist)


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["u8", "u16", "l16", "u32"])
def test_inputs_match_fixture(oracle, golden, name):
    n, s1, s2, dist = _case(golden, name)
    gen = oracle.uniform if dist == 0 else oracle.loguniform
    sig, filt = gen(n ** 3, s1), gen(n ** 3, s2)
    digest = hashlib.sha256(sig.tobytes() + filt.tobytes()).digest()
    assert digest == bytes(golden[name + "_sha256"])
    if name + "_sig" in golden:
        assert np.array_equal(sig.view(np.uint32), golden[name + "_sig"].view(np.uint32))


@pytest.mark.parametrize("name", ["u8", "u16", "l16", "u32"])
def test_conv_f64_matches_golden(oracle, golden, name):
    n, s1, s2, dist = _case(golden, name)
    gen = oracle.uniform if dist == 0 else oracle.loguniform
    sig, filt = gen(n ** 3, s1), gen(n ** 3, s2)
    out = oracle.conv3d_f64(n, sig, filt)
    ref = golden[name + "_conv"]
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("name", ["u8", "u16", "l16"])
def test_fft_f64_matches_golden(oracle, golden, name):
    sig = golden[name + "_sig"]
    out = oracle.fft3d_f64(sig, -1)
    ref = golden[name + "_fft"]
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-14


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_conv_f64_equals_brute_force(oracle, n):
    sig, filt = oracle.uniform(n ** 3, 100 + n), oracle.uniform(n ** 3, 200 + n)
    a = oracle.conv3d_f64(n, sig, filt)
    b = oracle.conv3d_brute_f64(n, sig, filt)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-13


def test_brute_force_is_circular_convolution(oracle):
    # independent restatement on a 4^3 case through numpy index arithmetic
    n = 4
    sig, filt = oracle.uniform(n ** 3, 5), oracle.uniform(n ** 3, 6)
    s = sig.astype(np.complex128).reshape(n, n, n)
    f = filt.astype(np.complex128).reshape(n, n, n)
    ref = np.zeros((n, n, n), np.complex128)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                ref += s[z, y, x] * np.roll(np.roll(np.roll(f, z, 0), y, 1), x, 2)
    out = oracle.conv3d_brute_f64(n, sig, filt).reshape(n, n, n) / n ** 3
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [8, 32, 64])
def test_fft_f64_vs_numpy_live(oracle, n):
    x = oracle.uniform(n ** 3, 7 + n)
    a = oracle.fft3d_f64(x, -1)
    b = np.fft.fftn(x.astype(np.complex128).reshape(n, n, n)).ravel()
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-14
    ai = oracle.fft3d_f64(x, +1)
    bi = np.fft.ifftn(x.astype(np.complex128).reshape(n, n, n)).ravel() * n ** 3
    assert np.linalg.norm(ai - bi) / np.linalg.norm(bi) < 1e-14


@pytest.mark.parametrize("n", [8, 32])
def test_kat_delta_and_shift(oracle, n):
    sig = oracle.uniform(n ** 3, 3)
    d = np.zeros(n ** 3, np.complex64)
    d[0] = 1
    out = oracle.conv3d_f64(n, sig, d)
    assert np.allclose(out, sig.astype(np.complex128) * n ** 3, rtol=0, atol=1e-9 * n ** 3)
    a, b, c = 1, 2, 3  # delta at (x=a, y=b, z=c) -> circular shift of the signal
    d[:] = 0
    d[(c * n + b) * n + a] = 1
    out = oracle.conv3d_f64(n, sig, d, scale=1.0 / n ** 3).reshape(n, n, n)
    ref = np.roll(sig.reshape(n, n, n), (c, b, a), axis=(0, 1, 2))
    assert np.allclose(out, ref, atol=1e-12)


def test_kat_single_tone_forward_sign(oracle):
    n = 16
    kx, ky, kz = 3, 5, 7
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    tone = np.exp(2j * np.pi * (kx * x + ky * y + kz * z) / n).ravel()
    spec = oracle.fft3d_f64(tone, -1)
    peak = (kz * n + ky) * n + kx
    assert abs(spec[peak] - n ** 3) < 1e-9
    spec[peak] = 0
    assert np.abs(spec).max() < 1e-9


def test_parseval_linearity_roundtrip(oracle):
    n = 16
    a, b = oracle.uniform(n ** 3, 1), oracle.uniform(n ** 3, 2)
    A, B = oracle.fft3d_f64(a, -1), oracle.fft3d_f64(b, -1)
    assert np.isclose(np.sum(np.abs(A) ** 2), n ** 3 * np.sum(np.abs(a.astype(np.complex128)) ** 2), rtol=1e-12)
    AB = oracle.fft3d_f64(a.astype(np.complex128) * 2 - 3j * b.astype(np.complex128), -1)
    assert np.allclose(AB, 2 * A - 3j * B, atol=1e-9)
    back = oracle.fft3d_f64(A, +1) / n ** 3
    assert np.allclose(back, a, atol=1e-12)


@pytest.mark.parametrize("n", [8, 16, 64])
def test_f32_restatement_close_to_f64(oracle, n):
    sig, filt = oracle.uniform(n ** 3, 9), oracle.uniform(n ** 3, 10)
    c32 = oracle.conv3d_f32(n, sig, filt)
    c64 = oracle.conv3d_f64(n, sig, filt)
    err = oracle.rel_l2(c32, c64)
    assert 1e-9 < err < 1e-6  # fp32 arithmetic, but no gross error
    assert oracle.snr_db(c32, c32) == float("inf")


def test_f32_batch_shares_filter(oracle):
    n = 8
    sig = oracle.uniform(3 * n ** 3, 4)
    filt = oracle.uniform(n ** 3, 5)
    out = oracle.conv3d_f32(n, sig, filt, batch=3)
    for b in range(3):
        one = oracle.conv3d_f32(n, np.ascontiguousarray(sig[b * n ** 3:(b + 1) * n ** 3]), filt)
        assert np.array_equal(out[b * n ** 3:(b + 1) * n ** 3], one)


def test_snr_matches_definition(oracle):
    ref = oracle.uniform(1000, 1)
    test = ref + np.complex64(1e-4) * oracle.uniform(1000, 2)
    r, t = ref.astype(np.complex128), test.astype(np.complex128)
    want = 10 * np.log10(np.sum(np.abs(r) ** 2) / np.sum(np.abs(r - t) ** 2))
    assert abs(oracle.snr_db(ref, test) - want) < 1e-3
    # 120 dB is rel-L2 1e-6
    assert abs(oracle.snr_db(ref, (ref * np.complex64(1 + 1e-6)).astype(np.complex64)) - 120) < 3


def test_ref01_is_glibc_rand(oracle):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    want = np.array([libc.rand() for _ in range(200)], np.float64)
    rand_max = 2147483647.0
    w = (want.astype(np.float32) / np.float32(rand_max)).astype(np.float32)
    got = oracle.ref01(100, batch=2)
    assert np.array_equal(got[:100].view(np.float32), w)
    assert np.array_equal(got[100:], got[:100])  # helper.cpp:166-171 copies instance 0


def test_uniform_range_and_determinism(oracle):
    a = oracle.uniform(1 << 16, 42)
    f = a.view(np.float32)
    assert f.min() >= -1.0 and f.max() < 1.0
    assert abs(f.mean()) < 0.01
    assert np.array_equal(a, oracle.uniform(1 << 16, 42))
    assert not np.array_equal(a, oracle.uniform(1 << 16, 43))
    m = np.abs(oracle.loguniform(1 << 14, 1))
    assert 9e-4 < m.min() and m.max() < 1.1e3


# ---- the FP32 restatement and the metrics, each pinned on its own ----------

@pytest.mark.parametrize("n", [8, 16, 32])
def test_fft_f32_close_to_f64_and_numpy(oracle, n):
    """orc_fft3d_f32 against the pinned FP64 transform and numpy, both
    directions: FP32 rounding only (a wrong sign, index or dropped stage is
    an O(1) error; a silent FP64 path would be below 1e-9)."""
    x = oracle.uniform(n ** 3, 300 + n)
    for sign in (-1, +1):
        a = oracle.fft3d_f32(x, sign)
        b = oracle.fft3d_f64(x, sign)
        err = oracle.rel_l2(a, b)
        assert 1e-9 < err < 1e-6, (sign, err)
    npref = np.fft.fftn(x.astype(np.complex128).reshape(n, n, n)).ravel()
    assert np.linalg.norm(oracle.fft3d_f32(x, -1) - npref) / np.linalg.norm(npref) < 1e-6


def test_fft_f32_single_tone_sign_and_amplitude(oracle):
    """A tone exp(+2 pi i k.r / N) lands in bin k with amplitude P = N^3 under
    the forward sign -; the backward transform (sign +) puts it in bin -k."""
    n = 16
    kx, ky, kz = 5, 2, 9
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    tone = np.exp(2j * np.pi * (kx * x + ky * y + kz * z) / n).ravel().astype(np.complex64)
    spec = oracle.fft3d_f32(tone, -1)
    peak = (kz * n + ky) * n + kx
    assert abs(spec[peak] - n ** 3) < 1e-2
    spec[peak] = 0
    assert np.abs(spec).max() < 1e-2
    back = oracle.fft3d_f32(tone, +1)
    mirror = (((n - kz) % n) * n + (n - ky) % n) * n + (n - kx) % n
    assert abs(back[mirror] - n ** 3) < 1e-2
    back[mirror] = 0
    assert np.abs(back).max() < 1e-2


def test_fft_f32_round_trip_and_thread_count(oracle):
    """IFFT(FFT(x)) = P x; every line is transformed by itself, so the result
    does not depend on the number of threads (bit-equal)."""
    n = 32
    x = oracle.uniform(n ** 3, 77)
    X1 = oracle.fft3d_f32(x, -1, nthreads=1)
    X4 = oracle.fft3d_f32(x, -1, nthreads=4)
    assert np.array_equal(X1.view(np.uint64), X4.view(np.uint64))
    back = oracle.fft3d_f32(X1, +1) / np.float32(n ** 3)
    assert oracle.rel_l2(back.astype(np.complex64), x) < 1e-6


@pytest.mark.parametrize("n", [8, 32])
def test_conv_f32_spec_equals_conv_f32(oracle, n):
    """The timed CPU baseline step (spectrum precomputed) performs exactly the
    operations of the FP32 restatement of the reference verifier: bit-equal
    output with H = fft3d_f32(filter), for one thread and for several."""
    sig, filt = oracle.uniform(n ** 3, 500 + n), oracle.uniform(n ** 3, 600 + n)
    H = oracle.fft3d_f32(filt, -1)
    want = oracle.conv3d_f32(n, sig, filt)
    for threads in (1, 3):
        out = np.empty_like(sig)
        oracle.conv3d_f32_spec(n, sig, H, out, nthreads=threads)
        assert np.array_equal(out.view(np.uint64), want.view(np.uint64))
    # and it is the convolution (pinned FP64 oracle)
    assert oracle.rel_l2(want, oracle.conv3d_f64(n, sig, filt)) < 1e-6


def test_rel_l2_matches_numpy(oracle):
    a = oracle.uniform(4096, 1)
    b = (a + np.complex64(3e-3) * oracle.uniform(4096, 2)).astype(np.complex64)
    a64, b64 = a.astype(np.complex128), b.astype(np.complex128)
    want = np.linalg.norm(b64 - a64) / np.linalg.norm(a64)
    assert abs(oracle.rel_l2(b, a) - want) <= 1e-12 * want  # f32 vs f32 reference
    ref64 = a64 + 1e-9j  # an FP64 reference that float32 cannot hold
    want64 = np.linalg.norm(b64 - ref64) / np.linalg.norm(ref64)
    assert abs(oracle.rel_l2(b, ref64) - want64) <= 1e-12 * want64
    # zero reference: the absolute norm; identical arrays: zero
    z = np.zeros(16, np.complex64)
    assert abs(oracle.rel_l2(b[:16].copy(), z) - np.linalg.norm(b64[:16])) < 1e-9
    assert oracle.rel_l2(a, a) == 0.0 and oracle.rel_l2(a, a64) == 0.0


def test_fftw_probe_and_threads(oracle):
    """The FFTW probe agrees with a direct dlopen of the library; the thread
    count is the OpenMP default (at least one)."""
    try:
        L = ctypes.CDLL("libfftw3f.so.3")
        have = all(hasattr(L, s) for s in ("fftwf_plan_many_dft", "fftwf_execute", "fftwf_destroy_plan"))
    except OSError:
        have = False
    assert oracle.fftw_available() == have
    assert oracle.max_threads() >= 1


def test_loguniform_matches_numpy_port(oracle):
    """orc_fill_loguniform against the numpy restatement of the generator
    (tests/golden/make_golden.py), value by value to one float ulp (libm's
    pow/cos/sin and numpy's may differ in the last bit of a double)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(os.path.dirname(__file__), "golden",
                                                                     "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    got = oracle.loguniform(1 << 12, 9)
    want = mg.loguniform(1 << 12, 9)
    assert np.allclose(got.view(np.float32), want.view(np.float32), rtol=2 ** -23, atol=0)
    assert np.array_equal(oracle.uniform(1 << 12, 9).view(np.uint32), mg.uniform(1 << 12, 9).view(np.uint32))
