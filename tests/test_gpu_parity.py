"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars (written here, per test): rel-L2 <= 1e-5 against the FP64 oracle for FP32
results; SNR > 120 dB against the FP32 restatement of the reference's
verifier (src/fft_conv3D.cpp:235-258) for the drop-in calls; closed-form KATs
at the full sizes (256^3, 512^3) where the oracle would be slow.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_L2_FP32 = 1e-5   # FP32 bar against the FP64 oracle
SNR_DB_MIN = 120.0   # the reference verifier's threshold


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128])
def test_dropin_conv_matches_oracle(gpu, oracle, n):
    print("case fpgaf_conv3D N=%d" % n, flush=True)
    sig, filt = oracle.uniform(n ** 3, 1000 + n), oracle.uniform(n ** 3, 2000 + n)
    out = np.zeros_like(sig)
    r = gpu.fpgaf_conv3D(n, sig, filt, out)
    assert r.valid
    assert r.sig_exec_t > 0 and r.siginv_exec_t > 0 and r.filter_exec_t > 0
    ref64 = oracle.conv3d_f64(n, sig, filt)
    assert oracle.rel_l2(out, ref64) <= REL_L2_FP32
    ref32 = oracle.conv3d_f32(n, sig, filt)
    assert oracle.snr_db(ref32, out) > SNR_DB_MIN


@pytest.mark.parametrize("n", [32, 64])
def test_dropin_svm_matches_oracle(gpu, oracle, n):
    print("case fpgaf_conv3D_svm N=%d" % n, flush=True)
    sig, filt = oracle.uniform(n ** 3, 3000 + n), oracle.uniform(n ** 3, 4000 + n)
    out = np.zeros_like(sig)
    r = gpu.fpgaf_conv3D_svm(n, sig, filt, out)
    assert r.valid and r.sig_pcie_wr_t == 0.0 and r.sig_pcie_rd_t == 0.0
    assert oracle.rel_l2(out, oracle.conv3d_f64(n, sig, filt)) <= REL_L2_FP32


@pytest.mark.parametrize("how_many", [1, 2, 3])
def test_dropin_svm_batch_distinct_instances(gpu, oracle, how_many):
    n = 32
    print("case fpgaf_conv3D_svm_batch N=%d how_many=%d" % (n, how_many), flush=True)
    sig = oracle.uniform(how_many * n ** 3, 5000 + how_many)   # every instance different
    filt = oracle.uniform(n ** 3, 6000)
    out = np.zeros_like(sig)
    r = gpu.fpgaf_conv3D_svm_batch(n, sig, filt, out, how_many)
    assert r.valid
    ref32 = oracle.conv3d_f32(n, sig, filt, batch=how_many)
    assert oracle.snr_db(ref32, out) > SNR_DB_MIN
    for b in range(how_many):
        s = slice(b * n ** 3, (b + 1) * n ** 3)
        ref = oracle.conv3d_f64(n, np.ascontiguousarray(sig[s]), filt)
        assert oracle.rel_l2(np.ascontiguousarray(out[s]), ref) <= REL_L2_FP32


def test_dropin_svm_batch_streams_groups(gpu, oracle):
    """fpgaf_conv3D_svm_batch at a size where every cube is its own group of
    the overlapped upload / passes / download pipeline and the device slots
    are reused (5 cubes, 3 slots); odd count."""
    n, how_many = 128, 5
    print("case fpgaf_conv3D_svm_batch N=%d how_many=%d (pipelined)" % (n, how_many), flush=True)
    sig = oracle.uniform(how_many * n ** 3, 5100)
    filt = oracle.uniform(n ** 3, 6100)
    out = np.zeros_like(sig)
    r = gpu.fpgaf_conv3D_svm_batch(n, sig, filt, out, how_many)
    assert r.valid and r.sig_exec_t > 0 and r.siginv_exec_t >= r.sig_exec_t
    assert r.filter_pcie_wr_t == 0 and r.sig_pcie_wr_t == 0 and r.sig_pcie_rd_t == 0
    for b in range(how_many):
        s = slice(b * n ** 3, (b + 1) * n ** 3)
        ref = oracle.conv3d_f64(n, np.ascontiguousarray(sig[s]), filt)
        assert oracle.rel_l2(np.ascontiguousarray(out[s]), ref) <= REL_L2_FP32, b


def test_dropin_reference_generator(gpu, oracle):
    """The reference's own inputs (rand()/RAND_MAX, helper.cpp:148-172) at its default N=64."""
    n = 64
    print("case ref01 N=64", flush=True)
    filt = oracle.ref01(n ** 3)
    sig = oracle.ref01(n ** 3)
    out = np.zeros_like(sig)
    assert gpu.fpgaf_conv3D(n, sig, filt, out).valid
    assert oracle.snr_db(oracle.conv3d_f32(n, sig, filt), out) > SNR_DB_MIN
    assert oracle.rel_l2(out, oracle.conv3d_f64(n, sig, filt)) <= REL_L2_FP32


def test_dropin_repeated_calls_and_sizes(gpu, oracle):
    """Workspace reuse across calls of different N (the reference rebuilds per call)."""
    for n in (64, 16, 128, 32):
        print("case repeat N=%d" % n, flush=True)
        sig, filt = oracle.uniform(n ** 3, 7000 + n), oracle.uniform(n ** 3, 8000 + n)
        out = np.zeros_like(sig)
        assert gpu.fpgaf_conv3D(n, sig, filt, out).valid
        assert oracle.rel_l2(out, oracle.conv3d_f64(n, sig, filt)) <= REL_L2_FP32


@pytest.mark.parametrize("n", [8, 16, 32, 64, 128, 256])
def test_fft3d_device_matches_numpy(gpu, oracle, n):
    print("case fft3d N=%d" % n, flush=True)
    x = oracle.uniform(n ** 3, 9000 + n)
    d = _t(x)
    y = d.clone()
    gpu.fft3d_device(n, d, y)
    ref = oracle.fft3d_f64(x, -1)
    got = _np(y)
    assert oracle.rel_l2(got, ref) <= REL_L2_FP32
    z = y.clone()
    gpu.fft3d_device(n, y, z, inverse=True)
    back = _np(z)
    assert oracle.rel_l2(back, x.astype(np.complex128) * n ** 3) <= REL_L2_FP32


@pytest.mark.parametrize("n", [32, 64, 128])
def test_device_conv_normalised(gpu, oracle, n):
    print("case device conv scale=1/N^3 N=%d" % n, flush=True)
    sig, filt = oracle.uniform(n ** 3, 11000 + n), oracle.uniform(n ** 3, 12000 + n)
    spec = gpu.Spectrum(n, _t(filt))
    ds = _t(sig)
    do = ds.clone().zero_()
    gpu.conv3d_device(spec, ds, do)
    ref = oracle.conv3d_f64(n, sig, filt, scale=1.0 / n ** 3)
    assert oracle.rel_l2(_np(do), ref) <= REL_L2_FP32
    # in place gives the same bits
    gpu.conv3d_device(spec, ds, ds)
    assert np.array_equal(_np(ds), _np(do))


def test_device_conv_loguniform(gpu, oracle):
    n = 64
    print("case loguniform N=64", flush=True)
    sig, filt = oracle.loguniform(n ** 3, 13001), oracle.loguniform(n ** 3, 13002)
    spec = gpu.Spectrum(n, _t(filt))
    ds = _t(sig)
    gpu.conv3d_device(spec, ds, ds)
    assert oracle.rel_l2(_np(ds), oracle.conv3d_f64(n, sig, filt, scale=1.0 / n ** 3)) <= REL_L2_FP32


def test_device_batch_shared_and_own_spectra(gpu, oracle):
    n, b = 64, 4
    print("case batch shared/own N=64 b=4", flush=True)
    sig = oracle.uniform(b * n ** 3, 14000)
    filts = oracle.uniform(b * n ** 3, 15000)
    ds = _t(sig)
    shared = gpu.Spectrum(n, _t(filts[:n ** 3]))
    own = gpu.Spectrum(n, _t(filts), count=b)
    o1, o2 = ds.clone(), ds.clone()
    gpu.conv3d_device(shared, ds, o1, how_many=b)
    gpu.conv3d_device(own, ds, o2, how_many=b)
    g1, g2 = _np(o1), _np(o2)
    for i in range(b):
        s = slice(i * n ** 3, (i + 1) * n ** 3)
        si = np.ascontiguousarray(sig[s])
        r1 = oracle.conv3d_f64(n, si, np.ascontiguousarray(filts[:n ** 3]), 1.0 / n ** 3)
        r2 = oracle.conv3d_f64(n, si, np.ascontiguousarray(filts[s]), 1.0 / n ** 3)
        assert oracle.rel_l2(np.ascontiguousarray(g1[s]), r1) <= REL_L2_FP32
        assert oracle.rel_l2(np.ascontiguousarray(g2[s]), r2) <= REL_L2_FP32


@pytest.mark.parametrize("plan", [1, 7])
def test_share_timed(gpu, oracle, plan):
    """cf_conv3d_share_timed: with one share it is the timed convolution of
    one cube (same output bits); with 8 shares every pass runs on 1/8 of its
    tiles and reports 1/8 of the bytes."""
    import torch
    api = gpu.api
    n = 256
    print("case share timed plan=%d" % plan, flush=True)
    P = n ** 3
    sig = torch.empty(P, dtype=torch.complex64, device="cuda")
    filt = torch.empty_like(sig)
    api.fill_uniform(sig, 2 * P, 3131)
    api.fill_uniform(filt, 2 * P, 3132)
    api.set_plan(n, plan)
    spec = gpu.Spectrum(n, filt)
    api.set_plan(n, api.CF_PLAN_AUTO)
    a, b = torch.empty_like(sig), torch.empty_like(sig)
    full = gpu.conv3d_device_timed(spec, sig, a)
    one = api.conv3d_share_timed(spec, sig, b, 1)
    assert torch.equal(a, b)
    eight = api.conv3d_share_timed(spec, sig, b, 8)
    assert [p["name"] for p in eight] == [p["name"] for p in full]
    for p8, p1 in zip(eight, one):
        assert p8["bytes"] * 8 == p1["bytes"] and 0 < p8["ms"] < p1["ms"]
    with pytest.raises(api.ConvError):
        api.conv3d_share_timed(spec, sig, b, 3)


def test_256_conv_matches_oracle(gpu, oracle):
    n = 256
    print("case 256^3 conv vs FP64 oracle", flush=True)
    sig, filt = oracle.uniform(n ** 3, 16001), oracle.uniform(n ** 3, 16002)
    spec = gpu.Spectrum(n, _t(filt))
    ds = _t(sig)
    do = ds.clone()
    gpu.conv3d_device(spec, ds, do)
    ref = oracle.conv3d_f64(n, sig, filt, scale=1.0 / n ** 3)
    got = _np(do)
    err = oracle.rel_l2(got, ref)
    assert err <= REL_L2_FP32
    # the FP64 device error norm agrees with the host one
    dref = _t(ref)
    dev_err = gpu.rel_l2_device(do, dref, n ** 3, ref_is_f64=True)
    assert abs(dev_err - err) <= 1e-3 * err


@pytest.mark.parametrize("n", [256, 512])
def test_full_size_delta_shift_and_roundtrip(gpu, oracle, n):
    import torch
    print("case KAT delta-shift / round trip N=%d" % n, flush=True)
    sig = oracle.uniform(n ** 3, 17000 + n)
    ds = _t(sig)
    d = torch.zeros_like(ds)
    a, b, c = 5, n - 3, 17
    d.view(torch.float32)[2 * ((c * n + b) * n + a)] = 1.0
    spec = gpu.Spectrum(n, d)
    out = torch.empty_like(ds)
    gpu.conv3d_device(spec, ds, out)
    want = torch.roll(ds.view(n, n, n), shifts=(c, b, a), dims=(0, 1, 2)).reshape(-1)
    err = gpu.rel_l2_device(out, want, n ** 3, ref_is_f64=False)
    assert err <= REL_L2_FP32
    y = torch.empty_like(ds)
    gpu.fft3d_device(n, ds, y)
    gpu.fft3d_device(n, y, y, inverse=True)
    y.mul_(1.0 / n ** 3)
    assert gpu.rel_l2_device(y, ds, n ** 3, ref_is_f64=False) <= REL_L2_FP32
    del ds, d, out, y, want
    torch.cuda.empty_cache()


def test_dropin_512_delta_shift(gpu, oracle):
    """The drop-in call at the largest edge (1 GiB cubes staged through the
    pinned chunk buffers, the autotuned 512^3 plan): a delta filter at (a, b, c)
    gives N^3 times the signal circularly shifted by (a, b, c)."""
    import torch
    n = 512
    print("case drop-in KAT delta-shift N=%d" % n, flush=True)
    sig = oracle.uniform(n ** 3, 17900)
    filt = np.zeros(n ** 3, np.complex64)
    a, b, c = 3, n - 7, 200
    filt[(c * n + b) * n + a] = 1.0
    out = np.empty_like(sig)
    r = gpu.fpgaf_conv3D(n, sig, filt, out)
    assert r.valid and r.sig_pcie_wr_t > 0 and r.sig_pcie_rd_t > 0
    del filt
    got = _t(out)
    want = torch.roll(_t(sig).view(n, n, n), shifts=(c, b, a), dims=(0, 1, 2)).reshape(-1).mul_(float(n ** 3))
    assert gpu.rel_l2_device(got, want, n ** 3, ref_is_f64=False) <= REL_L2_FP32
    del got, want, out, sig
    torch.cuda.empty_cache()


def test_rel_l2_device_matches_numpy(gpu, oracle):
    a = oracle.uniform(100003, 1)
    b = (a + np.complex64(1e-3) * oracle.uniform(100003, 2)).astype(np.complex64)
    want = np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b.astype(np.complex128))
    got = gpu.rel_l2_device(_t(a), _t(b), a.size, ref_is_f64=False)
    assert abs(got - want) <= 1e-12 * want
    got64 = gpu.rel_l2_device(_t(a), _t(b.astype(np.complex128)), a.size, ref_is_f64=True)
    assert abs(got64 - want) <= 1e-12 * want


def test_device_fill_matches_host_generator(gpu, oracle):
    import torch
    x = torch.empty(100001, dtype=torch.complex64, device="cuda")
    gpu.api.fill_uniform(x, 2 * x.numel(), 987654321)
    want = oracle.uniform(x.numel(), 987654321)
    assert np.array_equal(_np(x).view(np.uint32), want.view(np.uint32))
    # log-uniform: same variates; pow/sin/cos of the device and host math
    # libraries may differ in the last place of the double, so the floats
    # agree to a few ulp
    gpu.api.fill_loguniform(x, x.numel(), 424242)
    want = oracle.loguniform(x.numel(), 424242)
    got = _np(x)
    assert np.all(np.abs(got - want) <= 4e-7 * np.abs(want))
    m = np.abs(got)
    assert 0.9e-3 < m.min() and m.max() < 1.01e3


PLAN_CASES = [(32, 1), (32, 2), (64, 1), (64, 2), (128, 1), (128, 2), (256, 1)] + \
    [(256, p) for p in range(3, 12)] + [(512, 1)] + [(n, 12) for n in (64, 128, 256, 512)] + [(128, p) for p in range(3, 12)] + \
    [(n, p) for n in (128, 256) for p in (13, 14, 15, 16, 17, 18, 19)]


@pytest.mark.parametrize("n,plan", PLAN_CASES)
def test_every_plan_matches_oracle(gpu, oracle, n, plan):
    """Each schedule (5-pass axis, 3-pass plane, 3-pass four-step with its tile
    variants) on a batch of 2 with per-instance spectra and with a shared one."""
    import torch
    print("case plan=%d N=%d" % (plan, n), flush=True)
    api = gpu.api
    api.set_plan(n, plan)
    try:
        b = 2 if n < 512 else 1
        P = n ** 3
        sig = oracle.uniform(b * P, 20000 + n + plan)
        filt = oracle.uniform(b * P, 30000 + n + plan)
        ds, df = _t(sig), _t(filt)
        own = gpu.Spectrum(n, df, count=b)
        shared = gpu.Spectrum(n, df[:P])
        assert own.plan == plan and shared.plan == plan
        o1, o2 = torch.empty_like(ds), torch.empty_like(ds)
        gpu.conv3d_device(own, ds, o1, how_many=b)
        gpu.conv3d_device(shared, ds, o2, how_many=b)
        g1, g2 = _np(o1), _np(o2)
        for i in range(b):
            s = slice(i * P, (i + 1) * P)
            si = np.ascontiguousarray(sig[s])
            if n <= 256:
                r1 = oracle.conv3d_f64(n, si, np.ascontiguousarray(filt[s]), 1.0 / P)
                assert oracle.rel_l2(np.ascontiguousarray(g1[s]), r1) <= REL_L2_FP32
            if i == 0 or n <= 128:
                r2 = oracle.conv3d_f64(n, si, np.ascontiguousarray(filt[:P]), 1.0 / P)
                assert oracle.rel_l2(np.ascontiguousarray(g2[s]), r2) <= REL_L2_FP32
        # in-place call gives the same bits as out-of-place
        gpu.conv3d_device(shared, ds, ds, how_many=b)
        assert np.array_equal(_np(ds), g2)
        del own, shared, ds, df, o1, o2
        torch.cuda.empty_cache()
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)


@pytest.mark.parametrize("plan", [1, 3, 6, 12, 13, 16])
def test_dropin_256_on_each_plan(gpu, oracle, plan):
    n = 256
    print("case drop-in N=256 plan=%d" % plan, flush=True)
    gpu.api.set_plan(n, plan)
    try:
        sig, filt = oracle.uniform(n ** 3, 40001), oracle.uniform(n ** 3, 40002)
        ref32 = oracle.conv3d_f32(n, sig, filt)
        out = np.zeros_like(sig)
        r = gpu.fpgaf_conv3D(n, sig, filt, out)
        assert r.valid and min(r.filter_exec_t, r.sig_exec_t, r.siginv_exec_t, r.sig_pcie_wr_t, r.sig_pcie_rd_t) > 0
        assert oracle.snr_db(ref32, out) > SNR_DB_MIN
        out2 = np.zeros_like(sig)
        r = gpu.fpgaf_conv3D_svm(n, sig, filt, out2)
        assert r.valid and r.sig_pcie_wr_t == 0.0
        assert oracle.snr_db(ref32, out2) > SNR_DB_MIN
    finally:
        gpu.api.set_plan(n, gpu.api.CF_PLAN_AUTO)


def _spectrum_copy(spec, n, count):
    """Device copy of a spectrum handle's own buffer (its plan's layout)."""
    import ctypes
    import torch
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = torch.empty(count * n ** 3, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    assert hip.hipMemcpy(out.data_ptr(), spec.data_ptr(), out.numel() * 8, 3) == 0  # 3: device to device
    return _np(out)


def _spectrum_order(api, n, plan, spec64):
    """The FP64 spectrum (natural order, index (kz*N + ky)*N + kx) permuted into
    the layout a spectrum of `plan` keeps: natural for AXIS5 / PLANE3,
    ((kx/16)*N + ky)*16N + kz*16 + kx%16 for AXIS5Z, [k1][kx/GM][k2][kz][kx%GM]
    (ky = k1 + 16 k2) for the four-step plans."""
    name = api.PLAN_NAMES[plan]
    if name in ("axis5", "plane3"):
        return spec64
    if name == "axis5z":
        return np.ascontiguousarray(spec64.reshape(n, n, n // 16, 16).transpose(2, 1, 0, 3)).reshape(-1)
    tail = name.split("_g")[1].split("o")[0].split("h")[0]  # <GA><GB>[m<GM>][o|h|h4]
    gm = int(tail.split("m")[1]) if "m" in tail else int(tail[1])
    ref = spec64.reshape(n, n // 16, 16, n // gm, gm)  # (kz, k2, k1, kx/GM, kx%GM)
    return np.ascontiguousarray(ref.transpose(2, 3, 1, 0, 4)).reshape(-1)


@pytest.mark.parametrize("n,plan", [(64, 1), (64, 2), (64, 12), (128, 9), (256, 4), (256, 7), (256, 8), (256, 14)])
def test_spectrum_layout_of_each_plan(gpu, oracle, n, plan):
    """The filter spectrum a plan keeps (axis / plane passes: natural order;
    AXIS5Z: z-blocked; four-step: the convolution's first pass, then the
    forward half of its middle pass) equals the FP64 FFT3 of each filter
    permuted into that plan's layout, for both filters of a two-filter
    handle."""
    api = gpu.api
    print("case spectrum layout N=%d plan=%s" % (n, api.PLAN_NAMES[plan]), flush=True)
    P = n ** 3
    filts = [oracle.uniform(P, 83000 + 7 * n + k) for k in range(2)]
    api.set_plan(n, plan)
    try:
        spec = gpu.Spectrum(n, _t(np.concatenate(filts)), count=2)
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    assert spec.plan == plan
    got = _spectrum_copy(spec, n, 2)
    spec.close()
    for k, f in enumerate(filts):
        want = _spectrum_order(api, n, plan, oracle.fft3d_f64(f, -1))
        assert oracle.rel_l2(got[k * P:(k + 1) * P], want) <= REL_L2_FP32, k


@pytest.mark.parametrize("plan", [2, 5, 13])
def test_svm_batch_on_3pass_plans(gpu, oracle, plan):
    n = 64 if plan == 2 else 256
    hm = 3 if n == 64 else 2
    print("case svm batch plan=%d N=%d" % (plan, n), flush=True)
    gpu.api.set_plan(n, plan)
    try:
        sig = oracle.uniform(hm * n ** 3, 41000 + plan)
        filt = oracle.uniform(n ** 3, 42000 + plan)
        out = np.zeros_like(sig)
        assert gpu.fpgaf_conv3D_svm_batch(n, sig, filt, out, hm).valid
        assert oracle.snr_db(oracle.conv3d_f32(n, sig, filt, batch=hm), out) > SNR_DB_MIN
    finally:
        gpu.api.set_plan(n, gpu.api.CF_PLAN_AUTO)


def test_autotuner_picks_a_supported_plan(gpu):
    api = gpu.api
    for n in (64, 256):
        plan, ms = api.autotune(n)
        print("autotune N=%d -> %s %s" % (n, api.PLAN_NAMES[plan], ms), flush=True)
        assert api.plan_supported(n, plan)
        assert plan in ms and ms[plan] == min(ms.values())
        assert api.get_plan(n) == plan


def _mixed_instances(gpu, oracle, sizes, plans):
    """(sig, out, spectrum, host sig, host filt) per instance, each with its own
    filter; plans maps an edge to the plan its spectrum is made for (others:
    the autotuned plan)."""
    api = gpu.api
    inst = []
    for k, n in enumerate(sizes):
        if n in plans:
            api.set_plan(n, plans[n])
        s = oracle.uniform(n ** 3, 50000 + 10 * k)
        f = oracle.uniform(n ** 3, 60000 + 10 * k)
        spec = gpu.Spectrum(n, _t(f))
        if n in plans:
            api.set_plan(n, api.CF_PLAN_AUTO)
        ds = _t(s)
        inst.append((ds, ds.clone().zero_(), spec, s, f))
    return inst


def _check_instances(oracle, inst):
    for ds, do, spec, s, f in inst:
        n = spec.N
        got = _np(do)
        ref = oracle.conv3d_f64(n, s, f, 1.0 / n ** 3)
        assert oracle.rel_l2(got, ref) <= REL_L2_FP32, n


@pytest.mark.parametrize("mode,plan256", [(0, 1), (1, 4), (1, 12)])
def test_batch_mixed_sizes_replayed(gpu, oracle, mode, plan256):
    """Mixed edges with their own filters in one graph launch (persistent kernel
    or recorded per-instance passes); replayed with eager work in between, every
    replay checked.  The 256^3 instance runs the given plan."""
    import torch
    api = gpu.api
    sizes = [32, 64, 16, 128, 32, 256, 64, 32]
    print("case batch mode=%d plan256=%d sizes=%s" % (mode, plan256, sizes), flush=True)
    # the persistent kernel reads spectra in natural order (the autotuned plan
    # of an edge from 64 up may lay them out privately)
    plans = ({256: plan256, 128: api.CF_PLAN_AXIS5, 64: api.CF_PLAN_AXIS5} if mode == api.CF_BATCH_PERSISTENT
             else {256: plan256})
    inst = _mixed_instances(gpu, oracle, sizes, plans)
    b = gpu.api.Batch([(ds, do, spec) for ds, do, spec, _, _ in inst], mode=mode)
    for rep in range(3):
        for _, do, _, _, _ in inst:
            do.zero_()
        b.launch()
        b.check()
        _check_instances(oracle, inst)
        # eager work on the same device between replays
        x = torch.ones(1 << 20, device="cuda")
        assert float((x * 2).sum()) == 2.0 * (1 << 20)
    b.close()


@pytest.mark.parametrize("n,plan,count", [(64, 2, 3), (128, 7, 2), (128, 15, 2), (256, 7, 1), (256, 15, 1),
                                          (32, 2, 4)])
def test_persistent_batch_on_each_plan(gpu, oracle, n, plan, count):
    """The persistent kernel on each plan it runs (plane, four-step g24m2 and
    g24m2h): every instance against the FP64 oracle and bitwise against the
    same spectra in graph mode (the same pass code), over two replays."""
    import torch
    api = gpu.api
    P = n ** 3
    print("case persistent batch N=%d plan=%s count=%d" % (n, api.PLAN_NAMES[plan], count), flush=True)
    sig = [oracle.uniform(P, 85000 + 11 * n + k) for k in range(count)]
    flt = [oracle.uniform(P, 86000 + 11 * n + k) for k in range(count)]
    api.set_plan(n, plan)
    try:
        specs = [gpu.Spectrum(n, _t(f)) for f in flt]
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    ds = [_t(x) for x in sig]
    op = [torch.zeros_like(d) for d in ds]
    og = [torch.zeros_like(d) for d in ds]
    bp = api.Batch(list(zip(ds, op, specs)), mode=api.CF_BATCH_PERSISTENT)
    bg = api.Batch(list(zip(ds, og, specs)), mode=api.CF_BATCH_GRAPH)
    for rep in range(2):
        for o in op + og:
            o.zero_()
        bp.launch()
        bp.check()
        bg.launch()
        bg.check()
        for k in range(count):
            assert torch.equal(op[k].view(torch.int64), og[k].view(torch.int64)), (rep, k)
    for k in range(count):
        ref = oracle.conv3d_f64(n, sig[k], flt[k], 1.0 / P)
        assert oracle.rel_l2(_np(op[k]), ref) <= REL_L2_FP32, k
    bp.close()
    bg.close()


def test_persistent_batch_longer_than_one_listing(gpu, oracle):
    """A persistent batch with more than 512 work items per work-group (the
    kernel lists a work-group's items in LDS 512 positions at a time, so this
    takes a second listing): 16^3 convolutions of one spectrum on the axis
    plan, five items each, about 600 per work-group at two work-groups per CU;
    every output bitwise equal to graph mode, a few against the FP64 oracle."""
    import torch
    api = gpu.api
    n = 16
    P = n ** 3
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    count = min((600 * 2 * cus) // 5, 65536)
    print("case persistent batch of %d 16^3 instances" % count, flush=True)
    f = oracle.uniform(P, 87001)
    api.set_plan(n, api.CF_PLAN_AXIS5)
    try:
        spec = gpu.Spectrum(n, _t(f))
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    sig = torch.empty(count * P, dtype=torch.complex64, device="cuda")
    api.fill_uniform(sig, 2 * count * P, 87002)
    op = torch.zeros_like(sig)
    og = torch.zeros_like(sig)
    nb = 8 * P
    bp = api.Batch([(sig.data_ptr() + k * nb, op.data_ptr() + k * nb, spec) for k in range(count)],
                   mode=api.CF_BATCH_PERSISTENT)
    bg = api.Batch([(sig.data_ptr() + k * nb, og.data_ptr() + k * nb, spec) for k in range(count)],
                   mode=api.CF_BATCH_GRAPH)
    bp.launch()
    bp.check()
    bg.launch()
    bg.check()
    torch.cuda.synchronize()
    assert torch.equal(op.view(torch.int64), og.view(torch.int64))
    for k in (0, count // 2, count - 1):
        ref = oracle.conv3d_f64(n, _np(sig[k * P:(k + 1) * P]), f, 1.0 / P)
        assert oracle.rel_l2(_np(op[k * P:(k + 1) * P]), ref) <= REL_L2_FP32, k
    bp.close()
    bg.close()


def test_spectrum_views_outlive_a_closed_parent(gpu, oracle):
    """Closing a two-filter spectrum while views of it are open keeps its
    memory until the last view is closed: a view still convolves correctly."""
    n = 32
    P = n ** 3
    print("case spectrum views after parent close", flush=True)
    f = [oracle.uniform(P, 84000 + k) for k in range(2)]
    s = oracle.uniform(P, 84100)
    parent = gpu.Spectrum(n, _t(np.concatenate(f)), count=2)
    v0, v1 = parent.view(0), parent.view(1)
    parent.close()
    assert parent.handle
    d = _t(s)
    o = d.clone()
    gpu.conv3d_device(v1, d, o)
    assert oracle.rel_l2(_np(o), oracle.conv3d_f64(n, s, f[1], 1.0 / P)) <= REL_L2_FP32
    v0.close()
    v1.close()
    assert not parent.handle


@pytest.mark.parametrize("mode,plan128", [(0, 0), (1, 0), (1, 12)])
def test_batch_spectrum_views_grouped(gpu, oracle, mode, plan128):
    """Instances whose spectra are consecutive views of one multi-filter
    spectrum and whose signals / outputs are consecutive slices: the graph
    batch records each such run as one batched pass sequence (own filter per
    cube); the order interleaves edges, so runs break and resume."""
    import torch
    api = gpu.api
    counts = {32: 3, 64: 2, 128: 2}
    print("case batch spectrum views mode=%d plan128=%d counts=%s" % (mode, plan128, counts), flush=True)
    host, dev, specs = {}, {}, {}
    for n, c in counts.items():
        P = n ** 3
        s = [oracle.uniform(P, 81000 + 31 * n + k) for k in range(c)]
        f = [oracle.uniform(P, 82000 + 31 * n + k) for k in range(c)]
        host[n] = (s, f)
        dsig = _t(np.concatenate(s))
        dev[n] = (dsig, torch.zeros_like(dsig))
        if mode == api.CF_BATCH_PERSISTENT and n >= 64:  # the persistent kernel reads natural-order spectra
            api.set_plan(n, api.CF_PLAN_AXIS5)
        elif n == 128 and plan128:  # runs of a plan that goes through the workspace, batched
            api.set_plan(n, plan128)
        parent = gpu.Spectrum(n, _t(np.concatenate(f)), count=c)
        api.set_plan(n, api.CF_PLAN_AUTO)
        specs[n] = [parent.view(k) for k in range(c)]
        del parent  # the views keep it alive
    order = [(32, 0), (32, 1), (64, 0), (64, 1), (32, 2), (128, 0), (128, 1)]
    inst = []
    for n, k in order:
        P = n ** 3
        ds, do = dev[n]
        inst.append((ds[k * P:(k + 1) * P], do[k * P:(k + 1) * P], specs[n][k]))
    b = api.Batch(inst, mode=mode)
    for rep in range(2):
        for _, do in dev.values():
            do.zero_()
        b.launch()
        b.check()
        for n, c in counts.items():
            got = _np(dev[n][1]).reshape(c, -1)
            s, f = host[n]
            for k in range(c):
                ref = oracle.conv3d_f64(n, s[k], f[k], 1.0 / n ** 3)
                assert oracle.rel_l2(got[k], ref) <= REL_L2_FP32, (n, k, rep)
    b.close()
    # a view used directly equals its filter's own spectrum
    n = 32
    s, f = host[n]
    out = torch.zeros(n ** 3, dtype=torch.complex64, device="cuda")
    gpu.conv3d_device(specs[n][2], _t(s[2]), out)
    assert oracle.rel_l2(_np(out), oracle.conv3d_f64(n, s[2], f[2], 1.0 / n ** 3)) <= REL_L2_FP32
    with pytest.raises(api.ConvError):
        specs[n][0]._parent.view(counts[n])


@pytest.mark.parametrize("n,how_many,own,in_place", [(64, 20, False, False), (64, 20, True, True),
                                                     (32, 70, True, False), (128, 5, False, True)])
def test_host_pipeline_matches_oracle(gpu, oracle, n, how_many, own, in_place):
    """cf_conv3d_host: host-resident cubes through the overlapped pipeline.
    Groups of 16 MiB of whole cubes: 64^3 x20 -> groups of 8, 8, 4 (ragged);
    32^3 x70 -> 64 + 6; 128^3 x5 -> five groups of one over three slots."""
    print("case cf_conv3d_host N=%d how_many=%d own=%s in_place=%s" % (n, how_many, own, in_place), flush=True)
    P = n ** 3
    sig = oracle.uniform(how_many * P, 90000 + n + how_many)
    filt = oracle.uniform((how_many if own else 1) * P, 91000 + n)
    spec = gpu.Spectrum(n, _t(filt), count=how_many if own else 1)
    out = sig.copy() if in_place else np.zeros_like(sig)
    pass_ms, wall_ms = gpu.conv3d_host(spec, out if in_place else sig, out, how_many)
    assert 0 < pass_ms <= wall_ms
    for b in range(how_many):
        s = slice(b * P, (b + 1) * P)
        f = np.ascontiguousarray(filt[s] if own else filt)
        ref = oracle.conv3d_f64(n, np.ascontiguousarray(sig[s]), f, 1.0 / P)
        assert oracle.rel_l2(np.ascontiguousarray(out[s]), ref) <= REL_L2_FP32, b


@pytest.mark.parametrize("precision", [0, 1])
def test_host_pipeline_256_equals_device_path(gpu, oracle, precision):
    """4 x 256^3 through cf_conv3d_host equal, bit for bit, the same
    convolutions on device-resident data (same kernels, one cube per group,
    sixteen 8 MiB chunks per transfer)."""
    import torch
    api = gpu.api
    n, how_many = 256, 4
    print("case cf_conv3d_host 256^3 x4 precision=%d" % precision, flush=True)
    P = n ** 3
    if precision == api.CF_PREC_SPLIT_FP16:
        api.set_plan(256, api.CF_PLAN_SPLIT3_G24M2)
    dsig = torch.empty(how_many * P, dtype=torch.complex64, device="cuda")
    dfilt = torch.empty(how_many * P, dtype=torch.complex64, device="cuda")
    api.fill_uniform(dsig, 2 * how_many * P, 4242)
    api.fill_uniform(dfilt, 2 * how_many * P, 4343)
    spec = gpu.Spectrum(n, dfilt, count=how_many)
    api.set_plan(256, api.CF_PLAN_AUTO)
    dout = torch.empty_like(dsig)
    gpu.conv3d_device(spec, dsig, dout, how_many=how_many, precision=precision)
    want = _np(dout)
    hsig = _np(dsig)
    hout = np.zeros_like(hsig)
    gpu.conv3d_host(spec, hsig, hout, how_many, precision=precision)
    assert np.array_equal(hout.view(np.uint32), want.view(np.uint32))


def test_batch_graph_keeps_order_of_dependent_instances(gpu, oracle):
    """Instances of a graph batch whose buffers overlap keep the order they were
    given: a chain (the second reads the first's output) and two instances
    writing one output (the last one wins), with unrelated instances between."""
    import torch
    api = gpu.api
    n = 32
    P = n ** 3
    print("case batch graph dependent instances", flush=True)
    s1, s2 = oracle.uniform(P, 71001), oracle.uniform(P, 71002)
    f1, f2 = oracle.uniform(P, 71003), oracle.uniform(P, 71004)
    sp1, sp2 = gpu.Spectrum(n, _t(f1)), gpu.Spectrum(n, _t(f2))
    d1, mid, fin = _t(s1), torch.zeros(P, dtype=torch.complex64, device="cuda"), \
        torch.zeros(P, dtype=torch.complex64, device="cuda")
    # unrelated instances of other edges in between
    others = [(_t(oracle.uniform(m ** 3, 71100 + m)), m) for m in (64, 16, 64)]
    extra = [(x, torch.zeros_like(x), gpu.Spectrum(m, _t(oracle.uniform(m ** 3, 71200 + m)))) for x, m in others]
    chain = api.Batch([(d1, mid, sp1)] + extra + [(mid, fin, sp2)], mode=api.CF_BATCH_GRAPH)
    for rep in range(2):
        mid.zero_()
        fin.zero_()
        chain.launch()
        chain.check()
        r1 = oracle.conv3d_f64(n, s1, f1, 1.0 / P)
        r2 = oracle.conv3d_f64(n, np.ascontiguousarray(r1.astype(np.complex64)), f2, 1.0 / P)
        assert oracle.rel_l2(_np(fin), r2) <= REL_L2_FP32, rep
    chain.close()
    d2 = _t(s2)
    same = api.Batch([(d1, fin, sp1)] + extra + [(d2, fin, sp2)], mode=api.CF_BATCH_GRAPH)
    fin.zero_()
    same.launch()
    same.check()
    assert oracle.rel_l2(_np(fin), oracle.conv3d_f64(n, s2, f2, 1.0 / P)) <= REL_L2_FP32
    same.close()


def test_persistent_batch_refuses_overlapping_instances(gpu, oracle):
    """Instances of a persistent batch run concurrently: an output that
    overlaps another instance's signal or output, or part of its own signal,
    is refused (graph mode keeps the order and accepts a chain)."""
    import torch
    api = gpu.api
    n = 32
    P = n ** 3
    print("case persistent batch overlap check", flush=True)
    spec = gpu.Spectrum(n, _t(oracle.uniform(P, 73001)))
    buf = torch.zeros(3 * P, dtype=torch.complex64, device="cuda")
    a, b, c = buf[:P], buf[P:2 * P], buf[2 * P:]
    for inst in ([(a, b, spec), (b, c, spec)],        # chain: 2 reads 1's output
                 [(a, c, spec), (b, c, spec)],        # one output for two instances
                 [(a, buf[P // 2:P // 2 + P], spec)]):  # output half over its own signal
        with pytest.raises(gpu.ConvError, match="overlap"):
            api.Batch(inst, mode=api.CF_BATCH_PERSISTENT)
    api.Batch([(a, b, spec), (b, c, spec)], mode=api.CF_BATCH_GRAPH).close()
    api.Batch([(a, a, spec), (b, b, spec)], mode=api.CF_BATCH_PERSISTENT).close()  # in place: fine


def test_batch_shared_spectrum_in_place(gpu, oracle):
    n, k = 64, 4
    print("case batch shared spectrum in place", flush=True)
    f = oracle.uniform(n ** 3, 70001)
    gpu.api.set_plan(n, gpu.api.CF_PLAN_AXIS5)  # natural order, as the persistent batch reads it
    try:
        spec = gpu.Spectrum(n, _t(f))
    finally:
        gpu.api.set_plan(n, gpu.api.CF_PLAN_AUTO)
    sigs = [oracle.uniform(n ** 3, 70100 + i) for i in range(k)]
    ds = [_t(s) for s in sigs]
    b = gpu.api.Batch([(d, d, spec) for d in ds])
    b.launch()
    b.check()
    for d, s in zip(ds, sigs):
        assert oracle.rel_l2(_np(d), oracle.conv3d_f64(n, s, f, 1.0 / n ** 3)) <= REL_L2_FP32
    b.close()


def test_batch_stalled_dependency_is_reported(gpu, oracle):
    """The persistent kernel's bounded wait: a dependency made unsatisfiable
    from the host (with a short spin limit) ends the launch with an error
    instead of a hang; a batch built afterwards runs normally."""
    print("case batch forced stall", flush=True)
    n = 32
    f = oracle.uniform(n ** 3, 71001)
    spec = gpu.Spectrum(n, _t(f))
    d = _t(oracle.uniform(n ** 3, 71002))
    b = gpu.api.Batch([(d, d.clone(), spec)], stall_test_spin_limit=4096)
    b.launch()
    with pytest.raises(gpu.ConvError, match="dependency"):
        b.check()
    b.close()
    out = d.clone()
    ok = gpu.api.Batch([(d, out, spec)])
    ok.launch()
    ok.check()
    ok.close()


@pytest.mark.parametrize("plan", [4, 12])
def test_batch_rejects_other_plans_in_persistent_mode(gpu, oracle, plan):
    """The persistent batch runs the axis, plane and two four-step plans:
    spectra of the others (here g24 and the z-blocked plan) are refused."""
    api = gpu.api
    api.set_plan(256, plan)
    try:
        spec = gpu.Spectrum(256, _t(oracle.uniform(256 ** 3, 72001)))
    finally:
        api.set_plan(256, api.CF_PLAN_AUTO)
    d = _t(oracle.uniform(256 ** 3, 72002))
    with pytest.raises(gpu.ConvError, match="persistent batch runs"):
        api.Batch([(d, d, spec)], mode=api.CF_BATCH_PERSISTENT)


REL_L2_SPLIT = 1e-5  # stated bound of the split-FP16 MFMA form against the FP64 oracle


@pytest.mark.parametrize("dist", ["uniform", "loguniform"])
def test_split_fp16_256_matches_oracle(gpu, oracle, dist):
    """Split-FP16 MFMA form at 256^3, two instances with their own spectra:
    rel-L2 against the FP64 oracle within the stated bound, for U[-1,1) and
    for log-uniform magnitudes (1e-3..1e3, uniform phase); in place gives
    the same bits."""
    import torch
    n, b = 256, 2
    P = n ** 3
    print("case split-FP16 N=256 %s" % dist, flush=True)
    api = gpu.api
    gen = oracle.uniform if dist == "uniform" else oracle.loguniform
    sig, filt = gen(b * P, 18001), gen(b * P, 18002)
    api.set_plan(n, api.CF_PLAN_SPLIT3_G24M2)
    try:
        spec = gpu.Spectrum(n, _t(filt), count=b)
        ds = _t(sig)
        d16, d32 = torch.empty_like(ds), torch.empty_like(ds)
        gpu.conv3d_device(spec, ds, d16, how_many=b, precision=api.CF_PREC_SPLIT_FP16)
        gpu.conv3d_device(spec, ds, d32, how_many=b)
        g16, g32 = _np(d16), _np(d32)
        for i in range(b):
            s = slice(i * P, (i + 1) * P)
            ref = oracle.conv3d_f64(n, np.ascontiguousarray(sig[s]), np.ascontiguousarray(filt[s]), 1.0 / P)
            e16 = oracle.rel_l2(np.ascontiguousarray(g16[s]), ref)
            e32 = oracle.rel_l2(np.ascontiguousarray(g32[s]), ref)
            print("  instance %d: rel-L2 split-FP16 %.3e, FP32 %.3e" % (i, e16, e32), flush=True)
            assert e16 <= REL_L2_SPLIT
            assert e32 <= REL_L2_FP32
        gpu.conv3d_device(spec, ds, ds, how_many=b, precision=api.CF_PREC_SPLIT_FP16)
        assert np.array_equal(_np(ds), g16)
        del spec, ds, d16, d32
        torch.cuda.empty_cache()
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)


@pytest.mark.parametrize("dist,plan", [("uniform", "AXIS5Z"), ("loguniform", "AXIS5")])
def test_split_fp16_512_matches_oracle(gpu, oracle, dist, plan):
    """Split-FP16 MFMA form of the 512^3 axis passes on the natural (axis5)
    and the z-blocked (axis5z) spectrum layout: rel-L2 against the FP64
    oracle within the stated bound, for U[-1,1) and for log-uniform inputs;
    in place gives the same bits."""
    import torch
    n = 512
    P = n ** 3
    print("case split-FP16 N=512 %s plan %s" % (dist, plan), flush=True)
    api = gpu.api
    gen = oracle.uniform if dist == "uniform" else oracle.loguniform
    sig, filt = gen(P, 18101), gen(P, 18102)
    api.set_plan(n, getattr(api, "CF_PLAN_" + plan))
    try:
        spec = gpu.Spectrum(n, _t(filt))
        ds = _t(sig)
        d16 = torch.empty_like(ds)
        gpu.conv3d_device(spec, ds, d16, precision=api.CF_PREC_SPLIT_FP16)
        g16 = _np(d16)
        del d16
        ref = oracle.conv3d_f64(n, sig, filt, 1.0 / P)
        e16 = oracle.rel_l2(g16, ref)
        print("  rel-L2 split-FP16 %.3e" % e16, flush=True)
        assert e16 <= REL_L2_SPLIT
        del ref
        gpu.conv3d_device(spec, ds, ds, precision=api.CF_PREC_SPLIT_FP16)
        assert np.array_equal(_np(ds), g16)
        del spec, ds
        torch.cuda.empty_cache()
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)


def test_split_fp16_rejects_other_layouts(gpu, oracle):
    import torch
    print("case split-FP16 argument checks", flush=True)
    api = gpu.api
    f64 = _t(oracle.uniform(64 ** 3, 5))
    s64 = gpu.Spectrum(64, f64)
    x = torch.empty_like(f64)
    with pytest.raises(api.ConvError):
        gpu.conv3d_device(s64, f64, x, precision=api.CF_PREC_SPLIT_FP16)
    api.set_plan(256, api.CF_PLAN_SPLIT3_G24)  # middle tiles of 4 x: another spectrum layout
    try:
        f = _t(oracle.uniform(256 ** 3, 6))
        s = gpu.Spectrum(256, f)
        y = torch.empty_like(f)
        with pytest.raises(api.ConvError):
            gpu.conv3d_device(s, f, y, precision=api.CF_PREC_SPLIT_FP16)
        with pytest.raises(api.ConvError):
            gpu.conv3d_device(s, f, y, precision=7)
        del s, f, y
    finally:
        api.set_plan(256, api.CF_PLAN_AUTO)
