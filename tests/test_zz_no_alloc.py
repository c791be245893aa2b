"""The allocation-free entry point (cf_conv3d_device_noalloc) against
cf_conv3d_device, through the C ABI: bit-identical outputs on consecutive
calls that reuse the one scratch array (the second call smaller than the
first, the third a repeat of the first), the size limit, and a plan with no
intermediate."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _same(a, b):
    import torch
    torch.cuda.synchronize()
    return torch.equal(a.view(torch.int64), b.view(torch.int64))


def test_noalloc_matches_conv3d_device(gpu, oracle):
    import torch
    api = gpu.api
    print("case noalloc 256^3 (autotuned plan), then 128^3 x 2 (four-step), then 256^3 again", flush=True)
    s256 = torch.empty(256 ** 3, dtype=torch.complex64, device="cuda")
    f256 = torch.empty_like(s256)
    api.fill_uniform(s256, 2 * s256.numel(), 81001)
    api.fill_uniform(f256, 2 * f256.numel(), 81002)
    spec256 = gpu.Spectrum(256, f256)
    assert api.plan_supported(256, spec256.plan)
    want256, got256 = torch.empty_like(s256), torch.full_like(s256, float("nan"))
    gpu.conv3d_device(spec256, s256, want256)
    api.conv3d_device_noalloc(spec256, s256, got256)
    assert _same(got256, want256)

    n, b = 128, 2
    P = n ** 3
    api.set_plan(n, api.CF_PLAN_SPLIT3_G24M2)  # a plan with an intermediate in the scratch
    try:
        sig, filt = oracle.uniform(b * P, 81003), oracle.uniform(b * P, 81004)
        spec = gpu.Spectrum(n, _t(filt), count=b)
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    ds = _t(sig)
    want, got = torch.empty_like(ds), torch.full_like(ds, float("nan"))
    gpu.conv3d_device(spec, ds, want, how_many=b)
    api.conv3d_device_noalloc(spec, ds, got, how_many=b)  # smaller than the first call
    assert _same(got, want)
    host = got.cpu().numpy()
    for k in range(b):
        ref = oracle.conv3d_f64(n, np.ascontiguousarray(sig[k * P:(k + 1) * P]),
                                np.ascontiguousarray(filt[k * P:(k + 1) * P]), 1.0 / P)
        assert oracle.rel_l2(np.ascontiguousarray(host[k * P:(k + 1) * P]), ref) <= 1e-5

    got256.fill_(float("nan"))
    api.conv3d_device_noalloc(spec256, s256, got256)  # the scratch holds the 128^3 call's data now
    assert _same(got256, want256)


def test_noalloc_limit_and_plans_without_intermediate(gpu, oracle):
    import torch
    api = gpu.api
    print("case noalloc size limit (9 x 128^3; 5 x 128^3 in place with two intermediates) and the plane plan at 64^3",
          flush=True)
    n = 128
    api.set_plan(n, api.CF_PLAN_SPLIT3_G24M2)
    try:
        spec = gpu.Spectrum(n, _t(oracle.uniform(n ** 3, 82001)))
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    nine = torch.zeros(9 * n ** 3, dtype=torch.complex64, device="cuda")
    with pytest.raises(gpu.ConvError, match="scratch"):
        api.conv3d_device_noalloc(spec, nine, nine, how_many=9)  # 144 MiB of intermediate
    # two intermediates (the h form) in place: 2 x 16 MiB per cube, 8 cubes fit but not 5 in place
    api.set_plan(n, api.CF_PLAN_SPLIT3_G24M2H)
    try:
        sh = gpu.Spectrum(n, _t(oracle.uniform(n ** 3, 82004)))
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    five = nine[: 5 * n ** 3]
    with pytest.raises(gpu.ConvError, match="scratch"):
        api.conv3d_device_noalloc(sh, five, five, how_many=5)  # 160 MiB of intermediates in place
    four = _t(oracle.uniform(4 * n ** 3, 82005))
    want4 = torch.empty_like(four)
    gpu.conv3d_device(sh, four, want4, how_many=4)
    api.conv3d_device_noalloc(sh, four, four, how_many=4)  # 128 MiB in place: fits
    assert _same(four, want4)
    m = 64
    api.set_plan(m, api.CF_PLAN_PLANE3)  # no intermediate: any count
    try:
        sm = gpu.Spectrum(m, _t(oracle.uniform(m ** 3, 82002)))
    finally:
        api.set_plan(m, api.CF_PLAN_AUTO)
    x = _t(oracle.uniform(3 * m ** 3, 82003))
    want, got = torch.empty_like(x), torch.empty_like(x)
    gpu.conv3d_device(sm, x, want, how_many=3)
    api.conv3d_device_noalloc(sm, x, got, how_many=3)
    assert _same(got, want)
