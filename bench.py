"""Benchmark: 256^3 complex-FP32 FFT convolution on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2_256]

Default workload (the metric's configuration): a step is one convolution of
one device-resident 256^3 complex64 signal with a filter whose spectrum was
computed once before timing -- the schedule chosen by the library's
autotuner (3-pass four-step or 5-pass axis-by-axis), ending with the 1/N^3
scale.  Inputs are synthetic (seeded U[-1,1) re/im, generated in HBM by the
library).  Prints one JSON line.

Per-pass device times come from HIP events recorded by the library on the
launch stream during the timed loop, on one step in PROFILE_EVERY, or on
one step of a loop shorter than 2 * PROFILE_EVERY (the event packets between
the passes of an instrumented step cost the GPU tens of microseconds, so
instrumenting every step would slow the loop it measures); the roofline
entry is for the pass with the largest time.  The cpu_baseline leg times the oracle's FP32 CPU
convolution (OpenMP, all host threads) on a bounded sample of the same
workload and, on the same inputs, checks the GPU result against it and
against the FP64 oracle.  Beside the FP32 form the line carries the same
convolution in the split-FP16 MFMA form (its time, its speed relative to FP32
and its rel-L2 against the FP64 oracle), the second half of the metric.

Other workloads (reported lines, not the metric): c2_128 / c1_64 / c5_512 (one cube),
c3_64x64 (64 cubes of 64^3, one shared spectrum), c4_mixed (16x32^3 + 8x64^3 +
4x128^3 + 1x256^3, own spectra, shuffled), each comparing the persistent
one-launch kernel, the recorded graph and, where it applies, batched passes;
c5_split (4x256^3 + 1x512^3, own spectra, U[-1,1) and log-uniform inputs): the
split-FP16 MFMA form beside the FP32 form at both edges, rel-L2 of every
instance against the FP64 oracle.
"""
import argparse
import json
import math
import os
import random
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X datasheet HBM3E bandwidth
PROFILE_EVERY = 20  # steps of the timed loop that carry per-pass events: one in PROFILE_EVERY

SINGLE = {"c2_256": 256, "c2_128": 128, "c1_64": 64, "c5_512": 512}
SPLIT = {"c5_split": [256] * 4 + [512]}
BATCHES = {
    "c3_64x64": [64] * 64,
    "c4_mixed": [32] * 16 + [64] * 8 + [128] * 4 + [256],
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2_256", choices=sorted(SINGLE) + sorted(BATCHES) + sorted(SPLIT))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--plan", type=int, default=0, help="force a schedule (CF_PLAN_*); 0 = autotuned")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--host-calls", action="store_true",
                    help="also time the drop-in calls on host buffers (PCIe included; not the metric)")
    ap.add_argument("--power", action="store_true",
                    help="also sample GPU power (amdsmi) over a separate 2 s loop of the same convolution")
    return ap.parse_args()


def pocketfft_leg(n, sig, H, reps, threads, ref):
    """The same FP32 convolution (forward FFT, product with the given spectrum,
    unnormalised inverse) through scipy's pocketfft, an FFTW-class library,
    on as many threads: how far the oracle's radix-2 FFT is from a tuned CPU
    FFT on this host; its output is compared with the oracle's (ref).  None
    when scipy is not importable."""
    try:
        import scipy.fft as sf
    except ImportError:
        return None
    import numpy as np
    x = sig.reshape(n, n, n)
    h = H.reshape(n, n, n)

    def conv():
        y = sf.fftn(x, workers=threads)
        y *= h
        return sf.ifftn(y, workers=threads, norm="forward")
    conv()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = conv()
        times.append(time.perf_counter() - t0)
    assert out.dtype == np.complex64
    r = ref.reshape(n, n, n)
    err = float(np.linalg.norm((out - r).ravel()) / np.linalg.norm(r.ravel()))
    return {"value": 1.0 / statistics.median(times), "unit": "conv/s", "cores": threads, "rel_l2_vs_oracle_f32": err,
            "library": "scipy.fft (pocketfft) %s, complex64" % __import__("scipy").__version__,
            "sample": "%d reps of the same %d^3 convolution, median" % (reps, n)}


def cpu_baseline_leg(n, gpu_out_host, seed_sig, seed_filt, reps, split_out_host=None):
    """Times the oracle's FP32 CPU convolution (port of src/fft_conv3D.cpp:50-177)
    on `reps` convolutions of one cube, and checks the GPU output on the same
    inputs.  With batch_cpu_leg and split_check_leg (the batch and c5
    checkers), the only places bench.py uses the oracle."""
    import numpy as np
    from oracle import oracle as O
    O.build()
    P = n ** 3
    sig = O.uniform(P, seed_sig)
    filt = O.uniform(P, seed_filt)
    H = O.fft3d_f32(filt, -1)  # filter spectrum once, outside the timing (fft_conv3D.cpp:136-139)
    out = np.empty_like(sig)
    O.conv3d_f32_spec(n, sig, H, out)  # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.conv3d_f32_spec(n, sig, H, out)  # input re-copied every rep
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    # one thread as well (the reference CLI's default, src/helper.cpp:68), on fewer reps
    t1 = []
    for _ in range(max(1, min(2, reps))):
        t0 = time.perf_counter()
        O.conv3d_f32_spec(n, sig, H, out, nthreads=1)
        t1.append(time.perf_counter() - t0)
    res = {
        "value": 1.0 / med, "unit": "conv/s", "cores": O.max_threads(), "kind": "port",
        "sample": "%d reps of one %d^3 FP32 convolution (FFT, product, inverse FFT; filter spectrum "
                  "precomputed), median; oracle's own radix-2 FFT with OpenMP (FFTW3 %s)"
                  % (reps, n, "present" if O.fftw_available() else "not installed"),
        "value_1_thread": 1.0 / statistics.median(t1), "cpu_model": cpu_model(),
    }
    res["fftw_class"] = pocketfft_leg(n, sig, H, reps, O.max_threads(), out)
    check = {}
    if gpu_out_host is not None:
        first = np.ascontiguousarray(gpu_out_host[:P])
        ref32 = out * np.complex64(1.0 / P)  # the GPU path applies 1/N^3; exact power-of-two scale
        check["snr_db_vs_fp32_cpu"] = O.snr_db(np.ascontiguousarray(ref32), first)
        ref64 = O.conv3d_f64(n, sig, filt, 1.0 / P)
        check["rel_l2_vs_fp64"] = O.rel_l2(first, ref64)
        if split_out_host is not None:
            check["rel_l2_vs_fp64_split_fp16_mfma"] = O.rel_l2(np.ascontiguousarray(split_out_host[:P]), ref64)
    return res, check


def batch_cpu_leg(by_n, shared, seed, outs, reps):
    """cpu_baseline of a batch workload: the oracle's FP32 convolution of one
    cube of the batch's largest edge, timed as cpu_baseline_leg times it, and
    the checker: every instance's output in every mode against the FP64
    oracle, on the inputs the bench generated on the device (the oracle's
    generator is the device generator's bit-exact port)."""
    import numpy as np
    from oracle import oracle as O
    nmax = max(by_n)
    res, _ = cpu_baseline_leg(nmax, None, seed + nmax, seed + 7 * nmax, reps)
    worst = {mode: 0.0 for mode in outs}
    for n, cnt in by_n.items():
        P = n ** 3
        sigs = O.uniform(cnt * P, seed + n)  # the stream the bench filled the edge's signals from
        filts = O.uniform((1 if shared else cnt) * P, seed + 7 * n)
        for k in range(cnt):
            f = filts[:P] if shared else filts[k * P:(k + 1) * P]
            ref = O.conv3d_f64(n, np.ascontiguousarray(sigs[k * P:(k + 1) * P]), np.ascontiguousarray(f), 1.0 / P)
            for mode, o in outs.items():
                worst[mode] = max(worst[mode], O.rel_l2(np.ascontiguousarray(o[n][k * P:(k + 1) * P]), ref))
    return res, {"rel_l2_vs_fp64_max_by_mode": worst, "instances_checked": sum(by_n.values())}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_bandwidth(torch, nelem, reps=5):
    """GB/s of torch's device copy of nelem complex64 values (read + write)."""
    from convfpga_amd import api
    a = torch.empty(nelem, dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    api.fill_uniform(a, 2 * nelem, 99)  # random data: the memory system is faster on zeros
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2.0 * 8 * nelem / (ms * 1e-3) / 1e9


def base_line(a, metric, value, unit, ms_per_step, config):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32", "data": "synthetic (seeded U[-1,1) re/im generated on the GPU)", "config": config}


KERNEL_OF_PASS = {  # kernel of each pass name of the axis and plane schedules (N = edge)
    "x_fwd": "cf_row_fwd_{N}", "x_inv": "cf_row_inv_{N}", "y_fwd": "cf_coly_fwd_{N}", "y_inv": "cf_coly_inv_{N}",
    "z_fwd_mul_z_inv": "cf_colz_conv_{N}", "xy_fwd": "cf_plane_fwd_{N}", "xy_inv": "cf_plane_inv_{N}",
}


def kernel_of_pass(name, n, plan_name):
    """Kernel of pass `name` under a plan: four-step plans split3_g<GA><GB>[m<GM>][o|h|h<GBO>][c<GAO>]
    (o: second intermediate of x group GM; h: y2 pairs interleaved; h<GBO>: the
    same in x groups of GBO; c<GAO>: last-pass tiles of GAO y2 values, reading a
    second intermediate with k1 inside the x group).  A first pass of GA = 1
    shares the middle pass of GA = 2 (the middle pass depends on GB only)."""
    if plan_name.startswith("split3_"):
        v = plan_name.split("_")[1]
        v, _, gao = v.partition("c")
        g = v.split("m")[0]
        gi = g
        if v.endswith("o"):  # the last pass reads x groups of GM
            gi = g[:-1] + v[-2]
        elif v.endswith("h"):
            gi = g + "h"
        elif "h" in v:  # h<GBO>
            gi = g[:-1] + v.split("h")[1] + "h"
        vm = "g2" + v[2:] if v[1] == "1" else v
        if gao:  # its second intermediate has k1 inside the x group (kernels ...k)
            gi = "g" + gao + gi[2:] + "k"
            vm += "k"
        return {"xy1_fwd": "cf_xy1_fwd_%d_%s" % (n, g), "xy1_inv": "cf_xy1_inv_%d_%s" % (n, gi),
                "y2z_fwd_mul_inv": "cf_mid_%d_%s" % (n, vm)}.get(name, "")
    k = KERNEL_OF_PASS.get(name, "").format(N=n)
    if plan_name == "axis5z" and name in ("y_fwd", "y_inv", "z_fwd_mul_z_inv"):
        k = k.replace("_%d" % n, "_zb_%d" % n)  # the z-blocked column kernels
    return k


def run_single(a, torch, C, api):
    n = SINGLE[a.workload]
    P = n ** 3
    seed_sig, seed_filt = a.seed, a.seed + 1
    sig = torch.empty(P, dtype=torch.complex64, device="cuda")
    filt = torch.empty(P, dtype=torch.complex64, device="cuda")
    api.fill_uniform(sig, 2 * P, seed_sig)
    api.fill_uniform(filt, 2 * P, seed_filt)
    torch.cuda.synchronize()
    if a.plan:
        api.set_plan(n, a.plan)
        plan, tune_ms = a.plan, {}
    else:
        plan, tune_ms = api.autotune(n)  # times every supported schedule once, keeps the fastest
    spec = C.Spectrum(n, filt)  # filter setup, outside the timed region
    assert spec.plan == plan
    out = torch.empty_like(sig)
    # plain copy kernel of the same footprint: measured HBM bandwidth beside the
    # datasheet peak, and the calibration dispatch for the HBM byte counters
    copy_gbs = copy_bandwidth(torch, max(P, 1 << 24))

    # Per-pass HIP events on every PROFILE_EVERY-th call (or one call of a
    # loop shorter than 2 * PROFILE_EVERY) counted from the first warm-up
    # call, with the interval longer than the warm-up so that only timed steps
    # are instrumented (an instrumented step costs the GPU tens of
    # microseconds: the event packets between its passes).  Armed before the
    # warm-up, so that the GPU goes from the warm-up straight into the timed loop.
    every = PROFILE_EVERY if a.steps >= 2 * PROFILE_EVERY else a.steps
    every = max(every, a.warmup + 1)
    recorded = (a.warmup + a.steps) // every - a.warmup // every
    api.profile_begin(recorded, every)
    for _ in range(a.warmup):
        C.conv3d_device(spec, sig, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        C.conv3d_device(spec, sig, out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    passes, ncalls = api.profile_end()
    assert ncalls == recorded
    wall = t1 - t0
    # the same loop without any per-pass events
    t2 = time.perf_counter()
    for _ in range(a.steps):
        C.conv3d_device(spec, sig, out)
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - t2
    alg_bytes = sum(p["bytes"] for p in passes)  # per step
    for p in passes:
        p["GBps"] = p["bytes"] / (p["ms"] * 1e-3) / 1e9
        p["frac_of_peak"] = p["GBps"] / HBM_PEAK_GBS
    dom = max(passes, key=lambda p: p["ms"])
    kname = kernel_of_pass(dom["name"], n, api.PLAN_NAMES[plan])
    traffic = None  # HBM bytes per launch from the committed rocprof counter runs (FETCH_SIZE, WRITE_SIZE)
    tfile = os.path.join(ROOT, "profiles", "traffic_256.json")
    if n == 256 and os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(kname, {}).get("traffic_bytes")
    metric = ("3D convs/sec + % HBM roofline at 256\u00b3 c2c; split-MFMA vs FP32 at rel-L2\u22641e-5" if n == 256
              else "3D convs/sec + %% HBM roofline at %d^3 c2c" % n)
    line = base_line(a, metric, a.steps / wall, "conv/s",
                     1e3 * wall / a.steps,
                     {"workload": "conv3d_%d^3_c2c_fp32_batch1_precomputed_filter_spectrum" % n, "N": n, "batch": 1,
                      "plan": api.PLAN_NAMES[plan], "passes": len(passes), "scale": "1/N^3",
                      "algorithmic_bytes_per_step": alg_bytes})
    line["roofline"] = {"bound": "hbm", "kernel": kname, "pass": dom["name"], "achieved": dom["GBps"],
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": traffic}
    line["conv_roofline_frac"] = alg_bytes * a.steps / wall / 1e9 / HBM_PEAK_GBS
    # the same rate priced with the 5-pass model: 11 volume transfers per convolution
    line["conv_roofline_frac_5pass_model"] = 11 * 8.0 * P * a.steps / wall / 1e9 / HBM_PEAK_GBS
    line["device_ms_per_step"] = sum(p["ms"] for p in passes)
    line["conv_per_s_without_pass_events"] = a.steps / wall_plain
    line["steps_with_pass_events"] = recorded
    # for information: the compulsory floor (read signal and spectrum, write the
    # output: 3V at the HBM peak) and the arithmetic (two 3D FFTs, the product)
    floor_us = 3 * 8.0 * P / (HBM_PEAK_GBS * 1e9) * 1e6
    flops = 10.0 * P * math.log2(P) + 8.0 * P
    line["compulsory_floor_3V"] = {"us": floor_us, "frac_of_measured": floor_us / (1e6 * wall / a.steps)}
    line["flops_per_conv"] = flops
    line["gflops"] = flops * a.steps / wall / 1e9
    line["copy_kernel_GBps"] = copy_gbs
    line["autotune_ms_per_conv"] = {api.PLAN_NAMES[p]: round(v, 5) for p, v in tune_ms.items()}
    line["passes"] = [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in p.items()} for p in passes]
    # the timed convolution's output, checked against the oracle in the cpu_baseline leg
    host_out = out.cpu().numpy() if not a.no_cpu_baseline else None
    scratch = torch.empty_like(out)  # the legs below write their own outputs here
    line["own_filter"] = own_filter_leg(torch, C, n, sig, filt, scratch)
    if n <= 128:  # small cubes: all passes in one launch
        line["one_launch_persistent"] = one_launch_leg(a, torch, C, api, n, sig, filt, scratch, 7 * 8.0 * P)
    line["slab_share"] = share_leg(C, n, spec, sig, scratch)
    split_host = None
    if n == 256:
        line["split_fp16_mfma"], split_host = split_leg(a, torch, C, api, n, spec, sig, filt, wall_plain / a.steps)
    if a.host_calls:
        line["dropin_host_calls"] = dropin_host_leg(C, n, sig, filt)
        line["host_batch"] = host_batch_leg(torch, C, n, spec, 4)
    if a.power:
        line["power"] = power_leg(torch, C, spec, sig, out)
    return line, (n, host_out, seed_sig, seed_filt, split_host)


def split_leg(a, torch, C, api, n, spec, sig, filt, fp32_s):
    """The same convolution in the split-FP16 MFMA form (radix-16 DFTs as
    hi/lo FP16 complex GEMMs with FP32 accumulation), timed like the FP32
    loop without events; its output is checked against the FP64 oracle in the
    cpu_baseline leg."""
    sp = spec
    if spec.plan != api.CF_PLAN_SPLIT3_G24M2:  # the MFMA kernels read this plan's spectrum layout
        api.set_plan(n, api.CF_PLAN_SPLIT3_G24M2)
        sp = C.Spectrum(n, filt)
        api.set_plan(n, api.CF_PLAN_AUTO)
    out2 = torch.empty_like(sig)
    for _ in range(a.warmup):
        C.conv3d_device(sp, sig, out2, precision=api.CF_PREC_SPLIT_FP16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        C.conv3d_device(sp, sig, out2, precision=api.CF_PREC_SPLIT_FP16)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    res = {"ms_per_conv": 1e3 * t, "conv_per_s": 1.0 / t, "speed_vs_fp32": fp32_s / t,
           "fp32_ms_per_conv_same_loop": 1e3 * fp32_s}
    return res, (out2.cpu().numpy() if not a.no_cpu_baseline else None)


def power_leg(torch, C, spec, sig, out, seconds=2.0, idle_s=0.5):
    """Board power while the convolution runs back to back for `seconds`
    (outside the timed region), the idle power before it, and convolutions
    per joule from the GPU's energy counter (sampled power if it has none)."""
    from convfpga_amd.power import PowerMonitor, PowerUnavailable, SmiReader, torch_device_bdf
    try:
        rd = SmiReader(torch_device_bdf(0))
    except PowerUnavailable as e:
        return {"unavailable": str(e)}
    try:
        torch.cuda.synchronize()
        idle = PowerMonitor(rd, 0.01).start()
        time.sleep(idle_s)
        idle_s = idle.stop()
        mon = PowerMonitor(rd, 0.01).start()
        count, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            for _ in range(100):
                C.conv3d_device(spec, sig, out)
            torch.cuda.synchronize()
            count += 100
        s = mon.stop()
    finally:
        rd.close()
    joules = s.get("energy_J") or ((s["mean_W"] or 0.0) * s["seconds"])
    res = {"gpu": rd.bdf, "idle_W": idle_s["mean_W"], "busy_mean_W": s["mean_W"], "busy_max_W": s["max_W"],
           "energy_counter_mean_W": s.get("energy_mean_W"), "samples": s["samples"], "convs": count,
           "conv_per_s": count / s["seconds"], "conv_per_J": count / joules if joules else None}
    return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in res.items()}


def share_leg(C, n, spec, sig, out, shares=8, reps=20):
    """One GPU's share of the convolution decomposed over `shares` GPUs (slabs
    for the outer passes, pencils for the passes along z): every pass on
    1/shares of its tiles at the real shapes, exchanges not included.  The
    exchange volume is stated, not timed."""
    sums = None
    try:
        C.api.conv3d_share_timed(spec, sig, out, shares)
    except C.api.ConvError as e:  # plans with too few tiles per pass to split (small cubes)
        return {"shares": shares, "unavailable": str(e)}
    for _ in range(reps):
        ps = C.api.conv3d_share_timed(spec, sig, out, shares)
        if sums is None:
            sums = [[p["name"], 0.0, p["bytes"]] for p in ps]
        for acc, p in zip(sums, ps):
            acc[1] += p["ms"]
    passes = [{"name": nm, "us": round(1e3 * ms / reps, 2), "GBps": b / (ms / reps * 1e-3) / 1e9}
              for nm, ms, b in sums]
    V = 8.0 * n ** 3
    return {"shares": shares, "passes": passes, "us_sum": round(sum(p["us"] for p in passes), 2),
            "all_to_all_bytes_out_per_gpu": V / shares * (shares - 1) / shares, "all_to_alls_per_conv": 2}


def own_filter_leg(torch, C, n, sig, filt, out, reps=20):
    """The convolution with its own filter (device-resident): filter transform
    (spectrum creation, including its allocation) plus the convolution, per
    call, and the filter setup alone.  Reported beside the metric, whose
    spectrum is precomputed."""
    def one():
        sp = C.Spectrum(n, filt)
        C.conv3d_device(sp, sig, out)
        sp.close()
    one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        one()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        C.Spectrum(n, filt).close()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return {"ms_per_conv_with_filter_transform": 1e3 * (t1 - t0) / reps,
            "filter_setup_ms": 1e3 * (t2 - t1) / reps, "reps": reps}


def dropin_host_leg(C, n, sig, filt, reps=3):
    """The drop-in calls on pageable host buffers (what the reference's callers
    pass): every phase including the PCIe copies and the filter transform, per
    call.  Not the metric -- reported beside it."""
    import numpy as np
    hs, hf = sig.cpu().numpy(), filt.cpu().numpy()
    ho = np.empty_like(hs)
    res = {}
    for name, fn in (("fpgaf_conv3D", C.fpgaf_conv3D), ("fpgaf_conv3D_svm", C.fpgaf_conv3D_svm)):
        walls, fields = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            t = fn(n, hs, hf, ho)
            walls.append(time.perf_counter() - t0)
            if not t.valid:
                raise SystemExit("%s returned valid=false: %s" % (name, C.api.last_error()))
            fields.append(t.as_dict())
        k = min(range(reps), key=lambda i: walls[i])
        res[name] = {"wall_ms_best": round(1e3 * walls[k], 3), "conv_per_s": 1.0 / walls[k],
                     "fpga_t_ms": {f: round(v, 4) for f, v in fields[k].items() if f != "valid"}}
    return res


def host_batch_leg(torch, C, n, spec, count, reps=3):
    """`count` cubes in pageable host memory through cf_conv3d_host (uploads,
    passes and downloads overlapped); compare ms_per_cube with the sequential
    phases of fpgaf_conv3D in dropin_host_calls."""
    import numpy as np
    P = n ** 3
    d = torch.empty(count * P, dtype=torch.complex64, device="cuda")
    C.api.fill_uniform(d, 2 * count * P, 77)
    hs = d.cpu().numpy()
    ho = np.empty_like(hs)
    walls = []
    for _ in range(reps):
        walls.append(C.conv3d_host(spec, hs, ho, count)[1])
    w = min(walls)
    return {"cubes": count, "pipelined_wall_ms": round(w, 3), "ms_per_cube": round(w / count, 3),
            "conv_per_s": 1e3 * count / w, "host_bus_GBps_both_ways": 2 * 8.0 * P * count / (w * 1e-3) / 1e9}


def persistent_plan(api, n):
    """Plan of the spectra of persistent-mode batches: the plane plan at 32^3
    and 64^3, the four-step plan with interleaved y2 pairs at 128^3 and 256^3
    (the plans the persistent kernel runs), the axis plan otherwise."""
    return {32: api.CF_PLAN_PLANE3, 64: api.CF_PLAN_PLANE3, 128: api.CF_PLAN_SPLIT3_G24M2H,
            256: api.CF_PLAN_SPLIT3_G24M2H}.get(n, api.CF_PLAN_AXIS5)


def one_launch_leg(a, torch, C, api, n, sig, filt, out, alg_bytes):
    """The same convolution as one launch of the persistent kernel (a batch of
    one instance, all three passes in one kernel, stages ordered on the
    device), timed like the main loop; its output is checked against the main
    loop's (bitwise: the same passes)."""
    api.set_plan(n, persistent_plan(api, n))
    try:
        sp = C.Spectrum(n, filt)
    finally:
        api.set_plan(n, api.CF_PLAN_AUTO)
    o1 = torch.empty_like(out)
    b = api.Batch([(sig, o1, sp)], mode=api.CF_BATCH_PERSISTENT)
    for _ in range(a.warmup):
        b.launch()
    b.check()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        b.launch()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / a.steps
    b.check()
    o2 = torch.empty_like(out)
    C.conv3d_device(sp, sig, o2)
    torch.cuda.synchronize()
    same = bool(torch.equal(o1.view(torch.int64), o2.view(torch.int64)))
    b.close()
    return {"plan": api.PLAN_NAMES[sp.plan], "ms_per_conv": 1e3 * t, "conv_per_s": 1.0 / t,
            "roofline_frac_7V": alg_bytes / t / 1e9 / HBM_PEAK_GBS, "bitwise_equal_to_three_launches": same}


def run_batch(a, torch, C, api):
    sizes = list(BATCHES[a.workload])
    shared = a.workload == "c3_64x64"
    if not shared:
        random.Random(a.seed).shuffle(sizes)  # instance order shuffled by seed
    by_n = {}
    for n in sizes:
        by_n[n] = by_n.get(n, 0) + 1
    # signals / outputs of one edge are contiguous
    sig, out, filt = {}, {}, {}
    for n, cnt in by_n.items():
        P = n ** 3
        sig[n] = torch.empty(cnt * P, dtype=torch.complex64, device="cuda")
        out[n] = torch.empty_like(sig[n])
        filt[n] = torch.empty((1 if shared else cnt) * P, dtype=torch.complex64, device="cuda")
        api.fill_uniform(sig[n], 2 * cnt * P, a.seed + n)
        api.fill_uniform(filt[n], filt[n].numel() * 2, a.seed + 7 * n)
    torch.cuda.synchronize()
    def make_specs():
        # one multi-filter spectrum per edge, one view per instance: the graph
        # batch records consecutive views with consecutive signals as one
        # batched pass sequence
        specs = {}
        for n, cnt in by_n.items():
            if shared:
                specs[n] = [C.Spectrum(n, filt[n])]
            else:
                parent = C.Spectrum(n, filt[n], count=cnt)
                specs[n] = [parent.view(k) for k in range(cnt)]
        return specs

    def instances(specs):
        used = {n: 0 for n in by_n}
        res = []
        for n in sizes:  # in the (shuffled) instance order
            k = used[n]
            used[n] += 1
            P = n ** 3
            res.append((sig[n][k * P:(k + 1) * P], out[n][k * P:(k + 1) * P], specs[n][0 if shared else k]))
        return res

    # graph mode: spectra of each edge's autotuned plan; persistent mode: the
    # plans its kernel runs (PERSISTENT_PLANS)
    inst_g = instances(make_specs())
    for n in by_n:
        api.set_plan(n, persistent_plan(api, n))
    inst = instances(make_specs())
    for n in by_n:
        api.set_plan(n, api.CF_PLAN_AUTO)
    # graph mode records contiguous runs of one spectrum as batched passes: order by edge
    inst_sorted = sorted(inst_g, key=lambda t: (t[2].N, t[0].data_ptr()))
    results = {}

    def timed(launch, sync_check):
        for _ in range(a.warmup):
            launch()
        sync_check()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            launch()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sync_check()
        return (t1 - t0) / a.steps

    outs = {}  # each mode's outputs (cleared before it runs), checked in the cpu_baseline leg

    def run_mode(name, launch, sync_check):
        for n in by_n:
            out[n].zero_()
        results[name] = timed(launch, sync_check)
        if not a.no_cpu_baseline:
            torch.cuda.synchronize()
            outs[name] = {n: out[n].cpu().numpy() for n in by_n}

    bp = api.Batch([(s, o, h) for s, o, h in inst], mode=api.CF_BATCH_PERSISTENT)
    run_mode("persistent_one_launch", bp.launch, bp.check)
    bg = api.Batch([(s, o, h) for s, o, h in inst_sorted], mode=api.CF_BATCH_GRAPH)
    run_mode("graph", bg.launch, bg.check)
    if shared:  # the batch as one sequence of batched passes of the autotuned plan
        sp = C.Spectrum(64, filt[64][:64 ** 3])
        run_mode("batched_passes_" + api.PLAN_NAMES[sp.plan],
                 lambda: C.conv3d_device(sp, sig[64], out[64], how_many=len(sizes)), torch.cuda.synchronize)
    per_size = {}
    if not shared:  # per-edge throughput: each edge group alone, in each mode
        for mode, mname, group in ((api.CF_BATCH_GRAPH, "graph", inst_g),
                                   (api.CF_BATCH_PERSISTENT, "persistent_one_launch", inst)):
            per_size[mname] = {}
            for n, cnt in sorted(by_n.items()):
                b = api.Batch([(s, o, h) for s, o, h in group if h.N == n], mode=mode)
                w = timed(b.launch, b.check)
                per_size[mname]["%d^3 x%d" % (n, cnt)] = {"conv_per_s": cnt / w, "ms": 1e3 * w}
                b.close()
    best = min(results, key=lambda k: results[k])
    wall = results[best]

    def npass(plan):
        return 5 if plan in (api.CF_PLAN_AXIS5, api.CF_PLAN_AXIS5Z) else 3

    # algorithmic bytes of the executed schedule: 2 volumes per pass per cube,
    # plus one spectrum read per distinct spectrum
    if best.startswith("persistent"):
        passes_of = {id(h): npass(h.plan) for _, _, h in inst}
    elif best.startswith("batched"):
        passes_of = {id(h): npass(sp.plan) for _, _, h in inst}
    else:
        inst = inst_g
        passes_of = {id(h): npass(h.plan) for _, _, h in inst}
    alg = sum(2 * passes_of[id(h)] * 8.0 * h.N ** 3 for _, _, h in inst)
    alg += sum(8.0 * h.N ** 3 for h in {id(h): h for _, _, h in inst}.values())
    line = base_line(a, "batched 3D convs/sec (%s)" % a.workload, len(sizes) / wall, "conv/s", 1e3 * wall,
                     {"workload": a.workload, "instances": len(sizes),
                      "edges": {str(n): c for n, c in sorted(by_n.items())},
                      "filters": "one shared" if shared else "own per instance", "mode": best, "scale": "1/N^3"})
    line["modes_ms_per_batch"] = {k: round(1e3 * v, 4) for k, v in results.items()}
    line["roofline"] = {"bound": "hbm", "kernel": "cf_batch_run" if best.startswith("persistent") else best,
                        "achieved": alg / wall / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / wall / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "bytes_model": "algorithmic bytes of the executed schedule for the whole batch"}
    if per_size:
        line["per_size"] = per_size
    bp.close()
    bg.close()
    return line, ("batch", by_n, shared, a.seed, outs)


def run_split(a, torch, C, api):
    """4 x 256^3 + 1 x 512^3 convolutions with their own filters, for U[-1,1)
    and for log-uniform inputs (both signal and filter): the split-FP16 MFMA
    form beside the FP32 form, at both edges (256^3: the four-step passes;
    512^3: the five axis passes of the autotuned plan, natural or z-blocked).
    A step is the whole batch (one call for the four 256^3 cubes, one for the
    512^3 cube), inputs resident in HBM, filter spectra precomputed."""
    n1, b1, n2 = 256, 4, 512
    P1, P2 = n1 ** 3, n2 ** 3
    api.set_plan(n1, api.CF_PLAN_SPLIT3_G24M2)  # spectrum layout read by both forms
    dists = {}
    checks = {}
    for di, dist in enumerate(("uniform", "loguniform")):
        seed = a.seed + 100 * di
        sig1 = torch.empty(b1 * P1, dtype=torch.complex64, device="cuda")
        flt1 = torch.empty_like(sig1)
        sig2 = torch.empty(P2, dtype=torch.complex64, device="cuda")
        flt2 = torch.empty_like(sig2)
        for k, x in enumerate((sig1, flt1, sig2, flt2)):
            if dist == "uniform":
                api.fill_uniform(x, 2 * x.numel(), seed + k)
            else:
                api.fill_loguniform(x, x.numel(), seed + k)
        torch.cuda.synchronize()
        spec1 = C.Spectrum(n1, flt1, count=b1)
        spec2 = C.Spectrum(n2, flt2)  # autotuned plan for 512^3
        out1, out2 = torch.empty_like(sig1), torch.empty_like(sig2)
        entry = {}
        host = {"sig256": sig1.cpu().numpy(), "flt256": flt1.cpu().numpy(),
                "sig512": sig2.cpu().numpy(), "flt512": flt2.cpu().numpy()}
        for prec, pname in ((api.CF_PREC_FP32, "fp32"), (api.CF_PREC_SPLIT_FP16, "split_fp16_mfma")):
            def part1():
                C.conv3d_device(spec1, sig1, out1, how_many=b1, precision=prec)

            def part2():
                C.conv3d_device(spec2, sig2, out2, precision=prec)

            def timed(fn, steps):
                for _ in range(a.warmup):
                    fn()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    fn()
                torch.cuda.synchronize()
                return (time.perf_counter() - t0) / steps

            t_batch = timed(lambda: (part1(), part2()), a.steps)
            t1 = timed(part1, a.steps)
            t2 = timed(part2, max(1, a.steps // 4))
            entry[pname] = {"ms_per_batch": 1e3 * t_batch, "conv_per_s": (b1 + 1) / t_batch,
                            "ms_4x256": 1e3 * t1, "ms_512": 1e3 * t2,
                            "conv_per_s_256": b1 / t1, "ms_per_256_conv": 1e3 * t1 / b1}
            part1()
            part2()
            host["out256_" + pname] = out1.cpu().numpy()
            host["out512_" + pname] = out2.cpu().numpy()
        entry["split_vs_fp32_speed_256"] = entry["fp32"]["ms_4x256"] / entry["split_fp16_mfma"]["ms_4x256"]
        entry["split_vs_fp32_speed_512"] = entry["fp32"]["ms_512"] / entry["split_fp16_mfma"]["ms_512"]
        dists[dist] = entry
        checks[dist] = host
        del spec1, spec2, sig1, flt1, sig2, flt2, out1, out2
        torch.cuda.empty_cache()
    api.set_plan(n1, api.CF_PLAN_AUTO)
    plan512 = api.PLAN_NAMES[api.get_plan(n2)]
    V1, V2 = 8.0 * P1, 8.0 * P2
    alg = b1 * 7 * V1 + 11 * V2  # executed schedules: four-step 256^3 (7V), 5-pass 512^3 (11V)
    u = dists["uniform"]["split_fp16_mfma"]
    line = base_line(a, "split-MFMA vs FP32 FFT convolution, 4x256^3 + 1x512^3 (own filters)", u["conv_per_s"],
                     "conv/s", u["ms_per_batch"],
                     {"workload": "c5_split_4x256^3+1x512^3", "instances": b1 + 1, "filters": "own per instance",
                      "scale": "1/N^3", "plans": {"256": "split3_g24m2", "512": plan512},
                      "split_fp16_sizes": [256, 512]})
    line["dtype"] = "split-fp16 mfma (fp32 accumulate)"
    line["data"] = "synthetic: U[-1,1) and log-uniform magnitudes 1e-3..1e3 with uniform phase, generated on the GPU"
    line["by_distribution"] = {d: {k: ({kk: round(vv, 5) for kk, vv in v.items()} if isinstance(v, dict) else round(v, 4))
                                   for k, v in e.items()} for d, e in dists.items()}
    line["roofline"] = {"bound": "hbm", "kernel": "whole batch (split-FP16 256^3 + split-FP16 512^3)",
                        "achieved": alg / (u["ms_per_batch"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / (u["ms_per_batch"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                        "bytes_model": "4 x 7V (256^3 four-step) + 11V (512^3 five-pass)"}
    return line, checks


def split_check_leg(checks, reps):
    """Checker and CPU baseline of the c5 workload: rel-L2 of every instance
    of both forms against the FP64 oracle on the very inputs the GPU used
    (copied back), and the oracle's FP32 CPU convolution timed at 256^3."""
    import numpy as np
    from oracle import oracle as O
    O.build()
    acc = {}
    for dist, h in checks.items():
        e = {"fp32": [], "split_fp16_mfma": []}
        for inst, (n, sk, fk, b) in enumerate([(256, "sig256", "flt256", k) for k in range(4)] + [(512, "sig512", "flt512", 0)]):
            P = n ** 3
            sig = np.ascontiguousarray(h[sk][b * P:(b + 1) * P])
            flt = np.ascontiguousarray(h[fk][b * P:(b + 1) * P])
            ref = O.conv3d_f64(n, sig, flt, 1.0 / P)
            for form in e:
                got = np.ascontiguousarray(h["out%d_%s" % (n, form)][b * P:(b + 1) * P])
                e[form].append(O.rel_l2(got, ref))
            del ref
        acc[dist] = {k: [float("%.3e" % x) for x in v] for k, v in e.items()}
        acc[dist]["instances"] = ["256^3"] * 4 + ["512^3"]
    base, _ = cpu_baseline_leg(256, None, 1, 2, reps)
    return base, {"rel_l2_vs_fp64": acc}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); --gpus %d is not supported"
              % a.gpus, file=sys.stderr)
        return 2
    import torch
    import convfpga_amd as C
    from convfpga_amd import api

    assert torch.cuda.is_available(), "no GPU"
    torch.zeros(1, device="cuda")
    rc = C.fpga_initialize("gfx950", C.CODE_OBJECT, True)
    if rc != 0:
        raise SystemExit("fpga_initialize -> %d" % rc)
    if a.workload in SPLIT:
        line, checks = run_split(a, torch, C, api)
        line["cpu_baseline"] = None
        if not a.no_cpu_baseline:
            base, check = split_check_leg(checks, a.cpu_reps)
            line["cpu_baseline"] = base
            line.update(check)
    else:
        line, cpu = (run_single if a.workload in SINGLE else run_batch)(a, torch, C, api)
        line["cpu_baseline"] = None
        if not a.no_cpu_baseline:
            if cpu[0] == "batch":
                base, check = batch_cpu_leg(cpu[1], cpu[2], cpu[3], cpu[4], a.cpu_reps)
            else:
                base, check = cpu_baseline_leg(cpu[0], cpu[1], cpu[2], cpu[3], a.cpu_reps,
                                               cpu[4] if len(cpu) > 4 else None)
            line["cpu_baseline"] = base
            line.update(check)
    C.fpga_final()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
