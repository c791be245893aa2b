"""Benchmark: 256^3 complex-FP32 FFT convolution on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2_256]

A step is one convolution of one device-resident 256^3 complex64 signal with
a filter whose spectrum was computed once before timing (the 256^3 single-cube
configuration of the metric): forward x and y passes, the fused z pass
(z forward, product with the spectrum, z inverse), inverse y and x passes
with the 1/N^3 scale.  Inputs are synthetic (seeded U[-1,1) re/im, generated
in HBM by the library).  Prints one JSON line.

Per-pass device times come from HIP events recorded by the library on the
launch stream during the timed steps; the roofline entry is for the pass with
the largest time.  The cpu_baseline leg times the oracle's FP32 CPU
convolution (OpenMP, all host threads) on a bounded sample of the same
workload and, on the same inputs, checks the GPU result against it and
against the FP64 oracle.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X datasheet HBM3E bandwidth

WORKLOADS = {
    # name: (N, batch, shared spectrum)
    "c2_256": (256, 1, True),
    "c2_128": (128, 1, True),
    "c1_64": (64, 1, True),
    "c3_64x64": (64, 64, True),
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2_256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--plan", type=int, default=0, help="force a schedule (CF_PLAN_*); 0 = autotuned")
    return ap.parse_args()


def cpu_baseline_leg(n, batch, gpu_out_host, seed_sig, seed_filt, reps):
    """Times the oracle's FP32 CPU convolution (port of src/fft_conv3D.cpp:50-177)
    on `reps` convolutions of one cube, and checks the GPU output on the same
    inputs.  The only place bench.py uses the oracle."""
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
    threads = O.max_threads()
    res = {
        "value": 1.0 / med, "unit": "conv/s", "cores": threads, "kind": "port",
        "sample": "%d reps of one %d^3 FP32 convolution (FFT, product, inverse FFT; filter spectrum "
                  "precomputed), median; oracle's own radix-2 FFT with OpenMP (FFTW3 %s)"
                  % (reps, n, "present" if O.fftw_available() else "not installed"),
    }
    check = {}
    if gpu_out_host is not None:
        first = np.ascontiguousarray(gpu_out_host[:P])
        ref32 = out * np.complex64(1.0 / P)  # the GPU path applies 1/N^3; exact power-of-two scale
        check["snr_db_vs_fp32_cpu"] = O.snr_db(np.ascontiguousarray(ref32), first)
        ref64 = O.conv3d_f64(n, sig, filt, 1.0 / P)
        check["rel_l2_vs_fp64"] = O.rel_l2(first, ref64)
    return res, check


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


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); --gpus %d is not supported"
              % a.gpus, file=sys.stderr)
        return 2
    import torch
    import convfpga_amd as C
    from convfpga_amd import api

    n, batch, shared = WORKLOADS[a.workload]
    assert torch.cuda.is_available(), "no GPU"
    torch.zeros(1, device="cuda")
    rc = C.fpga_initialize("gfx950", C.CODE_OBJECT, False)
    if rc != 0:
        raise SystemExit("fpga_initialize -> %d" % rc)
    P = n ** 3
    seed_sig, seed_filt = 1, 2
    sig = torch.empty(batch * P, dtype=torch.complex64, device="cuda")
    filt = torch.empty(P, dtype=torch.complex64, device="cuda")
    api.fill_uniform(sig, 2 * P * batch, seed_sig)  # one stream over the batch: every cube differs
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
    copy_gbs = copy_bandwidth(torch, max(batch * P, 1 << 24))

    for _ in range(a.warmup):
        C.conv3d_device(spec, sig, out, how_many=batch)
    torch.cuda.synchronize()
    api.profile_begin(a.steps)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        C.conv3d_device(spec, sig, out, how_many=batch)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    passes, ncalls = api.profile_end()
    assert ncalls == a.steps

    wall = t1 - t0
    ms_per_step = 1e3 * wall / a.steps
    value = a.steps * batch / wall
    V = 8.0 * P
    alg_bytes = sum(p["bytes"] for p in passes)  # per step
    for p in passes:
        p["GBps"] = p["bytes"] / (p["ms"] * 1e-3) / 1e9
        p["frac_of_peak"] = p["GBps"] / HBM_PEAK_GBS
    dom = max(passes, key=lambda p: p["ms"])
    dev_ms = sum(p["ms"] for p in passes)

    line = {
        "metric": "3D convs/sec + % HBM roofline at 256^3 c2c",
        "value": value,
        "unit": "conv/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded U[-1,1) re/im generated on the GPU)",
        "config": {"workload": "conv3d_%d^3_c2c_fp32_batch%d_precomputed_filter_spectrum" % (n, batch),
                   "N": n, "batch": batch, "plan": api.PLAN_NAMES[plan], "passes": len(passes), "scale": "1/N^3",
                   "algorithmic_bytes_per_step": alg_bytes},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": dom["GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["GBps"] / HBM_PEAK_GBS, "traffic": None},
        "conv_roofline_frac": alg_bytes * a.steps / wall / 1e9 / HBM_PEAK_GBS,
        # the same rate priced with the 5-pass model (11 volume transfers per
        # convolution with its own spectrum, 2*5 per batched cube + 1)
        "conv_roofline_frac_5pass_model": (2 * 5 * batch + 1) * 8.0 * P * a.steps / wall / 1e9 / HBM_PEAK_GBS,
        "device_ms_per_step": dev_ms,
        "copy_kernel_GBps": copy_gbs,
        "autotune_ms_per_conv": {api.PLAN_NAMES[p]: round(v, 5) for p, v in tune_ms.items()},
        "passes": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in p.items()} for p in passes],
    }
    if not a.no_cpu_baseline:
        host_out = out[:P].cpu().numpy()
        base, check = cpu_baseline_leg(n, batch, host_out, seed_sig, seed_filt, a.cpu_reps)
        line["cpu_baseline"] = base
        line.update(check)
    else:
        line["cpu_baseline"] = None
    C.fpga_final()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
