"""bench.py's output contract: one JSON line with the metric, the roofline
and cpu_baseline objects and finite parity checks of the timed output."""
import json
import math
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + list(args), cwd=ROOT,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.gpu
def test_single_cube_line():
    print("case bench c1_64", flush=True)
    r = _bench("--workload", "c1_64", "--steps", "5", "--warmup", "2", "--cpu-reps", "2")
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["value"] > 0
    ro = d["roofline"]
    assert ro["bound"] == "hbm" and 0 < ro["frac"] < 1.5 and ro["unit"] == "GB/s"
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1
    assert math.isfinite(d["rel_l2_vs_fp64"]) and d["rel_l2_vs_fp64"] <= 1e-5
    assert d["snr_db_vs_fp32_cpu"] > 120


@pytest.mark.gpu
def test_batch_line_checks_every_instance():
    print("case bench c3_64x64", flush=True)
    r = _bench("--workload", "c3_64x64", "--steps", "3", "--warmup", "1", "--cpu-reps", "2")
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["value"] > 0 and d["roofline"]["bound"] == "hbm" and d["cpu_baseline"]["value"] > 0
    assert d["instances_checked"] == 64
    worst = d["rel_l2_vs_fp64_max_by_mode"]
    assert set(worst) == set(d["modes_ms_per_batch"])
    assert all(math.isfinite(v) and v <= 1e-5 for v in worst.values()), worst


@pytest.mark.gpu
def test_c4_mixed_line_checks_all_29_instances():
    """The mixed-size workload at its stated composition (16x32^3 + 8x64^3 +
    4x128^3 + 1x256^3, own filters, shuffled): every instance of every batch
    mode within 1e-5 of the FP64 oracle."""
    print("case bench c4_mixed", flush=True)
    r = _bench("--workload", "c4_mixed", "--steps", "3", "--warmup", "1", "--cpu-reps", "1")
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["config"]["edges"] == {"32": 16, "64": 8, "128": 4, "256": 1}
    assert d["config"]["instances"] == 29 and d["instances_checked"] == 29
    worst = d["rel_l2_vs_fp64_max_by_mode"]
    assert set(worst) == set(d["modes_ms_per_batch"]) and len(worst) >= 2
    assert all(math.isfinite(v) and v <= 1e-5 for v in worst.values()), worst
    assert d["value"] > 0 and d["roofline"]["bound"] == "hbm"


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_c5_split_line_checks_every_instance():
    """4x256^3 + 1x512^3 with own filters, for U[-1,1) and log-uniform
    inputs: every instance of both forms (FP32 and split-FP16 MFMA) within the
    stated bound of 1e-5 against the FP64 oracle, 512^3 included."""
    print("case bench c5_split", flush=True)
    r = _bench("--workload", "c5_split", "--steps", "2", "--warmup", "1", "--cpu-reps", "1", timeout=800)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    acc = d["rel_l2_vs_fp64"]
    assert set(acc) == {"uniform", "loguniform"}
    for dist, e in acc.items():
        assert e["instances"][:4] == ["256^3"] * 4 and e["instances"][4].startswith("512^3")
        for form in ("fp32", "split_fp16_mfma"):
            errs = e[form]
            assert len(errs) == 5, (dist, form)
            assert all(math.isfinite(v) and 0 < v <= 1e-5 for v in errs), (dist, form, errs)
    assert d["value"] > 0 and d["cpu_baseline"]["value"] > 0


def test_pass_kernel_names_of_every_four_step_plan():
    """The bench names each pass's kernel (for the per-kernel HBM bytes in
    profiles/traffic_256.json): the names of every four-step plan must be
    kernels of the built code object, and every 256^3 one must have its bytes."""
    sys.path.insert(0, ROOT)
    import bench
    from convfpga_amd import api
    blob = open(api.CODE_OBJECT, "rb").read()
    traffic = json.load(open(os.path.join(ROOT, "profiles", "traffic_256.json")))
    plans = [name for name in api.PLAN_NAMES.values() if name.startswith("split3_")]
    assert len(plans) == 16
    for n in (128, 256):
        for plan in plans:
            for p in ("xy1_fwd", "y2z_fwd_mul_inv", "xy1_inv"):
                k = bench.kernel_of_pass(p, n, plan)
                assert (k + "\0").encode() in blob, (plan, p, k)
                if n == 256:
                    assert k in traffic, (plan, p, k)
    assert bench.kernel_of_pass("xy1_inv", 256, "split3_g28m2oc4") == "cf_xy1_inv_256_g42k"
    assert bench.kernel_of_pass("y2z_fwd_mul_inv", 256, "split3_g28m2oc4") == "cf_mid_256_g28m2ok"
    # one-y2 first pass; the middle and last passes of G28M2OC4
    assert [bench.kernel_of_pass(p, 256, "split3_g18m2oc4") for p in ("xy1_fwd", "y2z_fwd_mul_inv", "xy1_inv")] == \
        ["cf_xy1_fwd_256_g18", "cf_mid_256_g28m2ok", "cf_xy1_inv_256_g42k"]
