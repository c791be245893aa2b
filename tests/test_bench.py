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
