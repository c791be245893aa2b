"""The persistent batch's overlap rule (convfpga_amd/csrc/batch_spans.h, used
by cf_batch_create): its O(n log n) sweep against a brute-force check over
every pair of instances on random cases (tests/native/batch_spans_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_overlap_sweep_matches_brute_force(tmp_path):
    exe = tmp_path / "batch_spans_check"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-Wextra", "-fsanitize=address,undefined",
                    "-I" + os.path.join(ROOT, "convfpga_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "batch_spans_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert r.stdout.startswith("ok"), r.stdout
