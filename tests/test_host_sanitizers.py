"""The host library under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only; the kernels are not involved): a sanitized copy of
libconvfpga.so is built in a temporary directory and driven through every
C-ABI path that runs without a GPU (tests/sanitize/host_checks.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(os.path.join(ROCM, "lib", "libamdhip64.so")),
                    reason="needs g++ and the HIP runtime library")
def test_host_paths_clean_under_asan_ubsan(tmp_path):
    csrc = os.path.join(ROOT, "convfpga_amd", "csrc")
    inc = os.path.join(ROOT, "include")
    lib = tmp_path / "libconvfpga.so"
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]
    subprocess.run(["g++", "-O1", "-g", "-fPIC", "-shared", "-std=c++17", "-fvisibility=hidden", "-w"] + san +
                   ["-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-I" + inc, "-I" + csrc,
                    os.path.join(csrc, "engine.cpp"), os.path.join(csrc, "capi.cpp"),
                    "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
                    "-o", str(lib)], check=True)
    exe = tmp_path / "host_checks"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17"] + san +
                   ["-I" + inc, os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"), "-L" + str(tmp_path),
                    "-lconvfpga", "-Wl,-rpath," + str(tmp_path), "-o", str(exe)], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    assert "ok (0 failures)" in r.stdout
