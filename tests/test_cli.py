"""The `convolution` command-line harness (convfpga_amd/cli/convolution.cpp):
flags, dispatch rules and report of the reference's sample application
(src/main.cpp:11-106, src/helper.cpp:59-110, 180-270)."""
import ctypes.util
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "convfpga_amd", "lib", "convolution")


def _run(*args, timeout=120):
    if not os.path.exists(CLI):
        pytest.skip("convolution not built")
    r = subprocess.run([CLI] + list(args), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout


def _have_fftw():
    try:
        ctypes.CDLL("libfftw3f.so.3")
        return True
    except OSError:
        return False


def test_help():
    rc, out = _run("--help")
    assert rc == 0
    for flag in ("-p, --path", "-n, --num", "-i, --iter", "-b, --batch", "-s, --usesvm", "-y, --noverify",
                 "-c, --cpu-only", "-t, --threads", "-w, --wisdomfile", "-e, --emulate"):
        assert flag in out


def test_bad_values_fail():
    assert _run("-n", "abc")[0] == 1
    assert _run("-i", "0")[0] == 1
    assert _run("extra")[0] == 1
    assert _run("--fftw-plan", "fastest")[0] == 1


def test_fftw_plan_mode_is_reported():
    # FFTW_PLAN of the reference's build (src/CMakeLists.txt:5-6), chosen at run time here
    for mode, label in (("estimate", "Estimate"), ("measure", "Measure"), ("patient", "Patient"),
                        ("exhaustive", "Exhaustive")):
        rc, out = _run("-c", "-n", "8", "--fftw-plan", mode)[:2]
        assert "FFTW Plan    = %s" % label in out


def test_non_svm_batch_is_rejected():
    # src/main.cpp:71-72: batches need the SVM calls
    rc, out = _run("-n", "16", "-b", "2")
    assert rc == 1 and "Non-SVM Batch not implemented" in out


def test_emulation_is_not_available():
    rc, out = _run("-e")
    assert rc == 1 and "emulation" in out


@pytest.mark.skipif(_have_fftw(), reason="FFTW present: the CPU-only mode runs")
def test_cpu_only_without_fftw():
    # src/main.cpp:28-31
    rc, out = _run("-c")
    assert rc == 1 and "FFTW not found" in out


@pytest.mark.gpu
@pytest.mark.parametrize("args,batch", [(("-n", "32", "-i", "3"), 1), (("-s", "-n", "32", "-i", "2"), 1),
                                        (("-s", "-b", "3", "-n", "16", "-i", "2"), 3)])
def test_runs_on_the_gpu(args, batch):
    print("case convolution %s" % " ".join(args), flush=True)
    rc, out = _run(*args)
    assert rc == 0, out
    assert "Measurements in ms" in out and "FFT + Conv" in out
    assert ("Total Comp per batch" in out) == (batch > 1)
    assert ("Deviation of runtimes" in out)
    if "-s" in args:
        assert "PCIe Host to Device" not in out
    else:
        assert "PCIe Host to Device" in out
