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
                 "-c, --cpu-only", "-t, --threads", "-w, --wisdomfile", "-e, --emulate", "--dump-out"):
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
@pytest.mark.parametrize("args,batch", [(("-n", "32"), 1), (("-s", "-n", "32"), 1), (("-s", "-b", "3", "-n", "32"), 3)])
def test_cli_output_matches_oracle(args, batch, oracle, tmp_path):
    """The numbers the harness computes, for the plain, the -s and the -s -b 3
    dispatches (src/main.cpp:67-74): its inputs come from the reference's
    generator (rand()/RAND_MAX in [0, 1], every instance a copy of the first,
    src/helper.cpp:148-172; drawn after fpga_initialize as in main.cpp:42-62,
    so where glibc's stream stands then is the runtime's business), and every
    output cube equals the unnormalised convolution N^3 * (sig (*) filter) of
    the FP64 oracle on exactly those inputs."""
    import numpy as np
    print("case convolution %s --dump-out" % " ".join(args), flush=True)
    dump = str(tmp_path / "dump.bin")
    rc, out = _run(*(args + ("-y", "--dump-out", dump)))
    assert rc == 0, out
    n = 32
    P = n ** 3
    data = np.fromfile(dump, dtype=np.complex64)
    assert data.size == P + 2 * batch * P
    filt, sig, got = data[:P], data[P:P + batch * P], data[P + batch * P:]
    for x in (filt, sig):
        f = x.view(np.float32)
        assert f.min() >= 0.0 and f.max() <= 1.0 and 0.45 < f.mean() < 0.55
    assert not np.array_equal(filt, sig[:P])  # two different draws
    for b in range(batch):
        assert np.array_equal(sig[b * P:(b + 1) * P], sig[:P])  # copies of the first (helper.cpp:166-171)
        ref = oracle.conv3d_f64(n, np.ascontiguousarray(sig[b * P:(b + 1) * P]), filt)
        assert oracle.rel_l2(np.ascontiguousarray(got[b * P:(b + 1) * P]), ref) <= 1e-5, b


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
