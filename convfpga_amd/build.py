"""Build libconvfpga for gfx950, in-tree.

Three artefacts under convfpga_amd/lib/:
  convfpga_gfx950.hsaco  the kernel code object (hipcc --genco), loaded by
                         fpga_initialize(path) -- the analogue of the
                         reference's .aocx bitstream;
  libconvfpga.so         the C-ABI host library (g++, HIP runtime API only),
                         exporting include/convfpga/convfpga.h and
                         include/convfpga/convfpga_ext.h;
  convolution            the command-line harness (cli/convolution.cpp), a
                         plain C++ client of libconvfpga.so.
Run as `python -m convfpga_amd.build` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HSACO = os.path.join(LIBDIR, "convfpga_%s.hsaco" % ARCH)
SHLIB = os.path.join(LIBDIR, "libconvfpga.so")
CLI = os.path.join(LIBDIR, "convolution")
CLI_SOURCES = [os.path.join(PKG, "cli", "convolution.cpp")]

KERNEL_SOURCES = [os.path.join(CSRC, "kernels", "conv3d_kernels.hip")]
HOST_SOURCES = [os.path.join(CSRC, "engine.cpp"), os.path.join(CSRC, "capi.cpp")]

# Packed FP32 instructions are disabled for all device code on this target.
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def build(verbose=False, extra_hip_flags=(), hsaco=None):
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    target = hsaco or HSACO
    tmp = target + ".tmp"
    _run([hipcc, "--genco"] + HIP_FLAGS + list(extra_hip_flags) + ["-I" + os.path.join(CSRC, "kernels")]
         + KERNEL_SOURCES + ["-o", tmp])
    os.replace(tmp, target)
    cxx = os.environ.get("CXX", "g++")
    tmp = SHLIB + ".tmp"
    _run([cxx, "-O2", "-fPIC", "-shared", "-std=c++17", "-fopenmp", "-fvisibility=hidden", "-Wall", "-Wextra",
          "-Wno-unused-parameter",
          "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-I" + INCLUDE, "-I" + CSRC]
         + HOST_SOURCES + ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
                           "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", tmp])
    os.replace(tmp, SHLIB)
    tmp = CLI + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-fopenmp", "-Wall", "-Wextra", "-I" + INCLUDE] + CLI_SOURCES
         + ["-L" + LIBDIR, "-lconvfpga", "-ldl", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, CLI)
    if verbose:
        print("built", HSACO, "and", SHLIB)
    return SHLIB, HSACO


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
