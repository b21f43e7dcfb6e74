"""In-tree build of the native parts.

* libddyn_hip.so  HIP kernels + C ABI (include/ddyn.h): ddyn_hip.hip (state,
  steps, adjoint), ddyn_image.hip (images and SEDs) and ddyn_mie.hip (Mie
  tables), hipcc for gfx950 with
  -O3, C++17, -ffp-contract=off (every fused multiply-add in the kernels is
  written out explicitly, so rounding is fixed) and the packed-FP32 target
  feature switched off.
* libddyn_host.so .param reader, time-step rule, data-file writer/reader,
  optical data, dust grid, grain tables and particle generation
  (include/ddyn_host.h), host C++ only.
* ddyn            command-line driver linking both.
* runmie          the Mie-table tool (refractive indices in, composition table
                  out), linking libddyn_hip.so.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
LIB = os.path.join(HERE, "libddyn_hip.so")
HOST_LIB = os.path.join(HERE, "libddyn_host.so")
CLI = os.path.join(HERE, "ddyn")
MIE_CLI = os.path.join(HERE, "runmie")
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INC]
# -ffp-contract=off: the host library's products and sums stay apart, as in
# the reference's x86-64 host code, so generated particles match it bit for bit
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-I" + INC]


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build_hip(verbose=False):
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + [os.path.join(CSRC, "ddyn_hip.hip"), os.path.join(CSRC, "ddyn_image.hip"),
                                os.path.join(CSRC, "ddyn_mie.hip"), "-o", tmp], "hipcc")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_host(verbose=False):
    tmp = HOST_LIB + ".tmp"
    _run([CXX] + HOST_FLAGS + ["-shared", os.path.join(CSRC, "ddyn_host.cpp"), os.path.join(CSRC, "ddyn_setup.cpp"),
                               "-o", tmp], "host library")
    os.replace(tmp, HOST_LIB)
    tmp = CLI + ".tmp"
    _run([CXX] + HOST_FLAGS + [os.path.join(CSRC, "ddyn_main.cpp"), "-o", tmp, "-L" + HERE,
                               "-lddyn_hip", "-lddyn_host", "-Wl,-rpath,$ORIGIN"], "ddyn driver")
    os.replace(tmp, CLI)
    tmp = MIE_CLI + ".tmp"
    _run([CXX] + HOST_FLAGS + [os.path.join(CSRC, "ddyn_runmie.cpp"), "-o", tmp, "-L" + HERE, "-lddyn_hip",
                               "-Wl,-rpath,$ORIGIN"], "runmie tool")
    os.replace(tmp, MIE_CLI)
    if verbose:
        print("built", HOST_LIB, ",", CLI, "and", MIE_CLI)
    return HOST_LIB


COPY_CALIB_SRC = os.path.join(ROOT, "copy_calib.hip")
COPY_CALIB = os.path.join(ROOT, "copy_calib")


def build_copy_calib(verbose=False):
    """The copy kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters
    (measurement only, not part of the library)."""
    tmp = COPY_CALIB + ".tmp"
    flags = [f for f in HIP_FLAGS if f not in ("-fPIC", "-shared")]
    _run([HIPCC] + flags + [COPY_CALIB_SRC, "-o", tmp], "hipcc copy_calib")
    os.replace(tmp, COPY_CALIB)
    if verbose:
        print("built", COPY_CALIB)


def build_all(verbose=False):
    build_hip(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
