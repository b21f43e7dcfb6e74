"""In-tree build of libddyn_hip.so for gfx950 with hipcc.

Flags: -O3, C++17, --offload-arch=gfx950, -ffp-contract=off (every fused
multiply-add in the kernels is written out explicitly, so rounding is fixed),
and the packed-FP32 target feature switched off.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = [os.path.join(HERE, "csrc", "ddyn_hip.hip")]
LIB = os.path.join(HERE, "libddyn_hip.so")
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def build_hip(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + SOURCES + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build_hip(verbose=True)
