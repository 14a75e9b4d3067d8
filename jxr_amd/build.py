"""Builds the gfx950 reconstruction library in-tree: jxr_amd/libjxr_hip.so.

hipcc compiles the kernels for gfx950 only; the host runtime and the
forward analysis are linked into the same shared library. Packed-FP32
instructions are disabled for every HIP compilation.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjxr_hip.so")
SOURCES = ["recon.hip", "runtime.cpp", "analysis.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose: bool = False) -> str:
    """Compiles every source into LIB (always, even when it exists)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [
        HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I", os.path.join(ROOT, "include"),
        *srcs, "-o", tmp,
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = res.stdout + res.stderr
    noise = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if verbose and noise:
        print("\n".join(noise), file=sys.stderr)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(noise))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
