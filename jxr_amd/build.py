"""Builds the gfx950 reconstruction library in-tree: jxr_amd/libjxr_hip.so.

hipcc compiles the kernels for gfx950 only; the host runtime and the
forward analysis are linked into the same shared library. Packed-FP32
instructions are disabled for every HIP compilation.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libjxr_hip.so")
SOURCES = ["recon.hip", "runtime.cpp", "analysis.cpp", "codestream.cpp"]
# The reconstruction kernels are scheduled with LLVM's iterative ILP
# strategy: at the kernels' fixed four waves per SIMD it orders their long
# dependency chains better than the default (about 1% less device time per
# batch for integer and float outputs alike).
SOURCE_FLAGS = {"recon.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose: bool = False) -> str:
    """Compiles every source into LIB (always, even when it exists): the
    sources in parallel, then one link."""
    tmpdir = tempfile.mkdtemp(prefix="jxr_hip_build_")
    flags = [
        HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I", os.path.join(ROOT, "include"),
    ]
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(tmpdir, src + ".o")
        objs.append(obj)
        procs.append(subprocess.Popen([*flags, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    outs = [(p.communicate()[0], p.returncode) for p in procs]
    tmp = LIB + ".tmp"
    link = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp],
                          capture_output=True, text=True)
    outs.append((link.stdout + link.stderr, link.returncode))
    noise = [l for out, _ in outs for l in out.splitlines() if "packed-fp32-ops" not in l]
    if verbose and noise:
        print("\n".join(noise), file=sys.stderr)
    if any(rc != 0 for _, rc in outs):
        shutil.rmtree(tmpdir, ignore_errors=True)
        raise RuntimeError("hipcc failed:\n" + "\n".join(noise))
    shutil.rmtree(tmpdir, ignore_errors=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
