"""CPU build of the HIP path (test infrastructure): g++ compiles
jxr_amd/csrc/{recon.hip, runtime.cpp, analysis.cpp, codestream.cpp} against the host
emulation in tests/emu, optionally with AddressSanitizer. The resulting
library exports the same C ABI, so the parity cases run through the real
kernel and runtime logic on the CPU."""
import ctypes as C
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
OUT = os.path.join(EMU, "build")


def build(asan: bool) -> str:
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libjxr_hip_emu_asan.so" if asan else "libjxr_hip_emu.so")
    csrc = os.path.join(ROOT, "jxr_amd", "csrc")
    srcs = [os.path.join(csrc, "recon.hip"), os.path.join(csrc, "runtime.cpp"),
            os.path.join(csrc, "analysis.cpp"), os.path.join(csrc, "codestream.cpp"),
            os.path.join(EMU, "emu_runtime.cpp")]
    newest = max(os.path.getmtime(s) for s in srcs + [os.path.join(EMU, "hip", "hip_runtime.h"),
                                                     os.path.join(csrc, "recon_abi.h"),
                                                     os.path.join(csrc, "encoded.h")])
    if os.path.exists(lib) and os.path.getmtime(lib) >= newest:
        return lib
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-fno-omit-frame-pointer",
           "-Wno-unused-function", "-I", EMU, "-I", os.path.join(ROOT, "include"), "-x", "c++", *srcs,
           "-o", lib + ".tmp", "-lpthread"]
    if asan:
        cmd.insert(1, "-fsanitize=address")
    subprocess.run(cmd, check=True)
    os.replace(lib + ".tmp", lib)
    return lib


def load(asan: bool):
    """Loads the emulated library with the production signatures and makes
    it the library jxr_amd's classes call."""
    from jxr_amd import abi
    path = build(asan)
    lib = C.CDLL(path)
    for name, (res, args) in abi.SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    abi._lib = lib
    return lib


def asan_runtime() -> str:
    """LD_PRELOAD value for the sanitized library: the ASan runtime, and the
    C++ runtime beside it so that ASan's exception interceptor finds the real
    __cxa_throw (the codestream reader reports malformed input by throwing
    inside the library)."""
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    cxx = subprocess.run(["g++", "-print-file-name=libstdc++.so.6"], capture_output=True, text=True,
                         check=True).stdout.strip()
    return f"{asan} {cxx}"
