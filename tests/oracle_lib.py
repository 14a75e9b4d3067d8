"""ctypes loader of the CPU oracle (oracle/libjxr_oracle.so), the checker of
the HIP path. Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

from jxr_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "libjxr_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    L = C.CDLL(ORACLE)
    i32p = C.POINTER(C.c_int32)
    for name, n in [("jxro_inverse_core_transform", 16), ("jxro_forward_core_transform", 16),
                    ("jxro_inverse_chroma_420", 4), ("jxro_forward_chroma_420", 4),
                    ("jxro_inverse_chroma_422", 8), ("jxro_forward_chroma_422", 8),
                    ("jxro_inverse_overlap_4", 4), ("jxro_forward_overlap_4", 4),
                    ("jxro_inverse_overlap_4x4", 16), ("jxro_forward_overlap_4x4", 16),
                    ("jxro_inverse_overlap_2x2", 4), ("jxro_forward_overlap_2x2", 4),
                    ("jxro_inverse_overlap_2", 2), ("jxro_forward_overlap_2", 2)]:
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = [i32p]
    L.jxro_upsample_pair.restype = C.c_int
    L.jxro_upsample_pair.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32, i32p]
    L.jxro_predict_high_pass.restype = C.c_int
    L.jxro_predict_high_pass.argtypes = [i32p, C.c_uint32, C.c_uint32, C.c_uint32]
    L.jxro_premultiply_unsigned.restype = C.c_int
    L.jxro_premultiply_unsigned.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
    L.jxro_premultiply_signed.restype = C.c_int
    L.jxro_premultiply_signed.argtypes = [C.c_int32, C.c_int32, C.c_int32, i32p]
    L.jxro_float32_bits.restype = C.c_int
    L.jxro_float32_bits.argtypes = [C.c_int32, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_uint32)]
    L.jxro_float16_bits.restype = C.c_int
    L.jxro_float16_bits.argtypes = [C.c_int32, C.c_uint32, C.POINTER(C.c_uint16)]
    u32p = C.POINTER(C.c_uint32)
    for name in ("jxro_overlap_plane", "jxro_subsampled_overlap_plane"):
        fn = getattr(L, name)
        fn.restype = C.c_int
        fn.argtypes = [i32p, C.c_uint32, C.c_uint32, u32p, C.c_uint32, u32p, C.c_uint32, C.c_int32]
    L.jxro_convert.restype = C.c_int
    L.jxro_convert.argtypes = [C.c_uint32, C.c_uint32, i32p, C.c_int32, C.c_int32, i32p]
    L.jxro_decode.restype = C.c_int
    L.jxro_decode.argtypes = [C.POINTER(abi.ImageDesc), C.c_void_p, C.c_uint64, u32p, C.c_char_p, C.c_size_t]
    L.jxro_reconstruct_plane.restype = C.c_int
    L.jxro_reconstruct_plane.argtypes = [C.POINTER(abi.ImageDesc), C.c_uint32, i32p, C.c_uint64, u32p, u32p,
                                         u32p, u32p, u32p]
    L.jxro_probe_axis.restype = C.c_int
    L.jxro_probe_axis.argtypes = [u32p, C.c_uint32, u32p, u32p, u32p, C.c_uint32]
    L.jxro_analyze_plane.restype = C.c_int
    L.jxro_analyze_plane.argtypes = [i32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32,
                                     u32p, C.c_uint32, u32p, C.c_uint32, i32p]
    L.jxro_decode_many.restype = C.c_int
    L.jxro_decode_many.argtypes = [C.POINTER(C.POINTER(abi.ImageDesc)), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_uint64), C.c_uint32, C.c_uint32]
    _lib = L
    return L


def op(name, values):
    """Runs one operator on a copy of `values`; returns (result, overflowed)."""
    v = np.ascontiguousarray(np.array(values, dtype=np.int32))
    err = getattr(lib(), name)(v.ctypes.data_as(C.POINTER(C.c_int32)))
    return v, bool(err)


class OracleError(RuntimeError):
    def __init__(self, code, stage, message):
        super().__init__(f"oracle error {code} stage {stage}: {message}")
        self.code, self.stage, self.message = code, stage, message


def decode(desc: abi.ImageDesc) -> np.ndarray:
    out = np.zeros(desc.surface.byte_len, np.uint8)
    stage = C.c_uint32()
    msg = C.create_string_buffer(256)
    code = lib().jxro_decode(C.byref(desc), out.ctypes.data, out.nbytes, C.byref(stage), msg, 256)
    if code:
        raise OracleError(code, stage.value, msg.value.decode())
    return out


def decode_status(desc: abi.ImageDesc):
    """(code, stage, bytes or None) without raising."""
    try:
        return 0, 0, decode(desc)
    except OracleError as e:
        return e.code, e.stage, None


def reconstruct_plane(desc: abi.ImageDesc, index: int):
    cap = 1 << 26
    buf = np.zeros(cap, np.int32)
    w, h, ox, oy, st = (C.c_uint32() for _ in range(5))
    code = lib().jxro_reconstruct_plane(C.byref(desc), index, buf.ctypes.data_as(C.POINTER(C.c_int32)), cap,
                                        C.byref(w), C.byref(h), C.byref(ox), C.byref(oy), C.byref(st))
    if code:
        raise OracleError(code, st.value, "reconstruct_plane")
    return buf[: w.value * h.value].reshape(h.value, w.value).copy()


def decode_many(descs, threads, outs=None):
    """Decodes every descriptor on `threads` host threads, into `outs` when
    given (reused buffers: no page faults inside a timed pass)."""
    n = len(descs)
    if outs is None:
        outs = [np.zeros(d.surface.byte_len, np.uint8) for d in descs]
    darr = (C.POINTER(abi.ImageDesc) * n)(*[C.pointer(d) for d in descs])
    oarr = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
    larr = (C.c_uint64 * n)(*[o.nbytes for o in outs])
    code = lib().jxro_decode_many(darr, oarr, larr, n, threads)
    if code:
        raise OracleError(code, 0, "decode_many")
    return outs


def convert(internal, output, values, packed_rgb=False, red_blue_not_swapped=True):
    """One pixel's colour conversion; returns (components, code)."""
    inp = np.zeros(4, np.int32)
    inp[: len(values)] = values
    out = np.zeros(4, np.int32)
    code = lib().jxro_convert(internal, output, inp.ctypes.data_as(C.POINTER(C.c_int32)), 1 if packed_rgb else 0,
                              1 if red_blue_not_swapped else 0, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.tolist(), code


def probe_axis(regions):
    """(bounds, positions) of the sample-scale probe axis of regions with
    these macroblock boundaries, or None when the partition is refused."""
    r = np.ascontiguousarray(np.array(regions, np.uint32))
    bounds = np.zeros(16, np.uint32)
    nb = C.c_uint32()
    cap = max(int(r[-1]) if r.size else 0, 1)
    pos = np.zeros(cap, np.uint32)
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    if lib().jxro_probe_axis(u32(r), r.size, u32(bounds), C.byref(nb), u32(pos), cap):
        return None
    total = int(r[-1] - r[0])
    return bounds[: nb.value].tolist(), pos[:total].tolist()


def analyze_plane(samples, geometry, overlap, column_bounds, row_bounds, doubled=False):
    """Forward analysis at unit steps: (mby, mbx, 256) coefficients."""
    bc, br = geometry
    s = np.ascontiguousarray(np.asarray(samples, np.int32))
    mby, mbx = s.shape[0] // (br * 4), s.shape[1] // (bc * 4)
    out = np.zeros((mby, mbx, 256), np.int32)
    cb = np.ascontiguousarray(np.array(column_bounds, np.uint32))
    rb = np.ascontiguousarray(np.array(row_bounds, np.uint32))
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    if lib().jxro_analyze_plane(i32(s), mbx, mby, bc, br, 1 if doubled else 0, overlap, u32(cb), cb.size, u32(rb),
                                rb.size, i32(out)):
        raise OracleError(1, 0, "analyze_plane")
    return out
