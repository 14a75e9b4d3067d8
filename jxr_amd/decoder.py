"""Host-side API over the C ABI, named after the reference's CUDA backend.

  HipDecodePlan      CudaDecodePlan      (crates/jxr-cuda/src/plan.rs:17-148)
  HipDecoderSession  CudaDecoderSession  (crates/jxr-cuda/src/session.rs:16-268)
  HipSubmission      CudaSubmission      (crates/jxr-cuda/src/submission.rs:11-118)
  ResidentHipImage   ResidentCudaImage   (crates/jxr-cuda/src/resident.rs:8-58)
  plan_hip_route     plan_cuda_route     (crates/jxr-cuda/src/route.rs:23-50)

Errors raise JxrHipError with the C status code; validation errors happen
before any device resource exists, and an arithmetic failure after
submission carries the first failing stage (no CPU retry).
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import abi
from .abi import JxrHipError, check, load


# ---------------------------------------------------------------------------
# encoded images (forward analysis into arenas)
# ---------------------------------------------------------------------------
def quality_parameters(quality: int, color: int, rgb_source: bool, sample_range: int, is_float: bool = False,
                       lowest_quality: int = 0):
    """The reference encoder's quality scale (crates/jxr-native/src/encode/
    quality.rs:101-499): ((DC, LP, HP) QPs of luma, (DC, LP, HP) QPs of
    components 1 and 2, overlap mode the quality picks) for a source whose
    coded samples span `sample_range` values for the 8-bit range (255 for
    8-bit, 65535 for 16-bit; sample_scale() gives it for any format).
    Components 1 and 2 of an RGB source are colour differences at the chroma
    sampling of `color`. QP 0 is lossless."""
    luma, chroma, ov = (C.c_uint32 * 3)(), (C.c_uint32 * 3)(), C.c_uint32()
    check(load().jxrhip_quality_parameters_scaled(int(quality), int(color), 1 if rgb_source else 0,
                                                  int(sample_range), 1 if is_float else 0, int(lowest_quality),
                                                  luma, chroma, C.byref(ov)), "quality parameters")
    return tuple(luma), tuple(chroma), int(ov.value)


def sample_scale(depth: int, output_color: int = abi.COLOR_RGB, shift_bits: int = 0, mantissa_length: int = 13,
                 exponent_bias: int = 4, alpha: bool = False):
    """SampleScale (crates/jxr-native/src/encode/quality.rs:194-292) of a
    source stored at `depth`: (range, is_float, lowest_quality) of its colour
    planes, or of its alpha plane."""
    r, f, lq = C.c_uint32(), C.c_uint32(), C.c_uint32()
    check(load().jxrhip_sample_scale(int(depth), int(output_color), int(shift_bits), int(mantissa_length),
                                     int(exponent_bias), 1 if alpha else 0, C.byref(r), C.byref(f), C.byref(lq)),
          "sample scale")
    return int(r.value), bool(f.value), int(lq.value)


class EncodedImage:
    """Arenas produced by jxrhip_encode for one seeded image."""

    def __init__(self, pixels: np.ndarray, *, color: int, overlap: int, tiles=None,
                 hard_tiles: bool = False, scaled: bool = False, qp=(0, 0, 0),
                 bands: int = abi.BANDS_ALL, alpha_mode: int = abi.ALPHA_NONE,
                 shift_bits: int = 0, mantissa_length: int = 13, exponent_bias: int = 4,
                 alpha_qp: int = 0, seed: int = 0, vary_quantizers: bool = False, chroma_qp=None):
        lib = load()
        pixels = np.ascontiguousarray(pixels)
        if pixels.ndim == 2:
            pixels = pixels[:, :, None]
        h, w, ch = pixels.shape
        kind = {np.dtype(np.uint8): abi.INPUT_U8, np.dtype(np.uint16): abi.INPUT_U16,
                np.dtype(np.float32): abi.INPUT_F32}[pixels.dtype]
        p = abi.EncodeParams()
        p.width, p.height, p.input_kind, p.input_channels = w, h, kind, ch
        p.color_format, p.alpha_mode, p.overlap = color, alpha_mode, overlap
        p.hard_tiles = 1 if hard_tiles else 0
        self._tiles = None
        if tiles is not None:
            cols, rows = tiles
            self._tiles = (np.asarray(cols, np.uint32), np.asarray(rows, np.uint32))
            p.tile_column_count = len(cols)
            p.tile_column_widths = self._tiles[0].ctypes.data_as(abi.u32p)
            p.tile_row_count = len(rows)
            p.tile_row_heights = self._tiles[1].ctypes.data_as(abi.u32p)
        p.scaled = 1 if scaled else 0
        p.qp_dc, p.qp_lp, p.qp_hp = qp
        p.bands = bands
        p.shift_bits, p.mantissa_length, p.exponent_bias = shift_bits, mantissa_length, exponent_bias
        p.alpha_qp, p.seed, p.vary_quantizers = alpha_qp, seed, 1 if vary_quantizers else 0
        if chroma_qp is not None:
            p.chroma_qp_set = 1
            p.qp_chroma_dc, p.qp_chroma_lp, p.qp_chroma_hp = chroma_qp
        handle = C.c_void_p()
        check(lib.jxrhip_encode(pixels.ctypes.data, pixels.nbytes, C.byref(p), C.byref(handle)), "encode")
        self._handle = handle
        self.width, self.height, self.channels = w, h, ch
        self.color_format = color
        self.component_count = 1 if color == abi.COLOR_LUMA else 3
        self.alpha_mode = alpha_mode if ch in (2, 4) else abi.ALPHA_NONE
        self.scaled = bool(scaled)
        self.input_dtype = pixels.dtype
        self.shift_bits, self.mantissa_length, self.exponent_bias = shift_bits, mantissa_length, exponent_bias
        self.overlap = overlap

    @classmethod
    def at_quality(cls, pixels: np.ndarray, quality: int, *, color: int, overlap=None, **kw):
        """Analysis at a quality of the reference encoder's scale: its
        parameters for the source's sample range, scaled arithmetic unless
        every component is lossless, and the quality's overlap mode unless
        `overlap` is given (crates/jxr-native/src/encode/quality.rs:101-150)."""
        dtype = np.asarray(pixels).dtype
        if dtype == np.float32:
            # float sources: the coded bit patterns' scale at the request's
            # precision (quality.rs:221-230)
            scale = sample_scale(abi.DEPTH_F32, mantissa_length=kw.get("mantissa_length", 13),
                                 exponent_bias=kw.get("exponent_bias", 4))
        else:
            scale = ({np.dtype(np.uint8): 255, np.dtype(np.uint16): 65535}[dtype], False, 0)
        rgb = pixels.ndim == 3 and pixels.shape[2] >= 3
        luma, chroma, ov = quality_parameters(quality, color, rgb, *scale)
        return cls(pixels, color=color, overlap=ov if overlap is None else overlap,
                   scaled=any(luma) or any(chroma), qp=luma, chroma_qp=chroma, **kw)

    @property
    def primary(self):
        return load().jxrhip_encoded_arena(self._handle, 0)

    @property
    def alpha(self):
        ptr = load().jxrhip_encoded_arena(self._handle, 1)
        return ptr if ptr else None

    def coefficient_bytes(self) -> int:
        n = self.primary.contents.coefficient_count * 4
        if self.alpha is not None:
            n += self.alpha.contents.coefficient_count * 4
        return n

    def close(self):
        if getattr(self, "_handle", None):
            load().jxrhip_encoded_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# output policy + surface layout
# ---------------------------------------------------------------------------
# bytes of one stored element: a sample, a packed pixel word, or a byte of
# eight one-bit samples
_STORE_BYTES = {abi.STORE_BITS: 1, abi.STORE_U8: 1, abi.STORE_U16: 2, abi.STORE_I16: 2, abi.STORE_I32: 4,
                abi.STORE_F16: 2, abi.STORE_F32: 4, abi.STORE_PACKED16: 2, abi.STORE_PACKED32: 4}
_LAYOUT_CHANNELS = {abi.LAYOUT_LUMA: 1, abi.LAYOUT_LUMA_ALPHA: 2, abi.LAYOUT_RGB: 3,
                    abi.LAYOUT_BGR: 3, abi.LAYOUT_RGBA: 4, abi.LAYOUT_BGRA: 4,
                    abi.LAYOUT_RGBX: 4, abi.LAYOUT_BGRX: 4, abi.LAYOUT_CMYK: 4, abi.LAYOUT_CMYKA: 5,
                    abi.LAYOUT_YUV: 3, abi.LAYOUT_YUVA: 4}


@dataclass
class OutputSpec:
    """OutputFormatRequest + SurfaceLayout choice for one image."""
    storage: int = abi.STORE_U8
    depth: int = abi.DEPTH_U8
    layout: int = abi.LAYOUT_RGB
    output_color: int = abi.COLOR_RGB
    planar: bool = False
    shift_bits: int = 0
    mantissa_length: int = 13
    exponent_bias: int = 4
    premultiply: bool = False
    red_blue_not_swapped: bool = False
    centering: tuple = (0, 0)
    crop: Optional[tuple] = None  # (x, y, w, h) in the coded plane
    row_padding: int = 0          # extra bytes per row (strided surfaces)
    components: int = 3           # colour channels of an N-component layout
    scale: int = abi.SCALE_FULL   # DecodeScale: full, quarter or sixteenth

    @property
    def bytes_per_sample(self):
        return _STORE_BYTES[self.storage]

    @property
    def channels(self):
        if self.layout == abi.LAYOUT_NCOMPONENT:
            return self.components
        if self.layout == abi.LAYOUT_NCOMPONENT_ALPHA:
            return self.components + 1
        return _LAYOUT_CHANNELS[self.layout]


def surface_layout(width, height, spec: OutputSpec, base: int = 0) -> abi.SurfaceLayout:
    """Tightly packed surface of a request (SurfaceLayout::for_output,
    crates/jxr-core/src/surface.rs:29-140): one interleaved plane, one plane
    per channel (NCHW), or native planar YUV with chroma planes at the coded
    sampling; packed formats store one word per pixel and one-bit luma a
    byte per eight pixels."""
    s = abi.SurfaceLayout()
    s.width, s.height = width, height
    b = spec.bytes_per_sample
    ch = spec.channels
    if spec.layout in (abi.LAYOUT_YUV, abi.LAYOUT_YUVA):
        s.planar = 1
        sub_y = 2 if spec.output_color == abi.COLOR_YUV420 else 1
        dims = [(width, height), (width // 2, height // sub_y), (width // 2, height // sub_y)]
        if spec.layout == abi.LAYOUT_YUVA:
            dims.append((width, height))
        s.plane_count = len(dims)
        off = base
        for c, (w, h) in enumerate(dims):
            pl = s.planes[c]
            pl.byte_offset = off
            pl.row_stride_bytes = w * b + spec.row_padding
            pl.width, pl.height, pl.channels = w, h, 1
            off += pl.row_stride_bytes * h
        s.byte_len = off
        return s
    if spec.storage in (abi.STORE_BITS, abi.STORE_PACKED16, abi.STORE_PACKED32):
        s.planar = 0
        s.plane_count = 1
        pl = s.planes[0]
        pl.byte_offset = base
        row = (width + 7) // 8 if spec.storage == abi.STORE_BITS else width * b
        pl.row_stride_bytes = row + spec.row_padding
        pl.width, pl.height, pl.channels = width, height, ch
        s.byte_len = base + pl.row_stride_bytes * height
        return s
    if spec.planar:
        s.planar = 1
        s.plane_count = ch
        off = base
        for c in range(ch):
            pl = s.planes[c]
            pl.byte_offset = off
            pl.row_stride_bytes = width * b + spec.row_padding
            pl.width, pl.height, pl.channels = width, height, 1
            off += pl.row_stride_bytes * height
        s.byte_len = off
    else:
        s.planar = 0
        s.plane_count = 1
        pl = s.planes[0]
        pl.byte_offset = base
        pl.row_stride_bytes = width * ch * b + spec.row_padding
        pl.width, pl.height, pl.channels = width, height, ch
        s.byte_len = base + pl.row_stride_bytes * height
    return s


def image_desc(enc: EncodedImage, spec: OutputSpec) -> abi.ImageDesc:
    d = abi.ImageDesc()
    d.abi_version = abi.ABI_VERSION
    d.color_format = enc.color_format
    d.component_count = enc.component_count
    d.alpha_mode = enc.alpha_mode
    d.scaled = 1 if enc.scaled else 0
    d.primary = enc.primary
    if enc.alpha_mode == abi.ALPHA_SEPARATE:
        d.separate_alpha = enc.alpha
    o = d.output
    yuv = enc.color_format in (abi.COLOR_YUV420, abi.COLOR_YUV422, abi.COLOR_YUV444)
    native = spec.output_color in (abi.COLOR_YUV420, abi.COLOR_YUV422)  # planar YUV keeps the sampling
    o.internal_color = enc.color_format if native else (abi.COLOR_YUV444 if yuv else enc.color_format)
    o.output_color = spec.output_color
    o.bit_depth = spec.depth
    o.shift_bits = spec.shift_bits
    o.mantissa_length = spec.mantissa_length
    o.exponent_bias = spec.exponent_bias
    o.has_alpha_format = 1 if enc.alpha_mode != abi.ALPHA_NONE else 0
    o.alpha_bit_depth = spec.depth
    o.alpha_shift_bits = spec.shift_bits
    o.alpha_mantissa_length = spec.mantissa_length
    o.alpha_exponent_bias = spec.exponent_bias
    o.alpha_scaled = 1 if enc.scaled else 0
    o.storage = spec.storage
    o.channel_layout = spec.layout
    o.channel_count = spec.channels
    o.scaled = 1 if enc.scaled else 0
    o.red_blue_not_swapped = 1 if spec.red_blue_not_swapped else 0
    o.premultiply_alpha = 1 if spec.premultiply else 0
    o.chroma_centering_x, o.chroma_centering_y = spec.centering
    cx, cy, cw, ch = reduced_crop(spec.crop if spec.crop else (0, 0, enc.width, enc.height), spec.scale)
    o.crop_x, o.crop_y, o.crop_width, o.crop_height = cx, cy, cw, ch
    d.scale = spec.scale
    d.surface = surface_layout(cw, ch, spec)
    return d


def reduced_crop(region, scale: int):
    """The output crop of a full-resolution region (x, y, w, h) at a decode
    scale: its origin divided and its far edges rounded up by 4 or 16
    (native_output_region, crates/jxr-native/src/prepare.rs:196-215; output_crop,
    crates/jxr-native/src/decode/output.rs:484-501)."""
    den = {abi.SCALE_FULL: 1, abi.SCALE_QUARTER: 4, abi.SCALE_SIXTEENTH: 16}[scale]
    x, y, w, h = region
    x0, y0 = x // den, y // den
    return x0, y0, -(-(x + w) // den) - x0, -(-(y + h) // den) - y0


# ---------------------------------------------------------------------------
# plans, sessions, submissions
# ---------------------------------------------------------------------------
class HipDecodePlan:
    """Validated reconstruction plan (CudaDecodePlan::from_prepared)."""

    def __init__(self, desc: abi.ImageDesc, keepalive=None):
        lib = load()
        self._desc = desc
        self._keep = keepalive
        h = C.c_void_p()
        check(lib.jxrhip_plan_create(C.byref(desc), C.byref(h)), "plan")
        self._h = h

    @classmethod
    def from_encoded(cls, enc: EncodedImage, spec: OutputSpec):
        return cls(image_desc(enc, spec), keepalive=enc)

    @classmethod
    def from_codestream(cls, codestream, spec: Optional["OutputSpec"] = None, decoded=None, threads: int = 0):
        """The plan of a parsed codestream (jxr_amd.codestream.Codestream):
        its tiles entropy-decoded on host threads unless `decoded` already
        holds them, and the file's own output format unless `spec` names
        another (PreparedReconstruction::cuda_plan, crates/jxr/src/
        reconstruction.rs:137-158). A reduced spec.scale reads only the bands
        that scale keeps; a spec.crop is a region of the output image."""
        spec = spec if spec is not None else codestream.default_output()
        if decoded is None:
            tiles = codestream.region_tiles(spec.crop) if spec.crop else None
            decoded = codestream.decode(scale=spec.scale, tiles=tiles, threads=threads)
        return cls(codestream.image_desc(decoded, spec), keepalive=(codestream, decoded))

    @property
    def handle(self):
        return self._h

    @property
    def desc(self):
        return self._desc

    @property
    def output_bytes(self) -> int:
        return load().jxrhip_plan_output_bytes(self._h)

    @property
    def bucket_key(self) -> int:
        """Shape bucket: plans with equal keys share a persistent launch."""
        return load().jxrhip_plan_bucket_key(self._h)

    @property
    def output_pixels(self) -> int:
        return load().jxrhip_plan_output_pixels(self._h)

    @property
    def reconstructed_coefficients(self) -> int:
        return load().jxrhip_plan_reconstructed_coefficients(self._h)

    @property
    def device_footprint(self) -> int:
        return load().jxrhip_plan_device_footprint(self._h)

    def decode_report(self, requested: int = abi.BACKEND_HIP, host_readback: bool = False) -> dict:
        """Route and per-stage ownership of a device decode
        (CudaDecodePlan::decode_report, crates/jxr-cuda/src/plan.rs:214-253)."""
        r = abi.DecodeReport()
        check(load().jxrhip_plan_decode_report(self.handle, requested, 1 if host_readback else 0, C.byref(r)),
              "decode_report")
        return _report(r)

    @property
    def algorithmic_bytes(self) -> int:
        return load().jxrhip_plan_algorithmic_bytes(self._h)

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_hip_route(reconstructed_coefficients: int, strict: bool, available: bool) -> bool:
    use = C.c_int32()
    check(load().jxrhip_route(reconstructed_coefficients, 1 if strict else 0, 1 if available else 0, C.byref(use)),
          "route")
    return bool(use.value)


def _report(r: abi.DecodeReport) -> dict:
    backends = {abi.BACKEND_AUTO: "Auto", abi.BACKEND_CPU: "Cpu", abi.BACKEND_HIP: "Hip"}
    return {"requested": backends.get(r.requested, r.requested), "selected": backends.get(r.selected, r.selected),
            "fallback": None if r.fallback == 0 else r.fallback,
            "stages": [(abi.DECODE_STAGES[x.stage], abi.EXECUTORS[x.executor]) for x in r.stages[: r.stage_count]]}


class ResidentHipImage:
    def __init__(self, session, handle, session_handle=None):
        self._s = session
        self._h = handle
        # the image keeps the native session alive, also after the session
        # object was closed
        self._sh = session_handle if session_handle is not None else session.handle

    @property
    def device_ptr(self) -> int:
        return load().jxrhip_image_device_ptr(self._h) or 0

    @property
    def nbytes(self) -> int:
        return load().jxrhip_image_bytes(self._h)

    def readback(self) -> np.ndarray:
        out = np.empty(self.nbytes, np.uint8)
        check(load().jxrhip_readback(self._sh, self._h, out.ctypes.data, out.nbytes), "readback")
        return out

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_image_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipSubmission:
    def __init__(self, session, handle):
        self._s = session
        self._h = handle
        self._sh = session.handle

    def is_complete(self) -> bool:
        return bool(load().jxrhip_submission_is_complete(self._h))

    def enqueue_consumer_wait(self, hip_stream: int):
        """Work queued next on `hip_stream` runs after this reconstruction
        (CudaSubmission::enqueue_consumer_wait)."""
        check(load().jxrhip_submission_enqueue_wait(self._h, C.c_void_p(hip_stream)), "enqueue_wait")

    def output_layout(self) -> abi.SurfaceLayout:
        out = abi.SurfaceLayout()
        check(load().jxrhip_submission_output_layout(self._h, C.byref(out)), "output_layout")
        return out

    def pending_device_ptr(self):
        """(pointer, bytes) of the output a consumer stream may read after
        enqueue_consumer_wait (CudaSubmission::pending_device_buffer)."""
        p, n = C.c_void_p(), C.c_uint64()
        check(load().jxrhip_submission_pending_device_ptr(self._h, C.byref(p), C.byref(n)), "pending_device_ptr")
        return p.value or 0, n.value

    def wait(self) -> Optional[ResidentHipImage]:
        lib = load()
        img = C.c_void_p()
        stage = C.c_uint32()
        h, self._h = self._h, None
        code = lib.jxrhip_submission_wait(h, C.byref(img), C.byref(stage))
        if code != abi.OK:
            raise JxrHipError(code, lib.jxrhip_last_error().decode(errors="replace"), stage.value)
        return ResidentHipImage(self._s, img, self._sh) if img.value else None

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                load().jxrhip_submission_destroy(self._h)
            except Exception:
                pass


class HipBatch:
    def __init__(self, session, handle, count):
        self._s = session
        self._h = handle
        self.count = count

    def is_complete(self) -> bool:
        return bool(load().jxrhip_batch_is_complete(self._h))

    def enqueue_consumer_wait(self, hip_stream: int):
        """Work queued next on `hip_stream` runs after every image of the batch."""
        check(load().jxrhip_batch_enqueue_wait(self._h, C.c_void_p(hip_stream)), "batch_enqueue_wait")

    def wait(self, raise_on_error=True) -> np.ndarray:
        """Returns each image's stage code (0 = ok)."""
        lib = load()
        stages = np.zeros(max(self.count, 1), np.uint32)
        h, self._h = self._h, None
        code = lib.jxrhip_batch_wait(h, stages.ctypes.data_as(abi.u32p))
        if code not in (abi.OK, abi.ERR_KERNEL_ARITHMETIC) or (code != abi.OK and raise_on_error):
            raise JxrHipError(code, lib.jxrhip_last_error().decode(errors="replace"))
        return stages[: self.count]

    def __del__(self):
        # a batch dropped unwaited is waited for, then freed (as a dropped
        # pending submission is in the reference, submission.rs:132-139)
        if getattr(self, "_h", None):
            try:
                self.wait(raise_on_error=False)
            except Exception:  # noqa: BLE001
                pass


class HipBatchGraph:
    """A captured persistent batch. launch() replays it and returns a
    HipBatch; one replay at a time. Keeps its plans alive."""

    def __init__(self, session, plans, device_ptrs):
        self._s = session
        self._plans = list(plans)
        n = len(plans)
        parr = (C.c_void_p * max(n, 1))(*[p.handle.value for p in plans])
        darr = (C.c_void_p * max(n, 1))(*[int(d) for d in device_ptrs])
        h = C.c_void_p()
        check(load().jxrhip_batch_graph_create(session._h, parr, n, darr, C.byref(h)), "batch_graph_create")
        self._h = h
        self.count = n

    def launch(self) -> HipBatch:
        b = C.c_void_p()
        check(load().jxrhip_batch_graph_launch(self._h, C.byref(b)), "batch_graph_launch")
        return HipBatch(self._s, b, self.count)

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_batch_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class HostBuffer:
    """Page-locked host memory of a session (jxrhip_host_alloc)."""

    def __init__(self, session, nbytes: int):
        self._s = session
        self._sh = session.handle
        p = C.c_void_p()
        check(load().jxrhip_host_alloc(session.handle, nbytes, C.byref(p)), "host_alloc")
        self._p = p
        self.nbytes = nbytes

    @property
    def ptr(self) -> int:
        return self._p.value or 0

    def view(self, dtype=np.uint8) -> np.ndarray:
        buf = (C.c_uint8 * self.nbytes).from_address(self.ptr)
        return np.frombuffer(buf, dtype=dtype)

    def close(self):
        if getattr(self, "_p", None):
            load().jxrhip_host_free(self._sh, self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory of a session (CudaDecoderSession::allocate_destination)."""

    def __init__(self, session, nbytes: int):
        self._s = session
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().jxrhip_allocate_destination(session.handle, self.nbytes, C.byref(p)), "allocate")
        self._p = p
        # the destination keeps the native session alive until it is freed,
        # also after the session object was closed
        self._sh = session.handle

    @property
    def ptr(self) -> int:
        return self._p.value or 0

    def fill(self, value: int):
        check(load().jxrhip_fill_destination(self._sh, self._p, value & 0xFF, self.nbytes), "fill")

    def to_host(self, offset: int = 0, nbytes: Optional[int] = None) -> np.ndarray:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        check(load().jxrhip_copy_to_host(self._sh, C.c_void_p(self.ptr + offset), out.ctypes.data, n),
              "copy_to_host")
        return out

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            load().jxrhip_free_destination(self._sh, self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipDecoderSession:
    """One device, its streams, upload cache and scratch pool."""

    def __init__(self, device_ordinal: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.jxrhip_session_create(device_ordinal, C.byref(h)), "session")
        self._h = h

    @staticmethod
    def is_available() -> bool:
        return bool(load().jxrhip_is_available())

    @classmethod
    def system_default(cls):
        return cls(0)

    @property
    def handle(self):
        return self._h

    def diagnostics(self) -> abi.Diagnostics:
        d = abi.Diagnostics()
        check(load().jxrhip_session_diagnostics(self._h, C.byref(d)), "diagnostics")
        return d

    @property
    def stream_count(self) -> int:
        return self.diagnostics().stream_count

    @property
    def device_ordinal(self) -> int:
        return self.diagnostics().device_ordinal

    def set_tile(self, tx: int, ty: int):
        check(load().jxrhip_session_set_tile(self._h, tx, ty), "set_tile")

    def set_occupancy(self, groups_per_cu: int):
        check(load().jxrhip_session_set_occupancy(self._h, groups_per_cu), "set_occupancy")

    def set_upload_budget(self, nbytes: int):
        check(load().jxrhip_session_set_upload_budget(self._h, nbytes), "set_upload_budget")

    def upload(self, plan: HipDecodePlan):
        check(load().jxrhip_session_upload(self._h, plan.handle), "upload")

    def clear_cache(self):
        check(load().jxrhip_session_clear_cache(self._h), "clear_cache")

    def submit(self, plan: HipDecodePlan) -> HipSubmission:
        h = C.c_void_p()
        check(load().jxrhip_submit(self._h, plan.handle, C.byref(h)), "submit")
        return HipSubmission(self, h)

    def submit_into(self, plan: HipDecodePlan, device_ptr: int, nbytes: int) -> HipSubmission:
        h = C.c_void_p()
        check(load().jxrhip_submit_into(self._h, plan.handle, C.c_void_p(device_ptr), nbytes, C.byref(h)),
              "submit_into")
        return HipSubmission(self, h)

    def decode_to_host(self, plan: HipDecodePlan) -> np.ndarray:
        lib = load()
        out = np.empty(plan.output_bytes, np.uint8)
        stage = C.c_uint32()
        code = lib.jxrhip_decode_to_host(self._h, plan.handle, out.ctypes.data, out.nbytes, C.byref(stage))
        if code != abi.OK:
            raise JxrHipError(code, lib.jxrhip_last_error().decode(errors="replace"), stage.value)
        return out

    def submit_batch(self, plans: Sequence[HipDecodePlan], device_ptrs: Sequence[int],
                     mode: int = abi.BATCH_PERSISTENT) -> HipBatch:
        n = len(plans)
        parr = (C.c_void_p * max(n, 1))(*[p.handle.value for p in plans])
        darr = (C.c_void_p * max(n, 1))(*[int(d) for d in device_ptrs])
        h = C.c_void_p()
        check(load().jxrhip_submit_batch(self._h, parr, n, darr, mode, C.byref(h)), "submit_batch")
        return HipBatch(self, h, n)

    def autotune(self, plans: Sequence[HipDecodePlan], device_ptrs: Sequence[int], repeats: int = 3):
        """Times the candidate tile shapes of every shape bucket of the batch,
        then fewer workgroups per CU at the best shape, and keeps the fastest
        per bucket (jxrhip_session_autotune). Returns one dict per bucket."""
        n = len(plans)
        parr = (C.c_void_p * max(n, 1))(*[p.handle.value for p in plans])
        darr = (C.c_void_p * max(n, 1))(*[int(d) for d in device_ptrs])
        cap = max(n, 1)
        out = (abi.TileChoice * cap)()
        nb = C.c_uint32()
        check(load().jxrhip_session_autotune(self._h, parr, n, darr, repeats, out, cap, C.byref(nb)), "autotune")
        return [{"bucket_key": c.bucket_key, "tile": (c.tile_mbs_x, c.tile_mbs_y), "images": c.images,
                 "candidates": c.candidates, "best_ms": c.best_ms, "default_ms": c.default_ms,
                 "groups_per_cu": c.groups_per_cu}
                for c in out[: min(nb.value, cap)]]

    def calibrate_valu(self, iterations: int = 20000):
        """Issue rate of the vector ALU (jxrhip_calibrate_valu): dict of ms,
        wave_instructions, simds and clock_khz of one timed launch."""
        r = abi.ValuRate()
        check(load().jxrhip_calibrate_valu(self._h, iterations, C.byref(r)), "calibrate_valu")
        return {"ms": r.ms, "wave_instructions": r.wave_instructions, "simds": r.simds, "clock_khz": r.clock_khz}

    def set_bucket_tile(self, bucket_key: int, tx: int, ty: int):
        check(load().jxrhip_session_set_bucket_tile(self._h, bucket_key, tx, ty), "set_bucket_tile")

    def bucket_tile(self, bucket_key: int):
        tx, ty = C.c_uint32(), C.c_uint32()
        check(load().jxrhip_session_bucket_tile(self._h, bucket_key, C.byref(tx), C.byref(ty)), "bucket_tile")
        return tx.value, ty.value

    def capture_batch(self, plans: Sequence[HipDecodePlan], device_ptrs: Sequence[int]) -> "HipBatchGraph":
        """A persistent batch captured into a HIP graph for repeated replay
        (jxrhip_batch_graph_create)."""
        return HipBatchGraph(self, plans, device_ptrs)

    def allocate_destination(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def allocate_host(self, nbytes: int) -> "HostBuffer":
        """Page-locked host memory (jxrhip_host_alloc): staging the session
        uploads from by DMA."""
        return HostBuffer(self, nbytes)

    def synchronize(self):
        check(load().jxrhip_session_synchronize(self._h), "synchronize")

    def copy_device(self, dst: int, src: int, nbytes: int):
        check(load().jxrhip_copy_device(self._h, C.c_void_p(dst), C.c_void_p(src), nbytes), "copy_device")

    def last_batch_timing(self):
        ms = C.c_double()
        launches = C.c_uint32()
        check(load().jxrhip_last_batch_timing(self._h, C.byref(ms), C.byref(launches)), "timing")
        return ms.value, launches.value

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
