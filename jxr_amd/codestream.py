"""Codestreams: the CPU half of a decode and the last step of an encode.

  Codestream.parse       parse_codestream        (crates/jxr-native/src/codestream.rs:27-63)
  Codestream.info        image_info              (crates/jxr-native/src/metadata.rs:14-40)
  Codestream.decode      decode_tiles            (crates/jxr-native/src/tile_decode/mod.rs:31-60)
  encode_codestream      encode_codestream       (crates/jxr-native/src/encode.rs)
  HipDecodePlan.from_codestream
                         JxrDecoder::decode's plan step, PreparedReconstruction::cuda_plan
                                                 (crates/jxr/src/reconstruction.rs:137-158)

All of it runs in libjxr_hip.so (codestream.cpp): headers, tile index,
Annex-A container, adaptive-VLC entropy decoding into coefficient arenas
(tiles on host threads) and the writer's mirror image. Errors raise
JxrHipError with JXRHIP_ERR_SYNTAX, _UNSUPPORTED or _OVERFLOW.
"""
import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import abi
from .abi import check, load


def encode_codestream(enc, *, frequency: bool = False, trim: Optional[int] = None, annex_a: bool = False,
                      profile_idc: int = 111, level_idc: int = 255) -> bytes:
    """The codestream of an EncodedImage's arenas: spatial (one packet per
    tile) or frequency (one per band of each tile) packets, flexbits trimmed
    by `trim` bits when given, inside an Annex-A container when `annex_a`
    (which a separate alpha plane needs)."""
    o = abi.CodestreamOptions()
    o.frequency_mode = 1 if frequency else 0
    o.trim_flexbits = abi.TRIM_NONE if trim is None else int(trim)
    o.annex_a = 1 if annex_a else 0
    o.profile_idc, o.level_idc = profile_idc, level_idc
    buf = C.POINTER(C.c_uint8)()
    n = C.c_uint64()
    check(load().jxrhip_encoded_codestream(enc._handle, C.byref(o), C.byref(buf), C.byref(n)), "encode codestream")
    try:
        return C.string_at(buf, n.value)
    finally:
        load().jxrhip_free_bytes(C.cast(buf, C.c_void_p))


class DecodedArenas:
    """Coefficient arenas entropy-decoded from a codestream (the primary one,
    with an integrated alpha plane after its components, and a separate
    alpha codestream's)."""

    def __init__(self, handle, codestream):
        self._h = handle
        self._cs = codestream  # the plan's image facts come from it

    # each arena pointer keeps this object, which owns what it points to
    @property
    def primary(self):
        ptr = load().jxrhip_decoded_arena(self._h, 0)
        ptr.owner = self
        return ptr

    @property
    def alpha(self):
        ptr = load().jxrhip_decoded_arena(self._h, 1)
        if not ptr:
            return None
        ptr.owner = self
        return ptr

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_decoded_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Codestream:
    """A parsed JPEG XR file (raw T.832 or Annex-A): headers, tile grid,
    index and container, ready for entropy decoding."""

    def __init__(self, data: bytes):
        data = bytes(data)
        h = C.c_void_p()
        check(load().jxrhip_codestream_parse(data, len(data), C.byref(h)), "parse")
        self._h = h
        info = abi.CodestreamInfo()
        check(load().jxrhip_codestream_get_info(h, C.byref(info)), "info")
        self.info = info
        cols = (C.c_uint32 * info.tile_column_count)()
        rows = (C.c_uint32 * info.tile_row_count)()
        check(load().jxrhip_codestream_tile_grid(h, cols, rows), "tile grid")
        self.tile_columns = list(cols)
        self.tile_rows = list(rows)

    @classmethod
    def parse(cls, data: bytes) -> "Codestream":
        return cls(data)

    @property
    def width(self) -> int:
        return self.info.width

    @property
    def height(self) -> int:
        return self.info.height

    @property
    def tile_count(self) -> int:
        return len(self.tile_columns) * len(self.tile_rows)

    def _mask(self, tiles):
        if tiles is None:
            return None
        if len(tiles) != self.tile_count:
            raise ValueError("one flag per tile")
        return (C.c_uint8 * len(tiles))(*[1 if t else 0 for t in tiles])

    def coefficient_count(self, scale: int = abi.SCALE_FULL, tiles: Optional[Sequence[bool]] = None):
        """(primary, separate alpha) coefficient counts of the arenas a
        decode with these arguments produces."""
        p, a = C.c_uint64(), C.c_uint64()
        check(load().jxrhip_codestream_coefficient_count(self._h, scale, self._mask(tiles), C.byref(p), C.byref(a)),
              "coefficient count")
        return p.value, a.value

    def decode(self, scale: int = abi.SCALE_FULL, tiles: Optional[Sequence[bool]] = None,
               threads: int = 0, into=None) -> DecodedArenas:
        """Entropy-decodes every tile, or those `tiles` (one flag per tile in
        raster order) selects, on up to `threads` host threads (0: one per
        core, at most sixteen). `into`, a HostBuffer (or any object with
        `ptr` and `nbytes`) of at least coefficient_count() values, takes the
        coefficients: pinned memory makes the upload a DMA."""
        mask = self._mask(tiles)
        h = C.c_void_p()
        if into is None:
            check(load().jxrhip_codestream_decode(self._h, scale, mask, threads, C.byref(h)), "entropy decode")
            return DecodedArenas(h, self)
        np_, na = self.coefficient_count(scale, tiles)
        if (np_ + na) * 4 > into.nbytes:
            raise ValueError("coefficient storage below the arenas' size")
        check(load().jxrhip_codestream_decode_into(self._h, scale, mask, threads, C.c_void_p(into.ptr), np_,
                                                   C.c_void_p(into.ptr + np_ * 4) if na else None, na, C.byref(h)),
              "entropy decode")
        d = DecodedArenas(h, self)
        d._storage = into
        return d

    def region_tiles(self, region) -> list:
        """Tiles a region (x, y, w, h) of the output needs, with the one
        macroblock halo that overlap filtering or chroma upsampling reads
        (prepare.rs reconstruction_bounds)."""
        x, y, w, h = region
        i = self.info
        halo = 1 if (i.overlap != 0 or i.color_format in (abi.COLOR_YUV420, abi.COLOR_YUV422)) else 0
        x0 = max(0, (i.margin_left + x) // 16 - halo)
        y0 = max(0, (i.margin_top + y) // 16 - halo)
        x1 = min(i.macroblocks_x, -(-(i.margin_left + x + w) // 16) + halo)
        y1 = min(i.macroblocks_y, -(-(i.margin_top + y + h) // 16) + halo)
        flags = []
        ty = 0
        for th in self.tile_rows:
            tx = 0
            for tw in self.tile_columns:
                flags.append(tx < x1 and tx + tw > x0 and ty < y1 and ty + th > y0)
                tx += tw
            ty += th
        return flags

    def default_output(self, planar: bool = False):
        """The output a decode of this file produces by default: its header's
        output colour and bit depth, all channels, alpha kept."""
        from .decoder import OutputSpec
        i = self.info
        depth = i.bit_depth
        storage = {abi.DEPTH_U8: abi.STORE_U8, abi.DEPTH_U16: abi.STORE_U16, abi.DEPTH_I16: abi.STORE_I16,
                   abi.DEPTH_F16: abi.STORE_F16, abi.DEPTH_I32: abi.STORE_I32, abi.DEPTH_F32: abi.STORE_F32}[depth]
        alpha = i.alpha_mode != abi.ALPHA_NONE
        color = i.output_color
        if color == abi.COLOR_LUMA:
            layout = abi.LAYOUT_LUMA_ALPHA if alpha else abi.LAYOUT_LUMA
        elif color == abi.COLOR_RGB:
            layout = abi.LAYOUT_RGBA if alpha else abi.LAYOUT_RGB
        elif color == abi.COLOR_NCOMPONENT:
            layout = abi.LAYOUT_NCOMPONENT_ALPHA if alpha else abi.LAYOUT_NCOMPONENT
        elif color in (abi.COLOR_CMYK, abi.COLOR_CMYK_DIRECT):
            layout = abi.LAYOUT_CMYKA if alpha else abi.LAYOUT_CMYK
        else:
            raise NotImplementedError(f"default output for colour {color}")
        return OutputSpec(storage=storage, depth=depth, layout=layout, output_color=color, planar=planar,
                          shift_bits=i.shift_bits, mantissa_length=i.mantissa_length,
                          exponent_bias=i.exponent_bias, centering=(i.chroma_centering_x, i.chroma_centering_y),
                          red_blue_not_swapped=bool(i.red_blue_not_swapped),
                          components=i.component_count)

    def image_desc(self, decoded: DecodedArenas, spec) -> abi.ImageDesc:
        """The reconstruction request for `decoded` (CudaDecodePlan::from_prepared's
        arguments): the file's facts plus the requested output; a crop in
        spec.crop is a region of the output image."""
        from .decoder import reduced_crop, surface_layout
        i = self.info
        d = abi.ImageDesc()
        d.abi_version = abi.ABI_VERSION
        d.color_format = i.color_format
        d.component_count = i.component_count
        d.alpha_mode = i.alpha_mode
        d.scaled = i.scaled
        d.primary = decoded.primary
        if i.alpha_mode == abi.ALPHA_SEPARATE:
            d.separate_alpha = decoded.alpha
        o = d.output
        yuv = i.color_format in (abi.COLOR_YUV420, abi.COLOR_YUV422, abi.COLOR_YUV444)
        native = spec.output_color in (abi.COLOR_YUV420, abi.COLOR_YUV422)
        o.internal_color = i.color_format if native else (abi.COLOR_YUV444 if yuv else i.color_format)
        o.output_color = spec.output_color
        o.bit_depth = spec.depth
        o.shift_bits = spec.shift_bits
        o.mantissa_length = spec.mantissa_length
        o.exponent_bias = spec.exponent_bias
        o.has_alpha_format = 1 if i.alpha_mode != abi.ALPHA_NONE else 0
        o.alpha_bit_depth = i.alpha_bit_depth if i.alpha_mode == abi.ALPHA_SEPARATE else spec.depth
        o.alpha_shift_bits = i.alpha_shift_bits
        o.alpha_mantissa_length = i.alpha_mantissa_length
        o.alpha_exponent_bias = i.alpha_exponent_bias
        o.alpha_scaled = i.alpha_scaled
        o.storage = spec.storage
        o.channel_layout = spec.layout
        o.channel_count = spec.channels
        o.scaled = i.scaled
        o.red_blue_not_swapped = 1 if spec.red_blue_not_swapped else 0
        o.premultiply_alpha = 1 if spec.premultiply else 0
        o.chroma_centering_x, o.chroma_centering_y = spec.centering
        region = spec.crop if spec.crop else (0, 0, i.width, i.height)
        x, y, w, h = region
        cx, cy, cw, ch = reduced_crop((x + i.margin_left, y + i.margin_top, w, h), spec.scale)
        o.crop_x, o.crop_y, o.crop_width, o.crop_height = cx, cy, cw, ch
        d.scale = spec.scale
        d.surface = surface_layout(cw, ch, spec)
        d.keepalive = (self, decoded)  # the arenas the descriptor points into
        return d

    def close(self):
        if getattr(self, "_h", None):
            load().jxrhip_codestream_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
