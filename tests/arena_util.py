"""Hand-built coefficient arenas for known-answer tests.

A DC-only macroblock with coefficient 16*v (all steps 1, no overlap)
reconstructs to the constant v, so a plane of per-macroblock values stands in
for the reconstructed planes the reference's output tests format."""
import ctypes as C
import itertools

import numpy as np

from jxr_amd import OutputSpec, abi, surface_layout

_ids = itertools.count(1 << 40)


class ManualArena:
    def __init__(self, planes, overlap=0, dc_scale=16, geometry=None, hard_tiles=False, tiles=None):
        """planes: list of 2-D int arrays (mby, mbx) of per-macroblock values."""
        mby, mbx = planes[0].shape
        n = len(planes)
        geometry = geometry or [(4, 4)] * n
        self.coef = np.array([int(v) * dc_scale for p in planes for v in p.ravel()], np.int32)
        total = n * mbx * mby
        self.offsets = np.arange(total, dtype=np.uint32)
        self.quant = np.ones(total * 3, np.uint32)
        self.bands = np.zeros(total, np.uint8)
        self.preds = np.zeros(total, np.uint8)
        ys, xs = np.mgrid[0:mby, 0:mbx]
        self.cx = np.tile(xs.ravel().astype(np.uint32), n)
        self.cy = np.tile(ys.ravel().astype(np.uint32), n)
        self.planes = (abi.CoefficientPlane * n)()
        for i in range(n):
            cp = self.planes[i]
            cp.coefficient_offset = i * mbx * mby
            cp.coefficient_count = mbx * mby
            cp.macroblock_offset = i * mbx * mby
            cp.macroblock_count = mbx * mby
            cp.block_columns, cp.block_rows = geometry[i]
        cols, rows = tiles if tiles else ([mbx], [mby])
        self.tw = np.array(cols, np.uint32)
        self.th = np.array(rows, np.uint32)
        a = abi.Arena()
        u32 = lambda arr: arr.ctypes.data_as(abi.u32p)
        a.coefficients = self.coef.ctypes.data_as(C.POINTER(C.c_int32))
        a.coefficient_count = self.coef.size
        a.macroblock_count = total
        a.coefficient_offsets = u32(self.offsets)
        a.quantizers = u32(self.quant)
        a.bands = self.bands.ctypes.data_as(C.POINTER(C.c_uint8))
        a.hp_predictions = self.preds.ctypes.data_as(C.POINTER(C.c_uint8))
        a.coded_x = u32(self.cx)
        a.coded_y = u32(self.cy)
        a.planes = self.planes
        a.plane_count = n
        a.overlap = overlap
        a.hard_tiles = 1 if hard_tiles else 0
        a.tile_column_count = len(self.tw)
        a.tile_column_widths = u32(self.tw)
        a.tile_row_count = len(self.th)
        a.tile_row_heights = u32(self.th)
        a.identity = next(_ids)
        self.arena = a
        self.mbx, self.mby = mbx, mby


class FullArena(ManualArena):
    """Arena of full macroblocks (all bands, no prediction, steps 1):
    planes is a list of arrays (mby, mbx, 256) holding each macroblock's
    coefficients block by block (block k at 16k, its DC/LP at 16k)."""

    def __init__(self, planes, overlap=0):
        mby, mbx = planes[0].shape[:2]
        n = len(planes)
        super().__init__([np.zeros((mby, mbx), np.int64)] * n, overlap=overlap)
        self.coef = np.concatenate([np.asarray(p, np.int64).reshape(-1) for p in planes]).astype(np.int32)
        total = n * mbx * mby
        self.offsets = (np.arange(total, dtype=np.uint64) * 256).astype(np.uint32)
        self.bands = np.full(total, abi.BANDS_ALL, np.uint8)
        for i in range(n):
            cp = self.planes[i]
            cp.coefficient_offset = i * mbx * mby * 256
            cp.coefficient_count = mbx * mby * 256
        a = self.arena
        a.coefficients = self.coef.ctypes.data_as(C.POINTER(C.c_int32))
        a.coefficient_count = self.coef.size
        a.coefficient_offsets = self.offsets.ctypes.data_as(abi.u32p)
        a.bands = self.bands.ctypes.data_as(C.POINTER(C.c_uint8))
        a.identity = next(_ids)


def desc_for(arena: ManualArena, color, components, spec: OutputSpec, crop, scaled=False,
             alpha_mode=abi.ALPHA_NONE, separate_alpha: ManualArena = None, alpha_scaled=None):
    d = abi.ImageDesc()
    d.abi_version = abi.ABI_VERSION
    d.color_format = color
    d.component_count = components
    d.alpha_mode = alpha_mode
    d.scaled = 1 if scaled else 0
    d.primary = C.pointer(arena.arena)
    if separate_alpha is not None:
        d.separate_alpha = C.pointer(separate_alpha.arena)
    o = d.output
    native = spec.output_color in (abi.COLOR_YUV420, abi.COLOR_YUV422)  # planar YUV keeps the sampling
    o.internal_color = abi.COLOR_YUV444 if color in (abi.COLOR_YUV420, abi.COLOR_YUV422) and not native else color
    o.output_color = spec.output_color
    o.bit_depth = spec.depth
    o.shift_bits = spec.shift_bits
    o.mantissa_length = spec.mantissa_length
    o.exponent_bias = spec.exponent_bias
    if alpha_mode != abi.ALPHA_NONE:
        o.has_alpha_format = 1
        o.alpha_bit_depth = spec.depth
        o.alpha_shift_bits = spec.shift_bits
        o.alpha_mantissa_length = spec.mantissa_length
        o.alpha_exponent_bias = spec.exponent_bias
        o.alpha_scaled = (1 if scaled else 0) if alpha_scaled is None else (1 if alpha_scaled else 0)
    o.storage = spec.storage
    o.channel_layout = spec.layout
    o.channel_count = spec.channels
    o.scaled = 1 if scaled else 0
    o.red_blue_not_swapped = 1
    o.premultiply_alpha = 1 if spec.premultiply else 0
    o.crop_x, o.crop_y, o.crop_width, o.crop_height = crop
    d.surface = surface_layout(crop[2], crop[3], spec)
    d._keep = (arena, separate_alpha)
    return d


def row_of_pixels(values_per_plane, **kw):
    """Arena whose pixel k (k = 0..n-1) on the output row holds values[k]:
    macroblock k is the constant values[k]; the crop takes the pixel at
    column 15 of the first macroblock and column 0 of the next ones."""
    n = len(values_per_plane[0])
    planes = [np.array(vals, np.int64).reshape(1, n) for vals in values_per_plane]
    return ManualArena(planes, **kw), n


class BlockArena:
    """Arena of explicit macroblocks, one list per plane. A macroblock is
    (coefficients, bands, (dc, lp, hp) steps) with its 256 coefficients
    block by block (block k's DC/low-pass value at 16k); the arena stores
    what its bands code (1, blocks or blocks*16 values). Macroblocks are in
    raster order over mbx x mby, placed at `origin` in the coded image, whose
    tile grid `tiles` = (column widths, row heights) must cover the
    window."""

    def __init__(self, planes, geometry, mbx, mby, overlap=0, tiles=None, hard_tiles=False, origin=(0, 0)):
        n = len(planes)
        coef, offsets, quant, bands, cx, cy = [], [], [], [], [], []
        self.planes = (abi.CoefficientPlane * n)()
        for i, mbs in enumerate(planes):
            bc, br = geometry[i]
            blocks = bc * br
            cp = self.planes[i]
            cp.coefficient_offset = len(coef)
            cp.macroblock_offset = len(offsets)
            cp.block_columns, cp.block_rows = bc, br
            assert len(mbs) == mbx * mby
            for m, (c, b, q) in enumerate(mbs):
                c = np.asarray(c, np.int64).reshape(-1)
                offsets.append(len(coef))
                if b == abi.BANDS_DC_ONLY:
                    coef.append(int(c[0]))
                elif b == abi.BANDS_NO_HIGH_PASS:
                    coef.extend(int(c[16 * k]) for k in range(blocks))
                else:
                    coef.extend(int(v) for v in c[: blocks * 16])
                quant.extend(q)
                bands.append(b)
                cx.append(origin[0] + m % mbx)
                cy.append(origin[1] + m // mbx)
            cp.coefficient_count = len(coef) - cp.coefficient_offset
            cp.macroblock_count = len(offsets) - cp.macroblock_offset
        self.coef = np.array(coef if coef else [0], np.int32)
        self.offsets = np.array(offsets, np.uint32)
        self.quant = np.array(quant, np.uint32)
        self.bands = np.array(bands, np.uint8)
        self.preds = np.zeros(len(bands), np.uint8)
        self.cx = np.array(cx, np.uint32)
        self.cy = np.array(cy, np.uint32)
        cols, rows = tiles if tiles else ([origin[0] + mbx], [origin[1] + mby])
        self.tw = np.array(cols, np.uint32)
        self.th = np.array(rows, np.uint32)
        a = abi.Arena()
        u32 = lambda arr: arr.ctypes.data_as(abi.u32p)
        a.coefficients = self.coef.ctypes.data_as(C.POINTER(C.c_int32))
        a.coefficient_count = len(coef)
        a.macroblock_count = len(offsets)
        a.coefficient_offsets = u32(self.offsets)
        a.quantizers = u32(self.quant)
        a.bands = self.bands.ctypes.data_as(C.POINTER(C.c_uint8))
        a.hp_predictions = self.preds.ctypes.data_as(C.POINTER(C.c_uint8))
        a.coded_x = u32(self.cx)
        a.coded_y = u32(self.cy)
        a.planes = self.planes
        a.plane_count = n
        a.overlap = overlap
        a.hard_tiles = 1 if hard_tiles else 0
        a.tile_column_count = len(self.tw)
        a.tile_column_widths = u32(self.tw)
        a.tile_row_count = len(self.th)
        a.tile_row_heights = u32(self.th)
        a.identity = next(_ids)
        self.arena = a
        self.mbx, self.mby = mbx, mby


def dc_macroblock(value, step=1, bands=abi.BANDS_DC_ONLY):
    """A macroblock whose DC is `value` and every other coefficient zero
    (the reference tests' dc_macroblock)."""
    c = np.zeros(256, np.int64)
    c[0] = value
    return (c, bands, (step, 1, 1))
