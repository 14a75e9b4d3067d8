"""Region decodes: the macroblock window a region needs, as its own arena.

A decode of an output region reconstructs only the tiles that intersect the
region's macroblocks widened by one macroblock where the overlap filter or
chroma interpolation reaches across (crates/jxr-native/src/prepare.rs:231-359:
`reconstruction_bounds`, `tile_plans`). The entropy decoder hands over the
coefficients of exactly those tiles; `RegionImage` builds that arena from a
whole-image arena, so the reconstruction sees what a region decode's prepare
stage would give it. The output crop stays in coded-plane coordinates.
"""
import ctypes as C
import itertools

import numpy as np

from . import abi

_ids = itertools.count(1 << 48)


def reconstruction_bounds(region, macroblocks, overlap, chroma_upsampling, margins=(0, 0)):
    """[x0, y0, x1, y1) in macroblocks (prepare.rs:322-359)."""
    x, y, w, h = region
    sx, sy = margins[0] + x, margins[1] + y
    ex, ey = sx + w, sy + h
    halo = 1 if (overlap != 0 or chroma_upsampling) else 0
    return (max(sx // 16 - halo, 0), max(sy // 16 - halo, 0),
            min(-(-ex // 16) + halo, macroblocks[0]), min(-(-ey // 16) + halo, macroblocks[1]))


def required_tiles(sizes, lo, hi):
    """Start and end (macroblocks) of the run of tiles along one axis that
    intersect [lo, hi) (prepare.rs:265-303, `required_for_reconstruction`)."""
    start = end = None
    pos = 0
    for s in (int(v) for v in sizes):
        if pos < hi and pos + s > lo:
            start = pos if start is None else start
            end = pos + s
        pos += s
    return start, end


def _array(ptr, n, dtype):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True) if n else np.zeros(0, dtype)


def _arena_arrays(src: abi.Arena):
    """Per-macroblock arrays of an arena (copies) and its coefficients (a view)."""
    n = src.macroblock_count
    return (_array(src.coded_x, n, np.uint32), _array(src.coded_y, n, np.uint32),
            _array(src.coefficient_offsets, n, np.uint64), _array(src.bands, n, np.uint8),
            _array(src.hp_predictions, n, np.uint8), _array(src.quantizers, n * 3, np.uint32).reshape(n, 3),
            np.ctypeslib.as_array(src.coefficients, shape=(src.coefficient_count,)))


class _WindowArena:
    """The macroblocks of one arena that lie in [x0, x1) x [y0, y1), in the
    arena's own order (component-major, tile-major), with their coefficients."""

    def __init__(self, src: abi.Arena, window, cache: bool, arrays=None):
        x0, y0, x1, y1 = window
        cx, cy, offs, bands, preds, quant, coef = arrays if arrays is not None else _arena_arrays(src)
        keep = (cx >= x0) & (cx < x1) & (cy >= y0) & (cy < y1)
        self.planes = (abi.CoefficientPlane * src.plane_count)()
        sel_all, counts_all = [], []
        mb_base = coef_base = 0
        for p in range(src.plane_count):
            sp = src.planes[p]
            mbs = np.arange(sp.macroblock_offset, sp.macroblock_offset + sp.macroblock_count)
            sel = mbs[keep[mbs]]
            blocks = sp.block_columns * sp.block_rows
            cnt = np.where(bands[sel] == abi.BANDS_DC_ONLY, 1,
                           np.where(bands[sel] == abi.BANDS_NO_HIGH_PASS, blocks, blocks * 16)).astype(np.uint64)
            dp = self.planes[p]
            dp.macroblock_offset, dp.macroblock_count = mb_base, sel.size
            dp.coefficient_offset, dp.coefficient_count = coef_base, int(cnt.sum())
            dp.block_columns, dp.block_rows = sp.block_columns, sp.block_rows
            mb_base += sel.size
            coef_base += int(cnt.sum())
            sel_all.append(sel)
            counts_all.append(cnt)
        sel = np.concatenate(sel_all)
        cnt = np.concatenate(counts_all)
        starts = offs[sel]
        new_offs = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint64)
        # gather each kept macroblock's run of coefficients
        idx = np.repeat(starts - new_offs, cnt.astype(np.int64)) + np.arange(int(cnt.sum()), dtype=np.uint64)
        self.coef = np.ascontiguousarray(coef[idx.astype(np.int64)], np.int32)
        self.offsets = new_offs.astype(np.uint32)
        self.quant = np.ascontiguousarray(quant[sel].reshape(-1))
        self.bands = np.ascontiguousarray(bands[sel])
        self.preds = np.ascontiguousarray(preds[sel])
        self.cx = np.ascontiguousarray(cx[sel])
        self.cy = np.ascontiguousarray(cy[sel])
        self.tw = _array(src.tile_column_widths, src.tile_column_count, np.uint32)
        self.th = _array(src.tile_row_heights, src.tile_row_count, np.uint32)
        a = abi.Arena()
        u32 = lambda arr: arr.ctypes.data_as(abi.u32p)
        a.coefficients = self.coef.ctypes.data_as(C.POINTER(C.c_int32))
        a.coefficient_count = self.coef.size
        a.macroblock_count = sel.size
        a.coefficient_offsets = u32(self.offsets)
        a.quantizers = u32(self.quant)
        a.bands = self.bands.ctypes.data_as(C.POINTER(C.c_uint8))
        a.hp_predictions = self.preds.ctypes.data_as(C.POINTER(C.c_uint8))
        a.coded_x = u32(self.cx)
        a.coded_y = u32(self.cy)
        a.planes = self.planes
        a.plane_count = src.plane_count
        a.overlap = src.overlap
        a.hard_tiles = src.hard_tiles
        a.tile_column_count = self.tw.size
        a.tile_column_widths = u32(self.tw)
        a.tile_row_count = self.th.size
        a.tile_row_heights = u32(self.th)
        a.identity = next(_ids) if cache else 0
        self.arena = a

    @property
    def coefficient_bytes(self):
        return self.coef.nbytes


class RegionImage:
    """An encoded image restricted to what a decode of `region` (x, y, w, h
    in the image) reconstructs. Usable wherever an EncodedImage is (its
    `primary`/`alpha` arenas, colour format, alpha mode, scaling); `region`
    is the crop to request. `cache` gives the arenas an identity so that
    the session's upload cache keeps them."""

    def __init__(self, enc, region, chroma_upsampling=None, cache=True):
        src = enc.primary.contents
        mbx = int(sum(_array(src.tile_column_widths, src.tile_column_count, np.uint32)))
        mby = int(sum(_array(src.tile_row_heights, src.tile_row_count, np.uint32)))
        if chroma_upsampling is None:
            chroma_upsampling = enc.color_format in (abi.COLOR_YUV420, abi.COLOR_YUV422)
        bx0, by0, bx1, by1 = reconstruction_bounds(region, (mbx, mby), src.overlap, chroma_upsampling)
        tx0, tx1 = required_tiles(_array(src.tile_column_widths, src.tile_column_count, np.uint32), bx0, bx1)
        ty0, ty1 = required_tiles(_array(src.tile_row_heights, src.tile_row_count, np.uint32), by0, by1)
        self.window = (tx0, ty0, tx1, ty1)
        # the source's per-macroblock arrays, read once per source image
        memo = getattr(enc, "_region_arrays", None)
        if memo is None:
            memo = (_arena_arrays(src), _arena_arrays(enc.alpha.contents) if enc.alpha is not None else None)
            try:
                enc._region_arrays = memo
            except AttributeError:
                pass
        self._primary = _WindowArena(src, self.window, cache, memo[0])
        self._alpha = None
        if enc.alpha is not None:
            self._alpha = _WindowArena(enc.alpha.contents, self.window, cache, memo[1])
        self.region = tuple(region)
        self.width, self.height = enc.width, enc.height
        self.channels = enc.channels
        self.color_format = enc.color_format
        self.component_count = enc.component_count
        self.alpha_mode = enc.alpha_mode
        self.scaled = enc.scaled
        self.overlap = enc.overlap
        self._source = enc

    @property
    def primary(self):
        return C.pointer(self._primary.arena)

    @property
    def alpha(self):
        return C.pointer(self._alpha.arena) if self._alpha is not None else None

    def coefficient_bytes(self) -> int:
        return self._primary.coefficient_bytes + (self._alpha.coefficient_bytes if self._alpha else 0)
