"""Hand-built arenas whose reconstructed samples exceed 2^27, the largest
magnitude for which the kernels' output stage runs plain i32 arithmetic.
Each block holds fifteen high-pass coefficients of one magnitude whose signs
follow the inverse core transform's basis at one sample position, so the
block's samples add up to about 3.4 times that magnitude while every
transform intermediate stays inside i32 (the oracle decodes them; the case
with the largest magnitude overflows only in the colour conversion)."""
import numpy as np

import oracle_lib as O
from arena_util import FullArena, desc_for
from jxr_amd import OutputSpec, abi

GRAY_I32 = OutputSpec(storage=abi.STORE_I32, depth=abi.DEPTH_I32, layout=abi.LAYOUT_LUMA,
                      output_color=abi.COLOR_LUMA)
RGB8 = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB)
RGB_I32 = OutputSpec(storage=abi.STORE_I32, depth=abi.DEPTH_I32, layout=abi.LAYOUT_RGB,
                     output_color=abi.COLOR_RGB)

_basis = None


def core_basis():
    """Row k: the inverse core transform of coefficient k alone (scaled)."""
    global _basis
    if _basis is None:
        b = np.zeros((16, 16))
        for k in range(16):
            x = np.zeros(16, np.int64)
            x[k] = 1 << 20
            out, err = O.op("jxro_inverse_core_transform", x)
            assert not err
            b[k] = np.asarray(out) / (1 << 20)
        _basis = b
    return _basis


def hp_block(pos, magnitude, sign=1):
    s = np.sign(core_basis()[:, pos])
    s[0] = 0
    return (s * magnitude * sign).astype(np.int64)


def random_plane(mby, mbx, magnitude, seed, fill=0.5):
    """Macroblocks whose blocks are, with probability `fill`, a high-pass
    block of the given magnitude, random position and sign; zero otherwise."""
    rng = np.random.default_rng(seed)
    p = np.zeros((mby, mbx, 256), np.int64)
    for my in range(mby):
        for mx in range(mbx):
            for k in range(16):
                if rng.random() < fill:
                    p[my, mx, 16 * k:16 * k + 16] = hp_block(int(rng.integers(16)), magnitude,
                                                             int(rng.choice([-1, 1])))
    return p


def large_gray(overlap, magnitude=3 << 25, seed=7):
    """Gray i32 image, 4x3 macroblocks, samples up to ~3.6 * 2^27."""
    arena = FullArena([random_plane(3, 4, magnitude, seed)], overlap=overlap)
    return desc_for(arena, abi.COLOR_LUMA, 1, GRAY_I32, (0, 0, 64, 48))


def mixed_gray(overlap, seed=11):
    """8x8 macroblocks: the left half reaches ~2.6 * 2^27, the right half
    stays small, so tiles of one image take both output paths."""
    big = random_plane(8, 4, 3 << 25, seed)
    small = random_plane(8, 4, 1 << 10, seed + 1)
    arena = FullArena([np.concatenate([big, small], axis=1)], overlap=overlap)
    return desc_for(arena, abi.COLOR_LUMA, 1, GRAY_I32, (0, 0, 128, 128))


def large_yuv(magnitude, spec=RGB_I32):
    """YUV 4:4:4, 2x2 macroblocks, luma positive and chroma negative at the
    same sample positions; at magnitude 5 * 2^26 the red channel's
    conversion leaves i32 (output stage error)."""
    def plane(sign):
        p = np.zeros((2, 2, 256), np.int64)
        for k in range(16):
            p[:, :, 16 * k:16 * k + 16] = hp_block(5, magnitude, sign)
        return p
    arena = FullArena([plane(1), plane(-1), plane(-1)], overlap=0)
    return desc_for(arena, abi.COLOR_YUV444, 3, spec, (0, 0, 32, 32))


def cases():
    """(name, desc, expected error stage or 0)."""
    out = []
    for ov in (0, 1, 2):
        out.append((f"large_gray_ov{ov}", large_gray(ov), 0))
        out.append((f"mixed_gray_ov{ov}", mixed_gray(ov), 0))
    out.append(("large_yuv_rgb_i32", large_yuv(3 << 26), 0))
    out.append(("large_yuv_rgb8", large_yuv(3 << 26, RGB8), 0))
    out.append(("yuv_conversion_overflow", large_yuv(5 << 26, RGB8), abi.STAGE_OUTPUT))
    return out
