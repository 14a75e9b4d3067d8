"""The benchmark workloads as seeded synthetic arenas.

  c1  512x512 RGB u8, lossless 4:4:4, one tile, overlap 1 -> interleaved RGB u8
  c2  4096x4096 gray u16 pink noise, 8x8 grid of 512-px soft tiles, overlap 2,
      quality 75 with all bands (flexbits kept) -> gray u16
  c3  256 RGB images, sides uniform in 128..2048 step 16, 8/16-bit at 50/50,
      4:2:0, overlap 1 or 2, quality 50..100 -> device-resident NCHW
  c4  128 x 1024x1024 RGBA f32 lognormal HDR + uniform alpha, lossless,
      separate alpha, overlap 1 -> planar f32 + alpha

Entropy coding is not modelled: the forward analysis produces the arenas the
entropy decoder would hand the reconstruction.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import abi, synth
from .decoder import EncodedImage, OutputSpec, quality_qp

SEED = 0x4A585231


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def rgb_out(depth, planar):
    if depth == 8:
        return OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB,
                          output_color=abi.COLOR_RGB, planar=planar)
    return OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_RGB,
                      output_color=abi.COLOR_RGB, planar=planar)


def c1():
    px = synth.gradient_noise(512, 512, SEED, depth=8, sigma=8.0)
    enc = EncodedImage(px, color=abi.COLOR_YUV444, overlap=1, scaled=False, qp=(0, 0, 0))
    return [(enc, rgb_out(8, False))]


def c2():
    px = synth.pink_noise(4096, 4096, SEED)
    qp = quality_qp(75)
    enc = EncodedImage(px, color=abi.COLOR_LUMA, overlap=2, tiles=([32] * 8, [32] * 8), hard_tiles=False,
                       scaled=True, qp=(qp, qp, qp + 4), bands=abi.BANDS_ALL)
    spec = OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_LUMA,
                      output_color=abi.COLOR_LUMA)
    return [(enc, spec)]


def c3_params(count=256, seed=SEED):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w = int(rng.integers(8, 129)) * 16
        h = int(rng.integers(8, 129)) * 16
        depth = 8 if rng.random() < 0.5 else 16
        ov = 1 if rng.random() < 0.5 else 2
        q = int(rng.integers(50, 101))
        out.append(dict(index=i, w=w, h=h, depth=depth, ov=ov, quality=q, seed=seed + i))
    return out


def _c3_one(p):
    px = synth.gradient_noise(p["w"], p["h"], p["seed"], depth=p["depth"], sigma=8.0)
    q = p["quality"]
    qp = quality_qp(q)
    enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=p["ov"], scaled=q < 100,
                       qp=(qp, qp, qp + (4 if q < 100 else 0)), seed=p["seed"])
    return enc, rgb_out(p["depth"], True)


def c3(count=256, seed=SEED):
    params = c3_params(count, seed)
    with ThreadPoolExecutor(_threads()) as ex:
        return list(ex.map(_c3_one, params))


def _c4_one(i):
    px = synth.hdr_rgba(1024, 1024, SEED + i)
    enc = EncodedImage(px, color=abi.COLOR_YUV444, overlap=1, scaled=False, qp=(0, 0, 0),
                       alpha_mode=abi.ALPHA_SEPARATE, mantissa_length=13, exponent_bias=4)
    spec = OutputSpec(storage=abi.STORE_F32, depth=abi.DEPTH_F32, layout=abi.LAYOUT_RGBA,
                      output_color=abi.COLOR_RGB, planar=True, mantissa_length=13, exponent_bias=4)
    return enc, spec


def c4(count=128):
    with ThreadPoolExecutor(_threads()) as ex:
        return list(ex.map(_c4_one, range(count)))


WORKLOADS = {
    "c1": ("512x512 RGB u8 lossless 4:4:4 ov1, one tile -> interleaved RGB u8", c1),
    "c2": ("4096x4096 gray u16 pink noise, 8x8 soft tiles, ov2, q75 all bands -> u16", c2),
    "c3": ("256 mixed RGB 4:2:0 images 128..2048 px, 8/16-bit, ov1/ov2, q50..100 -> NCHW", c3),
    "c4": ("128 x 1024x1024 RGBA f32 lossless, separate alpha, ov1 -> planar f32 + alpha", c4),
}
