"""The benchmark workloads as seeded synthetic arenas.

  c1  512x512 RGB u8, lossless 4:4:4, one tile, overlap 1 -> interleaved RGB u8
  c2  4096x4096 gray u16 pink noise, 8x8 grid of 512-px soft tiles, overlap 2,
      quality 75 (uniform DC/LP/HP parameters) with all bands (flexbits kept)
      -> gray u16
  c3  256 RGB images, sides uniform in 128..2048 step 16, 8/16-bit at 50/50,
      4:2:0, overlap 1 or 2, quality 50..100 -> device-resident NCHW
  c4  128 x 1024x1024 RGBA f32 lognormal HDR + uniform alpha, lossless,
      separate alpha, overlap 1 -> planar f32 + alpha
  c5  serving stream: Poisson arrivals at 2000/s, 70% RGB u8 256..1024 px
      q75, 20% 4096x4096 gray u16 8x8 tiles, 10% 512x512 regions of an
      8192x8192 RGB u8 source coded in 512-px tiles (c5_pool, c5_schedule)

Qualities resolve to quantizers by the reference encoder's scale
(`quality_parameters`). Entropy coding is not modelled: the forward analysis
produces the arenas the entropy decoder would hand the reconstruction.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import abi, synth
from .decoder import EncodedImage, OutputSpec

SEED = 0x4A585231


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def rgb_out(depth, planar):
    if depth == 8:
        return OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB,
                          output_color=abi.COLOR_RGB, planar=planar)
    return OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_RGB,
                      output_color=abi.COLOR_RGB, planar=planar)


def c1_pixels():
    return synth.gradient_noise(512, 512, SEED, depth=8, sigma=8.0)


def c1():
    enc = EncodedImage(c1_pixels(), color=abi.COLOR_YUV444, overlap=1, scaled=False, qp=(0, 0, 0))
    return [(enc, rgb_out(8, False))]


def c2():
    px = synth.pink_noise(4096, 4096, SEED)
    enc = EncodedImage.at_quality(px, 75, color=abi.COLOR_LUMA, overlap=2, tiles=([32] * 8, [32] * 8),
                                  hard_tiles=False, bands=abi.BANDS_ALL)
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
    enc = EncodedImage.at_quality(px, p["quality"], color=abi.COLOR_YUV420, overlap=p["ov"], seed=p["seed"])
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


# ---- c5: serving stream ------------------------------------------------------
C5_RATE = 2000.0      # requests per second (Poisson)
C5_MIX = (0.7, 0.2, 0.1)  # small RGB u8, 4096x4096 gray u16, 512x512 region of 8192x8192


def _c5_small(p):
    px = synth.gradient_noise(p["w"], p["h"], p["seed"], depth=8, sigma=8.0)
    enc = EncodedImage.at_quality(px, 75, color=abi.COLOR_YUV444, seed=p["seed"])
    return enc, rgb_out(8, False)


def _c5_gray(seed):
    px = synth.pink_noise(4096, 4096, seed)
    enc = EncodedImage.at_quality(px, 75, color=abi.COLOR_LUMA, overlap=2, tiles=([32] * 8, [32] * 8),
                                  hard_tiles=False, bands=abi.BANDS_ALL)
    return enc, OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_LUMA,
                           output_color=abi.COLOR_LUMA)


def c5_pool(small=96, gray=2, regions=64, seed=SEED + 5000, cache=True):
    """The distinct images serving requests draw from: `small` RGB u8 images
    of 256..1024 px a side (step 16) at quality 75, 4:4:4, overlap 1 ->
    interleaved RGB u8; `gray` 4096x4096 u16 pink-noise images in an 8x8
    grid of soft tiles, overlap 2, quality 75 -> u16; and `regions` 512x512
    windows at 16-aligned positions of one 8192x8192 RGB u8 source coded in
    16x16 tiles of 512 px, each as the arena a region decode prepares.
    Returns {"small": [...], "gray": [...], "region": [...]} of
    (image, spec). Without `cache` the arenas have no identity, so every
    request uploads its coefficients again."""
    from .prepare import RegionImage
    rng = np.random.default_rng(seed)
    ps = [dict(w=int(rng.integers(16, 65)) * 16, h=int(rng.integers(16, 65)) * 16, seed=seed + i)
          for i in range(small)]
    with ThreadPoolExecutor(_threads()) as ex:
        smalls = list(ex.map(_c5_small, ps))
        grays = list(ex.map(_c5_gray, [seed + 1000 + i for i in range(gray)]))
    src_px = synth.gradient_noise_large(8192, 8192, seed + 2000)
    source = EncodedImage.at_quality(src_px, 75, color=abi.COLOR_YUV444, tiles=([32] * 16, [32] * 16),
                                     hard_tiles=False)
    del src_px
    regs = []
    for _ in range(regions):
        x = int(rng.integers(0, (8192 - 512) // 16 + 1)) * 16
        y = int(rng.integers(0, (8192 - 512) // 16 + 1)) * 16
        reg = RegionImage(source, (x, y, 512, 512), cache=cache)
        spec = rgb_out(8, False)
        spec.crop = (x, y, 512, 512)
        regs.append((reg, spec))
    if not cache:
        for enc, _ in smalls + grays:
            enc.primary.contents.identity = 0
    return {"small": smalls, "gray": grays, "region": regs, "source": source}


def c5_schedule(seconds, rate=C5_RATE, mix=C5_MIX, counts=(96, 2, 64), seed=SEED + 6000):
    """Poisson arrivals over `seconds`: (arrival seconds, kind, pool index)."""
    rng = np.random.default_rng(seed)
    n = int(rng.poisson(rate * seconds))
    t = np.sort(rng.uniform(0.0, seconds, n))
    kinds = rng.choice(3, size=n, p=np.asarray(mix) / sum(mix))
    names = ("small", "gray", "region")
    return [(float(t[i]), names[k], int(rng.integers(0, counts[k]))) for i, k in enumerate(kinds)]


WORKLOADS = {
    "c1": ("512x512 RGB u8 lossless 4:4:4 ov1, one tile -> interleaved RGB u8", c1),
    "c2": ("4096x4096 gray u16 pink noise, 8x8 soft tiles, ov2, q75 all bands -> u16", c2),
    "c3": ("256 mixed RGB 4:2:0 images 128..2048 px, 8/16-bit, ov1/ov2, q50..100 -> NCHW", c3),
    "c4": ("128 x 1024x1024 RGBA f32 lossless, separate alpha, ov1 -> planar f32 + alpha", c4),
}
