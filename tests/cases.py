"""Seeded parity cases shared by the GPU tests: each builds arenas with the
forward analysis and an output policy, at sizes the oracle decodes in
milliseconds."""
import numpy as np

from jxr_amd import EncodedImage, OutputSpec, abi, synth


def quality_qp(quality: int) -> int:
    """A QP ramp for the parity cases, which exercise a spread of steps: 100
    is lossless (step 1), lower qualities coarser. (The benchmark workloads
    use the reference encoder's scale, jxr_amd.quality_parameters.)"""
    if quality >= 100:
        return 0
    return int(round(4 + (100 - quality) * 0.8))

RGB8 = dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB)
RGB16 = dict(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB)
GRAY16 = dict(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)
GRAY8 = dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)
RGBAF = dict(storage=abi.STORE_F32, depth=abi.DEPTH_F32, layout=abi.LAYOUT_RGBA, output_color=abi.COLOR_RGB)


def _tiles(w, h, cols, rows):
    mbx, mby = (w + 15) // 16, (h + 15) // 16
    def split(n, parts):
        base = [n // parts] * parts
        for i in range(n % parts):
            base[i] += 1
        return [b for b in base if b > 0]
    return split(mbx, cols), split(mby, rows)


def make(case):
    """Returns (EncodedImage, OutputSpec) for a case dict."""
    kind = case.get("image", "rgb8")
    w, h, seed = case["w"], case["h"], case.get("seed", 1)
    if kind == "rgb8":
        px = synth.gradient_noise(w, h, seed, depth=8)
    elif kind == "rgb16":
        px = synth.gradient_noise(w, h, seed, depth=16)
    elif kind == "gray16":
        px = synth.pink_noise(w, h, seed)
    elif kind == "gray8":
        px = synth.gradient_noise(w, h, seed, channels=1, depth=8)
    elif kind == "rgbaf":
        px = synth.hdr_rgba(w, h, seed)
    elif kind == "rgbf":
        px = np.ascontiguousarray(synth.hdr_rgba(w, h, seed)[..., :3] - np.float32(1.5))  # signed
    elif kind == "grayaf":
        px = np.ascontiguousarray(synth.hdr_rgba(w, h, seed)[..., 2:])
    elif kind == "grayf":
        px = np.ascontiguousarray(synth.hdr_rgba(w, h, seed)[..., 0])
    elif kind == "rgba8":
        px = synth.gradient_noise(w, h, seed, channels=4, depth=8)
    elif kind == "graya8":
        px = synth.gradient_noise(w, h, seed, channels=2, depth=8)
    elif kind == "graya16":
        px = synth.gradient_noise(w, h, seed, channels=2, depth=16)
    else:
        raise ValueError(kind)
    q = case.get("quality", 100)
    qp = (quality_qp(q), quality_qp(q), quality_qp(q) + (4 if q < 100 else 0))
    tiles = None
    if "tiles" in case:
        tiles = _tiles(w, h, *case["tiles"])
    enc = EncodedImage(
        px, color=case.get("color", abi.COLOR_YUV444), overlap=case.get("ov", 1), tiles=tiles,
        hard_tiles=case.get("hard", False), scaled=case.get("scaled", q < 100), qp=qp,
        bands=case.get("bands", abi.BANDS_ALL), alpha_mode=case.get("alpha", abi.ALPHA_NONE),
        alpha_qp=qp[0], seed=seed, vary_quantizers=case.get("vary", False),
        shift_bits=case.get("shift_bits", 0), mantissa_length=case.get("mantissa", 13),
        exponent_bias=case.get("bias", 4))
    spec = OutputSpec(**case.get("out", RGB8))
    if "crop" in case:
        spec.crop = case["crop"]
    spec.centering = case.get("centering", (0, 0))
    spec.scale = case.get("scale", abi.SCALE_FULL)
    return enc, spec, px


CASES = {
    "c1_rgb8_444_lossless_ov1": dict(w=128, h=96, ov=1),
    "rgb8_444_ov0": dict(w=112, h=80, ov=0, quality=80),
    "rgb8_444_ov2_hard": dict(w=144, h=112, ov=2, quality=70, tiles=(3, 2), hard=True, vary=True),
    "rgb8_444_ov2_soft_odd": dict(w=100, h=70, ov=2, quality=90, tiles=(2, 2)),
    "rgb8_420_ov1": dict(w=128, h=128, ov=1, quality=75, color=abi.COLOR_YUV420),
    "rgb8_420_ov2_hard": dict(w=160, h=96, ov=2, quality=60, color=abi.COLOR_YUV420, tiles=(4, 3), hard=True),
    "rgb8_420_ov2_hard_1mb": dict(w=80, h=64, ov=2, quality=85, color=abi.COLOR_YUV420, tiles=(5, 4), hard=True),
    "rgb8_420_ov2_soft": dict(w=96, h=80, ov=2, quality=100, color=abi.COLOR_YUV420, tiles=(2, 2)),
    "rgb16_420_ov2": dict(w=112, h=96, ov=2, quality=55, color=abi.COLOR_YUV420, image="rgb16", out=RGB16),
    "rgb8_422_ov2_hard": dict(w=128, h=80, ov=2, quality=75, color=abi.COLOR_YUV422, tiles=(3, 2), hard=True),
    "rgb8_422_ov1_cent": dict(w=96, h=64, ov=1, quality=75, color=abi.COLOR_YUV422, centering=(2, 0)),
    "rgb8_420_cent31": dict(w=96, h=64, ov=2, quality=65, color=abi.COLOR_YUV420, centering=(3, 1)),
    "gray16_ov2_q75": dict(w=256, h=192, ov=2, quality=75, image="gray16", color=abi.COLOR_LUMA, out=GRAY16,
                           tiles=(4, 3), vary=True),
    "gray16_ov2_hard": dict(w=192, h=160, ov=2, quality=75, image="gray16", color=abi.COLOR_LUMA, out=GRAY16,
                            tiles=(4, 3), hard=True),
    "gray8_dc_only": dict(w=64, h=48, ov=1, quality=70, image="gray8", color=abi.COLOR_LUMA, out=GRAY8,
                          bands=abi.BANDS_DC_ONLY),
    "rgb8_no_hp": dict(w=80, h=64, ov=2, quality=70, bands=abi.BANDS_NO_HIGH_PASS, color=abi.COLOR_YUV420),
    "rgb8_no_flexbits": dict(w=96, h=48, ov=1, quality=60, bands=abi.BANDS_NO_FLEXBITS, color=abi.COLOR_YUV422,
                             tiles=(2, 2), vary=True),
    "rgbaf_separate_ov1": dict(w=96, h=64, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE,
                               out=dict(RGBAF, planar=True)),
    "rgbaf_integrated_ov2": dict(w=80, h=48, ov=2, image="rgbaf", alpha=abi.ALPHA_INTEGRATED, quality=90,
                                 out=dict(RGBAF, planar=False)),
    "rgbaf_premultiply": dict(w=64, h=48, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE,
                              out=dict(RGBAF, premultiply=True)),
    "rgbaf_mant10_bias1": dict(w=64, h=48, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE, mantissa=10, bias=1,
                               out=dict(RGBAF, planar=True, mantissa_length=10, exponent_bias=1)),
    "rgbaf_mant16_bias6_lossy": dict(w=80, h=48, ov=2, image="rgbaf", alpha=abi.ALPHA_INTEGRATED, quality=90,
                                     mantissa=16, bias=6,
                                     out=dict(RGBAF, planar=False, mantissa_length=16, exponent_bias=6)),
    # the f32 conversion's edge paths: a coded subnormal normalised only as
    # far as the exponent allows (bias 126), exponents below the f32 range
    # (bias 200) and above it (bias -130), both clipped to the largest float;
    # |bias| above 2^23 is decoded by the general output path
    "rgbf_bgr_lossy": dict(w=80, h=48, ov=2, image="rgbf", quality=85,
                           out=dict(RGBAF, layout=abi.LAYOUT_BGR)),
    "rgbaf_bias126_planar": dict(w=64, h=48, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE,
                                 out=dict(RGBAF, planar=True, exponent_bias=126)),
    "rgbaf_bias200_lossy": dict(w=64, h=48, ov=2, image="rgbaf", alpha=abi.ALPHA_INTEGRATED, quality=80,
                                out=dict(RGBAF, exponent_bias=200)),
    "rgbaf_bias_m130": dict(w=48, h=48, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE,
                            out=dict(RGBAF, planar=True, exponent_bias=-130)),
    "rgbaf_bias_2p23": dict(w=48, h=32, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE,
                            out=dict(RGBAF, exponent_bias=1 << 23)),
    "rgbaf_bias_over_2p23": dict(w=48, h=32, ov=1, image="rgbaf", alpha=abi.ALPHA_SEPARATE,
                                 out=dict(RGBAF, exponent_bias=(1 << 23) + 1)),
    "grayaf_luma_alpha": dict(w=64, h=48, ov=1, image="grayaf", color=abi.COLOR_LUMA, alpha=abi.ALPHA_SEPARATE,
                              quality=90,
                              out=dict(RGBAF, layout=abi.LAYOUT_LUMA_ALPHA, output_color=abi.COLOR_LUMA)),
    "grayf_to_rgb_planar": dict(w=64, h=32, ov=2, image="grayf", color=abi.COLOR_LUMA,
                                out=dict(RGBAF, layout=abi.LAYOUT_RGB, planar=True)),
    "rgba8_bgra_premultiply": dict(w=80, h=64, ov=1, image="rgba8", alpha=abi.ALPHA_INTEGRATED, quality=80,
                                   out=dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_BGRA,
                                            output_color=abi.COLOR_RGB, premultiply=True)),
    "graya8_luma_alpha": dict(w=80, h=48, ov=1, quality=75, image="graya8", color=abi.COLOR_LUMA,
                              alpha=abi.ALPHA_SEPARATE,
                              out=dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_LUMA_ALPHA,
                                       output_color=abi.COLOR_LUMA)),
    "graya16_luma_alpha_integrated": dict(w=64, h=64, ov=2, quality=85, image="graya16", color=abi.COLOR_LUMA,
                                          alpha=abi.ALPHA_INTEGRATED,
                                          out=dict(storage=abi.STORE_U16, depth=abi.DEPTH_U16,
                                                   layout=abi.LAYOUT_LUMA_ALPHA, output_color=abi.COLOR_LUMA)),
    "rgb16_420_bgr": dict(w=96, h=80, ov=2, quality=70, image="rgb16", color=abi.COLOR_YUV420,
                          out=dict(RGB16, layout=abi.LAYOUT_BGR)),
    "rgb8_bgrx": dict(w=64, h=64, ov=1, quality=80,
                      out=dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_BGRX,
                               output_color=abi.COLOR_RGB)),
    "rgb16_i16_shift": dict(w=64, h=48, ov=1, quality=80, image="rgb16",
                            out=dict(storage=abi.STORE_I16, depth=abi.DEPTH_I16, layout=abi.LAYOUT_RGB,
                                     output_color=abi.COLOR_RGB, shift_bits=2)),
    "rgb16_i32_shift": dict(w=64, h=48, ov=2, quality=60, image="rgb16",
                            out=dict(storage=abi.STORE_I32, depth=abi.DEPTH_I32, layout=abi.LAYOUT_RGB,
                                     output_color=abi.COLOR_RGB, shift_bits=5)),
    "rgb8_f16": dict(w=64, h=48, ov=1, quality=80,
                     out=dict(storage=abi.STORE_F16, depth=abi.DEPTH_F16, layout=abi.LAYOUT_RGB,
                              output_color=abi.COLOR_RGB)),
    "rgb16_u16_shift": dict(w=64, h=48, ov=1, quality=70, image="rgb16",
                            out=dict(RGB16, shift_bits=3)),
    "gray8_to_rgb": dict(w=64, h=48, ov=1, quality=80, image="gray8", color=abi.COLOR_LUMA,
                         out=dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB,
                                  output_color=abi.COLOR_RGB)),
    "rgb8_420_nchw": dict(w=144, h=80, ov=2, quality=70, color=abi.COLOR_YUV420, out=dict(RGB8, planar=True)),
    "rgb8_region_crop": dict(w=160, h=128, ov=2, quality=75, color=abi.COLOR_YUV420, crop=(37, 21, 77, 58)),
    "rgb8_region_crop_444": dict(w=128, h=96, ov=1, quality=75, crop=(16, 32, 48, 40)),
    # packed words, one-bit luma and native planar YUV
    "rgb8_rgb555": dict(w=96, h=64, ov=1, quality=80,
                        out=dict(storage=abi.STORE_PACKED16, depth=abi.DEPTH_RGB555, layout=abi.LAYOUT_RGB,
                                 output_color=abi.COLOR_RGB)),
    "rgb8_420_rgb565": dict(w=112, h=80, ov=2, quality=70, color=abi.COLOR_YUV420, crop=(5, 3, 97, 70),
                            out=dict(storage=abi.STORE_PACKED16, depth=abi.DEPTH_RGB565, layout=abi.LAYOUT_RGB,
                                     output_color=abi.COLOR_RGB)),
    "rgb16_rgb101010": dict(w=80, h=64, ov=1, quality=75, image="rgb16",
                            out=dict(storage=abi.STORE_PACKED32, depth=abi.DEPTH_RGB101010, layout=abi.LAYOUT_RGB,
                                     output_color=abi.COLOR_RGB)),
    "rgb16_rgbe": dict(w=80, h=48, ov=2, quality=85, image="rgb16",
                       out=dict(storage=abi.STORE_PACKED32, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB,
                                output_color=abi.COLOR_RGBE)),
    "gray8_to_rgb555": dict(w=64, h=48, ov=1, quality=80, image="gray8", color=abi.COLOR_LUMA,
                            out=dict(storage=abi.STORE_PACKED16, depth=abi.DEPTH_RGB555, layout=abi.LAYOUT_RGB,
                                     output_color=abi.COLOR_RGB)),
    "gray8_bits_white": dict(w=160, h=48, ov=1, quality=70, image="gray8", color=abi.COLOR_LUMA,
                             out=dict(storage=abi.STORE_BITS, depth=abi.DEPTH_BIT1_WHITE, layout=abi.LAYOUT_LUMA,
                                      output_color=abi.COLOR_LUMA)),
    "gray8_bits_black_crop": dict(w=176, h=64, ov=2, quality=60, image="gray8", color=abi.COLOR_LUMA,
                                  crop=(13, 5, 150, 51),
                                  out=dict(storage=abi.STORE_BITS, depth=abi.DEPTH_BIT1_BLACK,
                                           layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)),
    "yuv420_planar_u8": dict(w=144, h=96, ov=2, quality=70, color=abi.COLOR_YUV420,
                             out=dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_YUV,
                                      output_color=abi.COLOR_YUV420)),
    "yuv420_planar_crop": dict(w=128, h=80, ov=1, quality=80, color=abi.COLOR_YUV420, crop=(7, 9, 101, 63),
                               out=dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_YUV,
                                        output_color=abi.COLOR_YUV420)),
    "yuv422_planar_u16": dict(w=96, h=64, ov=2, quality=75, color=abi.COLOR_YUV422, image="rgb16",
                              out=dict(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_YUV,
                                       output_color=abi.COLOR_YUV422)),
    "yuv420a_planar_i16": dict(w=80, h=64, ov=1, quality=80, color=abi.COLOR_YUV420, image="rgba8",
                               alpha=abi.ALPHA_SEPARATE,
                               out=dict(storage=abi.STORE_I16, depth=abi.DEPTH_I16, layout=abi.LAYOUT_YUVA,
                                        output_color=abi.COLOR_YUV420)),
}
