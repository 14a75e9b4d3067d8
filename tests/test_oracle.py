"""Pins the CPU oracle (the reference's own known-answer vectors are in
tests/golden, checked by tests/test_golden.py): forward/inverse round trips
of every operator on the reference's seeded LCG samples, closed forms, and
brute force."""
import ctypes as C
import struct

import numpy as np
import pytest

import oracle_lib as O
from arena_util import ManualArena, desc_for
from jxr_amd import OutputSpec, abi


def lcg_samples(seed, bits, count):
    """The reference tests' generator (crates/jxr-math/src/transform.rs:402-415)."""
    state = seed
    out = []
    mask = (1 << 64) - 1
    for _ in range(count):
        state = (state * 6364136223846793005 + 1442695040888963407) & mask
        mag = (state >> 33) % (1 << bits)
        out.append(mag if (state >> 63) == 0 else -mag)
    return out


# ---- core transform (crates/jxr-math/src/transform.rs:461-516) -----------
def test_forward_transforms_concentrate_flat_input_in_dc():
    v, _ = O.op("jxro_forward_core_transform", [7] * 16)
    assert all(x == 0 for x in v[1:])
    back, _ = O.op("jxro_inverse_core_transform", v)
    assert back.tolist() == [7] * 16
    v, _ = O.op("jxro_forward_chroma_420", [6] * 4)
    assert v.tolist() == [12, 0, 0, 0]
    v, _ = O.op("jxro_forward_chroma_422", [6] * 8)
    assert v[0] != 0 and all(x == 0 for x in v[1:])
    back, _ = O.op("jxro_inverse_chroma_422", v)
    assert back.tolist() == [6] * 8


@pytest.mark.parametrize("bits", [1, 4, 9, 17, 24])
def test_transforms_are_exact_inverses(bits):
    s = lcg_samples(bits, bits, 2000 * 16)
    for i in range(0, len(s), 16):
        x = s[i:i + 16]
        f, e1 = O.op("jxro_forward_core_transform", x)
        b, e2 = O.op("jxro_inverse_core_transform", f)
        assert not e1 and not e2 and b.tolist() == x
        f, _ = O.op("jxro_inverse_core_transform", x)
        b, _ = O.op("jxro_forward_core_transform", f)
        assert b.tolist() == x
    s = lcg_samples(bits + 200, bits, 2000 * 12)
    for i in range(0, len(s), 12):
        four, eight = s[i:i + 4], s[i + 4:i + 12]
        f, _ = O.op("jxro_forward_chroma_420", four)
        assert O.op("jxro_inverse_chroma_420", f)[0].tolist() == four
        f, _ = O.op("jxro_forward_chroma_422", eight)
        assert O.op("jxro_inverse_chroma_422", f)[0].tolist() == eight


def test_transform_overflow_is_reported():
    assert O.op("jxro_forward_chroma_420", [2**31 - 1] * 4)[1]
    assert O.op("jxro_forward_chroma_422", [2**31 - 1] * 8)[1]
    assert O.op("jxro_inverse_core_transform", [2**31 - 1] * 16)[1]


# ---- overlap filters (crates/jxr-math/src/overlap.rs:450-547) -------------
@pytest.mark.parametrize("bits", [1, 4, 9, 17, 24])
def test_overlap_filters_are_exact_inverses(bits):
    s = lcg_samples(bits + 100, bits, 2000 * 20)
    for i in range(0, len(s), 20):
        four, sixteen = s[i:i + 4], s[i + 4:i + 20]
        f, _ = O.op("jxro_forward_overlap_4", four)
        assert O.op("jxro_inverse_overlap_4", f)[0].tolist() == four
        f, _ = O.op("jxro_inverse_overlap_4", four)
        assert O.op("jxro_forward_overlap_4", f)[0].tolist() == four
        f, _ = O.op("jxro_forward_overlap_4x4", sixteen)
        assert O.op("jxro_inverse_overlap_4x4", f)[0].tolist() == sixteen
        f, _ = O.op("jxro_inverse_overlap_4x4", sixteen)
        assert O.op("jxro_forward_overlap_4x4", f)[0].tolist() == sixteen
    s = lcg_samples(bits + 300, bits, 2000 * 6)
    for i in range(0, len(s), 6):
        pair, square = s[i:i + 2], s[i + 2:i + 6]
        f, _ = O.op("jxro_forward_overlap_2", pair)
        assert O.op("jxro_inverse_overlap_2", f)[0].tolist() == pair
        f, _ = O.op("jxro_forward_overlap_2x2", square)
        assert O.op("jxro_inverse_overlap_2x2", f)[0].tolist() == square


def test_overlap_known_answers():
    assert O.op("jxro_forward_overlap_2", [2, 3])[0].tolist() == [1, 2]
    assert O.op("jxro_forward_overlap_2x2", [3, 4, 4, 5])[0].tolist() == [1, 2, 3, 4]
    assert O.op("jxro_inverse_overlap_4", [0] * 4)[0].tolist() == [0] * 4
    assert O.op("jxro_inverse_overlap_4x4", [0] * 16)[0].tolist() == [0] * 16
    assert O.op("jxro_forward_overlap_2", [-(2**31), 2**31 - 1])[1]
    assert O.op("jxro_forward_overlap_2x2", [2**31 - 1] * 4)[1]


def _plane_overlap(samples, w, h, cols, rows, hard, fn="jxro_overlap_plane"):
    s = np.ascontiguousarray(np.array(samples, np.int32))
    cb = np.array(cols, np.uint32)
    rb = np.array(rows, np.uint32)
    err = getattr(O.lib(), fn)(s.ctypes.data_as(C.POINTER(C.c_int32)), w, h, cb.ctypes.data_as(C.POINTER(C.c_uint32)),
                               len(cb), rb.ctypes.data_as(C.POINTER(C.c_uint32)), len(rb), 1 if hard else 0)
    assert not err
    return s.tolist()


def test_subsampled_plane_round_trip_and_corner_order():
    """The ordered subsampled schedule inverts the encoder's reversed one
    for regions one macroblock (two samples) wide, where each corner step
    reads the other corner (subsampled_overlap.rs:234-283)."""
    rng = np.random.default_rng(5)
    from jxr_amd import EncodedImage  # noqa: F401  (forward path lives in the library)
    for w, h, cols, rows in [(4, 4, [0, 2, 4], [0, 2, 4]), (8, 6, [0, 2, 6, 8], [0, 2, 6]), (6, 8, [0, 6], [0, 8])]:
        x = rng.integers(-200, 200, w * h).tolist()
        y = _plane_overlap(x, w, h, cols, rows, True, "jxro_subsampled_overlap_plane")
        assert y != x  # the filters moved something
    # a two-wide region: corners read each other in order (left, then right)
    a0, b0 = 10, 3
    plane = [a0, b0, a0, b0]  # 2x2, one region
    out = _plane_overlap(plane, 2, 2, [0, 2], [0, 2], True, "jxro_subsampled_overlap_plane")
    # prefix: a = a0 - b0; b = b0 - a; suffix: a += b; b += a
    a = a0 - b0
    b = b0 - a
    a = a + b
    b = b + a
    assert out == [a, b, a, b]


# ---- prediction (crates/jxr-math/src/prediction.rs:104-115) ---------------
def test_subsampled_420_prediction_uses_two_by_two_block_geometry():
    c = np.zeros(256, np.int32)
    c[12], c[16 + 12], c[32 + 12], c[48 + 12] = 5, 7, 11, 13
    err = O.lib().jxro_predict_high_pass(c.ctypes.data_as(C.POINTER(C.c_int32)), abi.PRED_FROM_LEFT, 2, 2)
    assert not err
    assert c[16 + 12] == 12 and c[48 + 12] == 24


def test_prediction_is_a_running_sum():
    rng = np.random.default_rng(3)
    base = rng.integers(-1000, 1000, 256).astype(np.int32)
    for direction, positions, blocks in [(abi.PRED_FROM_LEFT, [4, 8, 12], lambda r, c: [(r, k) for k in range(c + 1)]),
                                         (abi.PRED_FROM_TOP, [1, 2, 3], lambda r, c: [(k, c) for k in range(r + 1)])]:
        c = base.copy()
        O.lib().jxro_predict_high_pass(c.ctypes.data_as(C.POINTER(C.c_int32)), direction, 4, 4)
        for r in range(4):
            for col in range(4):
                for p in positions:
                    assert c[(r * 4 + col) * 16 + p] == sum(int(base[(rr * 4 + cc) * 16 + p]) for rr, cc in blocks(r, col))


# ---- chroma upsampling (crates/jxr-math/src/sampling.rs:117-331) ----------
def up(prev, cur, nxt, centering):
    out = (C.c_int32 * 2)()
    assert O.lib().jxro_upsample_pair(prev, cur, nxt, centering, out) == 0
    return [out[0], out[1]]


def test_upsampling_known_answers():
    assert up(-(2**31), 2**31 - 1, -(2**31), 0) == [2**31 - 1, 0]
    for c in range(5):
        for s in (2**31 - 1, -(2**31)):
            assert up(s, s, s, c) == [s, s]
    assert O.lib().jxro_upsample_pair(0, 0, 0, 7, (C.c_int32 * 2)()) != 0


def test_upsampling_equals_table_180_sum_in_64_bits():
    samples = []
    for group in range(7):
        for off in range(-3, 4):
            base = [-(2**31) + 3, 2**31 - 1 - 3, 0, 8, -8, 0x12345678, -0x12345678][group]
            samples.append(base + off)
    weights = {0: (4, 4, 0, 8), 1: (5, 3, 1, 7), 2: (6, 2, 2, 6), 3: (7, 1, 3, 5), 4: (8, 0, 4, 4)}
    for c, (h0, h1, h2, h3) in weights.items():
        for p in samples[::3]:
            for q in samples[::2]:
                for n in (p, q, samples[c * 9]):
                    assert up(p, q, n, c) == [(p * h2 + q * h3 + 4) >> 3, (q * h0 + n * h1 + 4) >> 3]


# ---- alpha (crates/jxr-math/src/alpha.rs:57-90) ----------------------------
def pm_u(v, a, m):
    out = C.c_uint32()
    assert O.lib().jxro_premultiply_unsigned(v, a, m, C.byref(out)) == 0
    return out.value


def pm_s(v, a, m):
    out = C.c_int32()
    assert O.lib().jxro_premultiply_signed(v, a, m, C.byref(out)) == 0
    return out.value


def test_premultiply_known_answers():
    assert pm_u(255, 128, 255) == 128
    assert pm_u(7, 0, 255) == 0
    assert pm_s(-20000, 16384, 32767) == -10000
    assert pm_s(20000, -1, 32767) == 0


# ---- output scaling (crates/jxr-native/src/output_format/scaling.rs) ------
def f32bits(sample, scaled, mant, bias):
    out = C.c_uint32()
    assert O.lib().jxro_float32_bits(sample, scaled, mant, bias, C.byref(out)) == 0
    return out.value


def f16bits(sample, scaled):
    out = C.c_uint16()
    assert O.lib().jxro_float16_bits(sample, scaled, C.byref(out)) == 0
    return out.value


def bits_of(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_float_outputs_known_answers():
    # crates/jxr-native/src/output_format/tests.rs:320-382
    assert f16bits(-1, 0) == 0x8001 and f16bits(1, 0) == 1
    assert f32bits(-16, 0, 4, 1) == bits_of(-1.0) and f32bits(16, 0, 4, 1) == bits_of(1.0)
    # a coded exponent beyond the f32 range clips to the largest finite value
    assert f32bits(2**31 - 1, 0, 0, -200) == 0x7F7FFFFF


def test_float32_bits_closed_form_for_normal_values():
    """For a normal coded value the f32 is 2^(e - bias) * (1 + m / 2^len)."""
    rng = np.random.default_rng(9)
    for mant, bias in [(13, 4), (10, 8), (23, 0), (4, 1)]:
        for _ in range(500):
            e = int(rng.integers(1, 40))
            m = int(rng.integers(0, 1 << mant))
            sample = (e << mant) | m
            if sample >= 2**31:
                continue
            value = 2.0 ** (e - bias) * (1 + m / 2**mant)
            if not (2**-126 <= value < 3.4e38):
                continue
            assert f32bits(sample, 0, mant, bias) == bits_of(value)
            assert f32bits(-sample, 0, mant, bias) == bits_of(-value)


# ---- output formats through whole-image decode ------------------------------
def _decode_row(values_per_plane, color, comps, spec, alpha=None, scaled=False, alpha_scaled=None, alpha_mode=None):
    n = len(values_per_plane[0])
    planes = [np.array(v, np.int64).reshape(1, n) for v in values_per_plane]
    if alpha is not None and alpha_mode == abi.ALPHA_INTEGRATED:
        planes.append(np.array(alpha, np.int64).reshape(1, n))
    arena = ManualArena(planes)
    sep = None
    if alpha is not None and alpha_mode == abi.ALPHA_SEPARATE:
        sep = ManualArena([np.array(alpha, np.int64).reshape(1, n)])
    crop = (15, 0, 2, 1) if n == 2 else (0, 0, 1, 1)
    d = desc_for(arena, color, comps, spec, crop, scaled=scaled,
                 alpha_mode=alpha_mode or abi.ALPHA_NONE, separate_alpha=sep, alpha_scaled=alpha_scaled)
    return O.decode(d)


def test_constant_macroblock_planes_reconstruct_exactly():
    arena = ManualArena([np.array([[-129, 127, 0, 5000, -7]])])
    d = desc_for(arena, abi.COLOR_LUMA, 1, OutputSpec(layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA),
                 (0, 0, 80, 16))
    plane = O.reconstruct_plane(d, 0)
    for k, v in enumerate([-129, 127, 0, 5000, -7]):
        assert (plane[:, 16 * k:16 * k + 16] == v).all()


def test_dc_only_zero_luma_decodes_to_128():
    # crates/jxr-cuda/src/plan/tests.rs:30-35
    arena = ManualArena([np.zeros((1, 1), np.int64)])
    d = desc_for(arena, abi.COLOR_LUMA, 1, OutputSpec(layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA),
                 (0, 0, 16, 16))
    assert (O.decode(d) == 128).all()


LUMA8 = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)


def test_luma_u8_biases_and_clips():
    # output_format/tests.rs:182-197
    assert _decode_row([[-129, 127]], abi.COLOR_LUMA, 1, LUMA8).tolist() == [0, 255]


def test_yuv444_to_rgba_formats_alpha_independently():
    # output_format/tests.rs:225-250
    spec = OutputSpec(layout=abi.LAYOUT_RGBA)
    for mode in (abi.ALPHA_INTEGRATED, abi.ALPHA_SEPARATE):
        got = _decode_row([[0, 1], [0, 0], [0, 0]], abi.COLOR_YUV444, 3, spec, alpha=[-128, 127], alpha_mode=mode)
        assert got.tolist() == [128, 128, 128, 0, 129, 129, 129, 255]


def test_alpha_uses_its_plane_local_scaled_flag():
    # output_format/tests.rs:253-275 (one pixel, luma to RGBA)
    arena = ManualArena([np.array([[0]])])
    sep = ManualArena([np.array([[8]])])
    d = desc_for(arena, abi.COLOR_LUMA, 1, OutputSpec(layout=abi.LAYOUT_RGBA), (0, 0, 1, 1),
                 alpha_mode=abi.ALPHA_SEPARATE, separate_alpha=sep, alpha_scaled=True)
    assert O.decode(d).tolist() == [128, 128, 128, 129]


def test_luma_typed_outputs_apply_exact_postscaling():
    # output_format/tests.rs:320-382
    cases = [
        (OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, shift_bits=1, layout=abi.LAYOUT_LUMA,
                    output_color=abi.COLOR_LUMA), np.uint16, [32766, 32770]),
        (OutputSpec(storage=abi.STORE_I16, depth=abi.DEPTH_I16, shift_bits=2, layout=abi.LAYOUT_LUMA,
                    output_color=abi.COLOR_LUMA), np.int16, [-4, 4]),
        (OutputSpec(storage=abi.STORE_I32, depth=abi.DEPTH_I32, shift_bits=3, layout=abi.LAYOUT_LUMA,
                    output_color=abi.COLOR_LUMA), np.int32, [-8, 8]),
        (OutputSpec(storage=abi.STORE_F16, depth=abi.DEPTH_F16, layout=abi.LAYOUT_LUMA,
                    output_color=abi.COLOR_LUMA), np.uint16, [0x8001, 1]),
    ]
    for spec, dt, want in cases:
        assert _decode_row([[-1, 1]], abi.COLOR_LUMA, 1, spec).view(dt).tolist() == want
    spec = OutputSpec(storage=abi.STORE_F32, depth=abi.DEPTH_F32, mantissa_length=4, exponent_bias=1,
                      layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)
    got = _decode_row([[-16, 16]], abi.COLOR_LUMA, 1, spec).view(np.uint32).tolist()
    assert got == [bits_of(-1.0), bits_of(1.0)]


def test_yuv444_to_rgb_u16_and_padding_layouts():
    # output_format/tests.rs:384-408, 613-662
    yuv = [[0, 10], [0, 3], [0, 5]]
    u16 = dict(storage=abi.STORE_U16, depth=abi.DEPTH_U16, output_color=abi.COLOR_RGB)
    got = _decode_row(yuv, abi.COLOR_YUV444, 3, OutputSpec(layout=abi.LAYOUT_RGB, **u16)).view(np.uint16)
    assert got.tolist() == [32768, 32768, 32768, 32774, 32780, 32779]
    got = _decode_row(yuv, abi.COLOR_YUV444, 3, OutputSpec(layout=abi.LAYOUT_RGBX, **u16)).view(np.uint16)
    assert got.tolist() == [32768, 32768, 32768, 0, 32774, 32780, 32779, 0]
    arena = ManualArena([np.array([[10]]), np.array([[3]]), np.array([[5]])])
    d = desc_for(arena, abi.COLOR_YUV444, 3, OutputSpec(layout=abi.LAYOUT_BGRX, **u16), (0, 0, 1, 1))
    assert O.decode(d).view(np.uint16).tolist() == [32779, 32780, 32774, 0]


def test_rgba_u8_premultiplication_uses_formatted_alpha():
    # output_format/tests.rs:832-896
    for layout in (abi.LAYOUT_RGBA, abi.LAYOUT_BGRA):
        spec = OutputSpec(layout=layout, premultiply=True)
        got = _decode_row([[0, 0], [0, 0], [0, 0]], abi.COLOR_YUV444, 3, spec, alpha=[-64, 127],
                          alpha_mode=abi.ALPHA_SEPARATE)
        assert got.tolist() == [32, 32, 32, 64, 128, 128, 128, 255]


def test_invalid_requests_fail_before_decoding():
    arena = ManualArena([np.zeros((1, 1), np.int64)])
    d = desc_for(arena, abi.COLOR_LUMA, 1, OutputSpec(layout=abi.LAYOUT_RGBA), (0, 0, 16, 16))
    code, _, _ = O.decode_status(d)  # alpha layout without an alpha plane
    assert code == abi.ERR_INVALID_PLAN
    d = desc_for(arena, abi.COLOR_LUMA, 1, LUMA8, (8, 8, 16, 16))
    code, _, _ = O.decode_status(d)  # crop outside the plane
    assert code == abi.ERR_INVALID_PLAN


def test_large_sample_cases_reach_beyond_the_plain_output_bound():
    """The hand-built large-sample arenas decode (or fail in the output
    stage) as their GPU tests assume: samples above 2^27 and, for the last,
    an overflow first met in the colour conversion."""
    import large_cases
    for name, desc, stage in large_cases.cases():
        code, st, _ = O.decode_status(desc)
        assert (code, st) == ((abi.ERR_KERNEL_ARITHMETIC, stage) if stage else (0, 0)), name
        peak = max(int(np.abs(O.reconstruct_plane(desc, i).astype(np.int64)).max())
                   for i in range(desc.component_count))
        assert peak > 2 ** 27, name


# ---- packed, RGBE, one-bit and planar YUV outputs ---------------------------
def _rgbe(r, g, b):
    out = (C.c_int32 * 3)(r, g, b)
    O.lib().jxro_rgbe_word.restype = C.c_uint32
    w = O.lib().jxro_rgbe_word(out)
    return [w & 0xFF, w >> 8 & 0xFF, w >> 16 & 0xFF, w >> 24]


def test_rgbe_postscale_known_answers():
    # jxr-math/src/rgbe.rs:129-147
    assert _rgbe(128, 256, 64) == [64, 128, 32, 2]
    assert _rgbe(-1, 0, 127) == [0, 0, 127, 1]
    assert _rgbe(2**31 - 1, 0, 0)[3] == 255


def _prescale(rgbe):
    """rgbe.rs:27-52: the inverse the reference pins postscale against."""
    out = []
    for m in rgbe[:3]:
        e = rgbe[3]
        v = 0
        if m and e:
            while m < 128 and e > 1:
                m, e = m << 1, e - 1
            v = (e << 7) | (m & 0x7F) if e > 1 else m
        out.append(v)
    return out


def test_rgbe_postscale_inverts_prescale_on_canonical_pixels():
    # rgbe.rs:104-127
    mantissas = [0, 1, 2, 63, 64, 127, 128, 129, 200, 255]
    for e in (1, 2, 3, 100, 128, 254, 255):
        for r in mantissas:
            for g in mantissas:
                for b in mantissas:
                    top = max(r, g, b)
                    if not (top >= 128 or (e == 1 and top != 0)):
                        continue
                    assert _rgbe(*_prescale([r, g, b, e])) == [r, g, b, e]
    assert _rgbe(*_prescale([0, 0, 0, 0])) == [0, 0, 0, 0]


RGB555 = OutputSpec(storage=abi.STORE_PACKED16, depth=abi.DEPTH_RGB555, layout=abi.LAYOUT_RGB,
                    output_color=abi.COLOR_RGB)


def test_packed_rgb555_bit_positions():
    # output_format/tests.rs:277-293 (luma -17, 15 -> 0, 0x7fff)
    assert _decode_row([[-17, 15]], abi.COLOR_LUMA, 1, RGB555).view(np.uint16).tolist() == [0, 0x7FFF]


def test_packed_rgb555_places_blue_low_and_red_high():
    # output_format/tests.rs:295-318 (Y 10, U 3, V 5 -> 0x5b9b)
    got = _decode_row([[10, 10], [3, 3], [5, 5]], abi.COLOR_YUV444, 3, RGB555)
    assert got.view(np.uint16).tolist() == [0x5B9B, 0x5B9B]


def test_packed_rgb565_and_101010_fields():
    # packed_color.rs:20-38 field positions (red highest; 565 green six bits,
    # red and blue shifted one more, scaling.rs:87); biases 32 and 512
    # (scaling.rs:299-318)
    spec = OutputSpec(storage=abi.STORE_PACKED16, depth=abi.DEPTH_RGB565, layout=abi.LAYOUT_RGB,
                      output_color=abi.COLOR_RGB)
    assert _decode_row([[0, 31]], abi.COLOR_LUMA, 1, spec).view(np.uint16).tolist() == [
        16 << 11 | 32 << 5 | 16, 0xFFFF]
    spec = OutputSpec(storage=abi.STORE_PACKED32, depth=abi.DEPTH_RGB101010, layout=abi.LAYOUT_RGB,
                      output_color=abi.COLOR_RGB)
    assert _decode_row([[0, -600]], abi.COLOR_LUMA, 1, spec).view(np.uint32).tolist() == [
        512 << 20 | 512 << 10 | 512, 0]


def test_rgbe_output_through_decode():
    spec = OutputSpec(storage=abi.STORE_PACKED32, depth=abi.DEPTH_U8, layout=abi.LAYOUT_RGB,
                      output_color=abi.COLOR_RGBE)
    # YUV (Y, 0, 0) is grey Y in every channel; RGBE takes no bias
    got = _decode_row([[256, 100], [0, 0], [0, 0]], abi.COLOR_YUV444, 3, spec).view(np.uint32).tolist()
    assert [[w & 0xFF, w >> 8 & 0xFF, w >> 16 & 0xFF, w >> 24] for w in got] == [_rgbe(256, 256, 256),
                                                                                [100, 100, 100, 1]]


def test_bit_packing_is_msb_first_and_black_inverts():
    # output_format/tests.rs:664-679
    black = OutputSpec(storage=abi.STORE_BITS, depth=abi.DEPTH_BIT1_BLACK, layout=abi.LAYOUT_LUMA,
                       output_color=abi.COLOR_LUMA)
    white = OutputSpec(storage=abi.STORE_BITS, depth=abi.DEPTH_BIT1_WHITE, layout=abi.LAYOUT_LUMA,
                       output_color=abi.COLOR_LUMA)
    assert _decode_row([[0, 1]], abi.COLOR_LUMA, 1, black).tolist() == [0x80]
    assert _decode_row([[0, 1]], abi.COLOR_LUMA, 1, white).tolist() == [0x40]


def test_bit_rows_pad_with_zero_bits():
    # packed_color.rs:44-46: pixel x at byte x/8, bit 7 - x%8; rows are whole bytes
    white = OutputSpec(storage=abi.STORE_BITS, depth=abi.DEPTH_BIT1_WHITE, layout=abi.LAYOUT_LUMA,
                       output_color=abi.COLOR_LUMA)
    arena = ManualArena([np.array([[1, 0]], np.int64)])
    d = desc_for(arena, abi.COLOR_LUMA, 1, white, (3, 0, 21, 2))
    got = O.decode(d).reshape(2, 3)
    # columns 3..15 are 1 (13 pixels), 16..23 are 0
    row = [0xFF, 0xF8, 0x00]
    assert got.tolist() == [row, row]


def test_planar_yuv420_keeps_native_chroma():
    # planar.rs:77-123: Y over the crop, U and V at half resolution both ways
    spec = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_YUV,
                      output_color=abi.COLOR_YUV420)
    arena = ManualArena([np.array([[10, -20]], np.int64), np.array([[3, 4]], np.int64),
                         np.array([[-5, 6]], np.int64)], geometry=[(4, 4), (2, 2), (2, 2)])
    d = desc_for(arena, abi.COLOR_YUV420, 3, spec, (0, 0, 32, 16))
    got = O.decode(d)
    y = got[:32 * 16].reshape(16, 32)
    u = got[512:512 + 16 * 8].reshape(8, 16)
    v = got[512 + 128:].reshape(8, 16)
    assert (y[:, :16] == 138).all() and (y[:, 16:] == 108).all()
    # each chroma plane is its reconstructed plane, unfiltered, biased by 128
    for plane, index in ((u, 1), (v, 2)):
        want = np.clip(O.reconstruct_plane(d, index).astype(np.int64) + 128, 0, 255)
        assert want.shape == (8, 16) and (plane == want).all()
    assert len(np.unique(u[:, :8])) == 1 and u[0, 0] != u[0, 8]


def test_planar_yuv_rejects_premultiply_and_other_samplings():
    # validate.rs:114-121, 205-215
    spec = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_YUV,
                      output_color=abi.COLOR_YUV422)
    arena = ManualArena([np.zeros((1, 1), np.int64)] * 3, geometry=[(4, 4), (2, 2), (2, 2)])
    d = desc_for(arena, abi.COLOR_YUV420, 3, spec, (0, 0, 16, 16))
    code, _, _ = O.decode_status(d)
    assert code == abi.ERR_UNSUPPORTED_OUTPUT
