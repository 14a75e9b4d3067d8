"""The codestream reader and writer on the CPU: the reference's codestream
fixtures (headers, the DC-only luma codestream, a T.835-encoded image and
its expected pixels), the reference's entropy vectors placed in a crafted
packet, round trips through the writer for every layout the encoder makes,
and the failures a malformed or unsupported codestream gives.

Pixels come from the oracle (the checker) decoding what the reader made; the
GPU half of the same comparisons is tests/test_gpu_codestream.py."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from codestream_util import (GOLDEN, arena_arrays, arena_differences, bits_to_bytes, codestream,
                             dc_only_luma_codestream, minimal_header)
from jxr_amd import EncodedImage, JxrHipError, abi, synth
from jxr_amd.abi import load
from jxr_amd.codestream import Codestream, encode_codestream


# ---- headers (crates/jxr-native/tests/header_parse.rs) ---------------------
def test_parses_minimal_image_and_plane_headers():
    cs = Codestream(minimal_header() + bytes([0xFD]))
    i = cs.info
    assert (i.width, i.height, i.component_count) == (16, 16, 1)
    assert i.alpha_mode == abi.ALPHA_NONE
    assert i.color_format == abi.COLOR_LUMA and i.output_color == abi.COLOR_LUMA and i.bit_depth == abi.DEPTH_U8
    assert i.bands == abi.BANDS_ALL and i.short_header == 1 and i.frequency_mode == 0
    assert (i.macroblocks_x, i.macroblocks_y, i.tile_column_count, i.tile_row_count) == (1, 1, 1, 1)
    # no PROFILE_LEVEL_INFO: the advanced profile at level 255 (metadata.rs profile_and_level)
    assert (i.profile_idc, i.level_idc) == (111, 255)


def test_rejects_bad_signature():
    data = bytearray(minimal_header() + bytes([0xFD]))
    data[0] = ord("X")
    with pytest.raises(JxrHipError) as e:
        Codestream(bytes(data))
    assert e.value.code == abi.ERR_SYNTAX


def test_rejects_reserved_overlap_mode():
    data = bytearray(minimal_header() + bytes([0xFD]))
    data[9] |= 0b11
    with pytest.raises(JxrHipError) as e:
        Codestream(bytes(data))
    assert e.value.code == abi.ERR_SYNTAX and "OVERLAP_MODE" in e.value.message


def test_truncation_is_a_syntax_error():
    with pytest.raises(JxrHipError) as e:
        Codestream(minimal_header()[:10])
    assert e.value.code == abi.ERR_SYNTAX and "truncated" in e.value.message


# ---- the DC-only luma codestream (crates/jxr/tests/common/mod.rs:39-46) ----
def luma_u8(cs, dec):
    return O.decode(cs.image_desc(dec, cs.default_output()))


def test_dc_only_codestream_decodes_to_mid_gray():
    cs = Codestream(dc_only_luma_codestream(False, GOLDEN["zero_dc_macroblock"]["bytes"]))
    dec = cs.decode()
    a = arena_arrays(dec.primary)
    assert a["coefficients"].tolist() == [0] and a["bands"].tolist() == [abi.BANDS_DC_ONLY]
    assert a["quantizers"].tolist() == [[1, 1, 1]]
    out = luma_u8(cs, dec)
    assert out.size == 256 and np.all(out == GOLDEN["zero_dc_macroblock"]["expected_u8"])


def test_dc_only_codestream_honours_its_scaled_flag():
    # NONZERO_DC_MACROBLOCK: no ABS_LEVEL flag, eight model bits 00011001
    # (25), sign 1: a DC of -25. A scaled plane doubles the luma step of QP 1
    # (quantizer.rs quant_map: mantissa 1, exponent 1).
    mb = GOLDEN["nonzero_dc_macroblock"]["bytes"]
    outs = []
    for scaled, step in ((False, 1), (True, 2)):
        cs = Codestream(dc_only_luma_codestream(scaled, mb))
        dec = cs.decode()
        a = arena_arrays(dec.primary)
        assert a["coefficients"].tolist() == [-25]
        assert a["quantizers"].tolist() == [[step, step, step]]
        outs.append(luma_u8(cs, dec))
    # the fixture makes the scaled flag observable (public_parse.rs:185-200)
    assert not np.array_equal(outs[0], outs[1])


# ---- a file the T.835 reference encoder wrote --------------------------------
def test_reference_encoder_image_decodes_to_its_expected_pixels():
    g = GOLDEN["scaled_three_component"]
    cs = Codestream(bytes.fromhex(g["hex"]))
    i = cs.info
    assert i.annex_a == 1 and (i.width, i.height) == (g["width"], g["height"])
    assert i.color_format == abi.COLOR_NCOMPONENT and i.component_count == g["components"]
    assert i.scaled == 1 and i.output_color == abi.COLOR_NCOMPONENT
    dec = cs.decode()
    out = O.decode(cs.image_desc(dec, cs.default_output())).reshape(-1, 3)
    assert out.shape == (256, 3)
    assert np.all(out == np.array(g["expected_u8_pixel"]))


# ---- the reference's entropy vectors in one crafted packet -------------------
def crafted_packet(bands_code: int, body: str) -> bytes:
    """16x16 Y-only image with image-level quantizers of one for the bands
    `bands_code` codes, and one tile packet whose macroblock bits are `body`."""
    fields = [(0, 3), (0, 1), (bands_code, 4), (1, 1), (1, 8)]
    if bands_code <= 2:
        fields += [(0, 1), (1, 1), (1, 8)]
    if bands_code <= 1:
        fields += [(0, 1), (1, 1), (1, 8)]
    return codestream(0x01, fields, bytes([0xFD, 0, 0, 1, 0x5A]) + bits_to_bytes(body))


DC_ZERO = "0" + "0" * 8          # no ABS_LEVEL, eight zero model bits: DC 0
LP_REFINE_ZERO = "0000" * 15     # four LP model bits of each position, all zero


def test_lp_vector_places_minus_one_at_position_four():
    v = GOLDEN["entropy_vectors"]["lp_single_minus_one"]["bits"]
    # DC; LP flag 1, the block, refinement (the coded part -1 carries the sign)
    cs = Codestream(crafted_packet(2, DC_ZERO + "1" + v + LP_REFINE_ZERO))
    a = arena_arrays(cs.decode().primary)
    low = a["coefficients"]
    assert low.size == 16 and low[4] == -1 << 4 and np.count_nonzero(low) == 1


def test_hp_vector_and_initial_pattern_prediction():
    ev = GOLDEN["entropy_vectors"]
    lp = ev["lp_single_minus_one"]["bits"]
    hp = ev["hp_two_coefficients"]["bits"]
    # The LP -1 at position 4 (energy down the luma) makes the HP direction
    # "from the left", whose scan is the horizontal one (initial order 0, 4,
    # 1, 5, ...). CBPHP: NUM_CBPHP 1 (no group), which the initial predictor
    # turns into all sixteen blocks coded. Block 0 carries the reference's
    # two-coefficient vector (scan indices 2 and 3: positions 1 and 5); every
    # other block one +1 at scan index 1 (position 4). HP model bits start
    # at zero: no flexbits.
    assert ev["cbphp_initial_prediction"]["predicted"] == 0xFFFF
    body = DC_ZERO + "1" + lp + LP_REFINE_ZERO + "1" + hp + "000100" * 15
    cs = Codestream(crafted_packet(0, body))
    a = arena_arrays(cs.decode().primary)
    blocks = a["coefficients"].reshape(16, 16)
    assert a["hp_predictions"].tolist() == [abi.PRED_FROM_LEFT]
    want = np.zeros((16, 16), np.int64)
    want[4, 0] = -16  # the LP coefficient, block 4's DC/LP value
    want[0, 1], want[0, 5] = -2, 1
    want[1:, 4] = 1
    assert np.array_equal(blocks, want)


# ---- round trips through the writer ------------------------------------------
CASES = [
    # (name, colour, channels, dtype, overlap, qp, bands, tiles, frequency, alpha)
    ("gray8_lossless", abi.COLOR_LUMA, 1, np.uint8, 1, (0, 0, 0), abi.BANDS_ALL, None, False, None),
    ("gray16_q_tiles_ov2", abi.COLOR_LUMA, 1, np.uint16, 2, (20, 24, 30), abi.BANDS_ALL, ([2, 3], [1, 2, 2]), False,
     None),
    ("rgb444_lossless_freq", abi.COLOR_YUV444, 3, np.uint8, 1, (0, 0, 0), abi.BANDS_ALL, None, True, None),
    ("rgb420_q_tiles", abi.COLOR_YUV420, 3, np.uint8, 2, (30, 36, 40), abi.BANDS_ALL, ([3, 2], [5]), False, None),
    ("rgb422_q16_freq_tiles", abi.COLOR_YUV422, 3, np.uint16, 1, (10, 14, 20), abi.BANDS_ALL, ([1, 4], [2, 3]), True,
     None),
    ("rgb420_noflex", abi.COLOR_YUV420, 3, np.uint8, 1, (20, 20, 20), abi.BANDS_NO_FLEXBITS, None, False, None),
    ("rgb444_nohp_freq", abi.COLOR_YUV444, 3, np.uint8, 0, (30, 30, 30), abi.BANDS_NO_HIGH_PASS, None, True, None),
    ("gray_dconly", abi.COLOR_LUMA, 1, np.uint8, 2, (30, 30, 30), abi.BANDS_DC_ONLY, None, False, None),
    ("rgba_integrated", abi.COLOR_YUV444, 4, np.uint8, 1, (16, 20, 24), abi.BANDS_ALL, ([2, 3], [5]), False,
     abi.ALPHA_INTEGRATED),
    ("graya16_integrated_freq", abi.COLOR_LUMA, 2, np.uint16, 2, (0, 0, 0), abi.BANDS_ALL, None, True,
     abi.ALPHA_INTEGRATED),
    ("rgba_separate", abi.COLOR_YUV420, 4, np.uint8, 1, (24, 28, 30), abi.BANDS_ALL, None, False, abi.ALPHA_SEPARATE),
]


def make(case, width=80, height=72, seed=3):
    name, color, ch, dtype, ov, qp, bands, tiles, freq, alpha = case
    px = synth.gradient_noise(width, height, seed=seed, depth=8 if dtype == np.uint8 else 16).astype(dtype)
    if ch == 1:
        px = px[..., 0]
    elif ch == 2:
        px = np.dstack([px[..., 0], px[..., 2]])
    elif ch == 4:
        px = np.dstack([px, px[..., 1:2] // 2 + px[..., 0:1] // 3])
    kw = dict(color=color, overlap=ov, qp=qp, scaled=any(qp), tiles=tiles, bands=bands, seed=seed,
              vary_quantizers=bool(qp[0]))
    if alpha is not None:
        kw.update(alpha_mode=alpha, alpha_qp=qp[1])
    return px, EncodedImage(px, **kw)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_round_trip_reproduces_the_arenas(case):
    px, enc = make(case)
    bands, freq, alpha = case[6], case[8], case[9]
    data = encode_codestream(enc, frequency=freq, annex_a=alpha == abi.ALPHA_SEPARATE)
    cs = Codestream(data)
    assert cs.info.frequency_mode == int(freq) and cs.info.bands == bands
    assert cs.info.alpha_mode == (alpha if alpha is not None else abi.ALPHA_NONE)
    dec = cs.decode(threads=4)
    want, got = arena_arrays(enc.primary), arena_arrays(dec.primary)
    if bands == abi.BANDS_NO_FLEXBITS:
        # without flexbits a decoder keeps each high-pass coefficient's coded
        # part only: the low bits the model leaves are zero
        diff = set(arena_differences(want, got, bands))
        assert diff <= {"coefficients"}, diff
        w, g = want["coefficients"], got["coefficients"]
        assert np.all(np.abs(g) <= np.abs(w)) and np.all((g == 0) | (np.sign(g) == np.sign(w)))
        # a second pass through the codec changes nothing
        again = Codestream(data).decode()
        assert not arena_differences(got, arena_arrays(again.primary))
    else:
        assert not arena_differences(want, got, bands)
    if alpha == abi.ALPHA_SEPARATE:
        assert not arena_differences(arena_arrays(enc.alpha), arena_arrays(dec.alpha), bands)


@pytest.mark.parametrize("case", [c for c in CASES if c[5] == (0, 0, 0)], ids=lambda c: c[0])
def test_lossless_round_trip_returns_the_source_pixels(case):
    px, enc = make(case)
    cs = Codestream(encode_codestream(enc, frequency=case[8]))
    dec = cs.decode()
    spec = cs.default_output()
    out = O.decode(cs.image_desc(dec, spec)).view(px.dtype).reshape(px.shape)
    assert np.array_equal(out, px)


def test_trimmed_flexbits_drop_only_low_bits():
    px, enc = make(CASES[3])
    full = Codestream(encode_codestream(enc)).decode()
    for trim in (0, 2, 4):
        data = encode_codestream(enc, trim=trim)
        cs = Codestream(data)
        assert cs.info.trim_flexbits_present == 1
        got = arena_arrays(cs.decode().primary)["coefficients"]
        want = arena_arrays(full.primary)["coefficients"]
        assert np.all(np.abs(want - got) < (1 << trim))
        if trim == 0:
            assert np.array_equal(got, want)


def test_tile_selection_decodes_only_the_selected_tiles():
    px, enc = make(CASES[1])
    cs = Codestream(encode_codestream(enc))
    full = arena_arrays(cs.decode().primary)
    mask = [False, True, False, False, True, False]
    part = arena_arrays(cs.decode(tiles=mask).primary)
    # columns of 2 and 3 macroblocks, rows of 1, 2 and 2: tile 1 is columns
    # 2..4 of row 0, tile 4 columns 0..1 of rows 3..4
    keep = ((full["coded_x"] >= 2) & (full["coded_y"] == 0)) | ((full["coded_x"] < 2) & (full["coded_y"] >= 3))
    assert np.array_equal(part["coded_x"], full["coded_x"][keep])
    assert np.array_equal(part["coded_y"], full["coded_y"][keep])
    for m_part, m_full in zip(range(part["offsets"].size), np.nonzero(keep)[0]):
        n = 256
        assert np.array_equal(part["coefficients"][part["offsets"][m_part]:part["offsets"][m_part] + n],
                              full["coefficients"][full["offsets"][m_full]:full["offsets"][m_full] + n])
    with pytest.raises(JxrHipError) as e:
        cs.decode(tiles=[False] * 6)
    assert e.value.code == abi.ERR_SYNTAX


def test_threads_do_not_change_the_arena():
    case = CASES[4][:7] + (([2, 3, 5], [3, 5]),) + CASES[4][8:]
    px, enc = make(case, width=160, height=128)
    cs = Codestream(encode_codestream(enc, frequency=True))
    one = arena_arrays(cs.decode(threads=1).primary)
    many = arena_arrays(cs.decode(threads=8).primary)
    assert not arena_differences(one, many)


class NumpyStorage:
    """Caller storage for Codestream.decode(into=...): pageable numpy memory
    (the ABI takes any host memory; pinned blocks only make uploads DMA)."""

    def __init__(self, values):
        self.array = np.full(values, 0x5A5A5A5A, np.int32)
        self.ptr = self.array.ctypes.data
        self.nbytes = self.array.nbytes


@pytest.mark.parametrize("case", [CASES[1], CASES[4], CASES[8], CASES[10]], ids=lambda c: c[0])
def test_decode_into_caller_storage_equals_the_owned_arena(case):
    px, enc = make(case)
    cs = Codestream(encode_codestream(enc, frequency=case[8], annex_a=case[9] == abi.ALPHA_SEPARATE))
    owned = cs.decode()
    n_primary, n_alpha = cs.coefficient_count()
    assert n_primary == owned.primary.contents.coefficient_count
    assert n_alpha == (owned.alpha.contents.coefficient_count if owned.alpha else 0)
    store = NumpyStorage(n_primary + n_alpha + 7)
    into = cs.decode(threads=3, into=store)
    got = arena_arrays(into.primary)
    assert not arena_differences(arena_arrays(owned.primary), got)
    # the arena's coefficients are the caller's storage, written in place
    assert C.addressof(into.primary.contents.coefficients.contents) == store.ptr
    assert np.array_equal(store.array[:n_primary], got["coefficients"])
    if n_alpha:
        assert not arena_differences(arena_arrays(owned.alpha), arena_arrays(into.alpha))
        assert C.addressof(into.alpha.contents.coefficients.contents) == store.ptr + 4 * n_primary
    assert np.all(store.array[n_primary + n_alpha:] == 0x5A5A5A5A)


def test_decode_into_counts_selected_tiles_and_reduced_scales():
    px, enc = make(CASES[4])
    cs = Codestream(encode_codestream(enc, frequency=True))
    mask = [True, False, False, True]
    for scale in (abi.SCALE_FULL, abi.SCALE_QUARTER, abi.SCALE_SIXTEENTH):
        owned = cs.decode(scale=scale, tiles=mask)
        n, na = cs.coefficient_count(scale, mask)
        assert na == 0 and n == owned.primary.contents.coefficient_count
        store = NumpyStorage(n)
        into = cs.decode(scale=scale, tiles=mask, into=store)
        assert not arena_differences(arena_arrays(owned.primary), arena_arrays(into.primary))


def test_decode_into_rejects_storage_below_the_arena():
    px, enc = make(CASES[0])
    cs = Codestream(encode_codestream(enc))
    n, _ = cs.coefficient_count()
    lib = load()
    store = NumpyStorage(n)
    h = C.c_void_p()
    assert lib.jxrhip_codestream_decode_into(cs._h, abi.SCALE_FULL, None, 1, C.c_void_p(store.ptr), n - 1, None, 0,
                                             C.byref(h)) == abi.ERR_BUFFER_TOO_SMALL
    assert not h.value
    assert lib.jxrhip_codestream_decode_into(cs._h, abi.SCALE_FULL, None, 1, None, n, None, 0,
                                             C.byref(h)) == abi.ERR_INVALID_ARGUMENT
    with pytest.raises(ValueError):
        cs.decode(into=NumpyStorage(n - 1))


def test_reduced_decodes_keep_the_dc_and_lp_bands():
    px, enc = make(CASES[4])
    cs = Codestream(encode_codestream(enc, frequency=True))
    full = arena_arrays(cs.decode().primary)
    for scale, bands in ((abi.SCALE_QUARTER, abi.BANDS_NO_HIGH_PASS), (abi.SCALE_SIXTEENTH, abi.BANDS_DC_ONLY)):
        red = arena_arrays(cs.decode(scale=scale).primary)
        assert set(red["bands"].tolist()) == {bands}
        for (fo, _, fm, fn, bc, br), (ro, _, rm, rn, _, _) in zip(full["planes"], red["planes"]):
            blocks = bc * br
            for k in range(fn):
                src = full["coefficients"][full["offsets"][fm + k]:][: blocks * 16: 16]
                dst = red["coefficients"][red["offsets"][rm + k]:]
                n = blocks if bands == abi.BANDS_NO_HIGH_PASS else 1
                assert np.array_equal(dst[:n], src[:n])


def test_reduced_decodes_need_frequency_packets_and_the_lp_band():
    px, enc = make(CASES[1])
    spatial = Codestream(encode_codestream(enc))
    with pytest.raises(JxrHipError) as e:
        spatial.decode(scale=abi.SCALE_QUARTER)
    assert e.value.code == abi.ERR_UNSUPPORTED
    px, enc = make(CASES[7])
    dc_only = Codestream(encode_codestream(enc, frequency=True))
    with pytest.raises(JxrHipError) as e:
        dc_only.decode(scale=abi.SCALE_QUARTER)
    assert e.value.code == abi.ERR_UNSUPPORTED
    assert dc_only.decode(scale=abi.SCALE_SIXTEENTH) is not None


def test_truncated_packets_fail_with_a_syntax_error():
    px, enc = make(CASES[0])
    data = encode_codestream(enc)
    for cut in (len(data) - 1, len(data) // 2, len(data) - 40):
        cs = Codestream(data[:cut])
        with pytest.raises(JxrHipError) as e:
            cs.decode()
        assert e.value.code == abi.ERR_SYNTAX


def test_separate_alpha_needs_a_container():
    px, enc = make(CASES[10])
    with pytest.raises(JxrHipError) as e:
        encode_codestream(enc)
    assert e.value.code == abi.ERR_UNSUPPORTED
    cs = Codestream(encode_codestream(enc, annex_a=True))
    i = cs.info
    assert i.annex_a == 1 and i.alpha_mode == abi.ALPHA_SEPARATE and i.alpha_bytes > 0
    assert bytes(i.pixel_format_guid)[:4] == bytes([0x24, 0xC3, 0xDD, 0x6F])
