"""Plan validation on the host (no GPU): every malformed arena, window,
quantizer, prediction, crop and surface is refused by jxrhip_plan_create
before any device resource exists, with the reference's error class
(crates/jxr-cuda/src/plan.rs:255-354,405-721: spans, windows, duplicate
macroblocks, FromTopLeft prediction, crop and surface checks), and the
oracle refuses the same requests where it validates them."""
import numpy as np
import pytest

import oracle_lib as O
from arena_util import ManualArena, desc_for
from jxr_amd import HipDecodePlan, OutputSpec, abi
from jxr_amd.abi import JxrHipError

LUMA8 = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)


def luma_arena(mbx=3, mby=2, **kw):
    return ManualArena([np.arange(mbx * mby, dtype=np.int64).reshape(mby, mbx)], **kw)


def luma_desc(arena, crop=None, spec=LUMA8):
    crop = crop or (0, 0, arena.mbx * 16, arena.mby * 16)
    return desc_for(arena, abi.COLOR_LUMA, 1, spec, crop)


def refused(desc, code=abi.ERR_INVALID_PLAN):
    with pytest.raises(JxrHipError) as e:
        HipDecodePlan(desc)
    assert e.value.code == code, e.value
    return str(e.value)


def test_valid_arena_is_accepted_and_agrees_with_the_oracle():
    a = luma_arena()
    plan = HipDecodePlan(luma_desc(a))
    assert plan.output_bytes == 48 * 32
    assert O.decode_status(luma_desc(a))[0] == 0


def test_zero_quantizer_step():
    a = luma_arena()
    a.quant[4] = 0
    assert "zero quantizer step" in refused(luma_desc(a))


def test_top_left_prediction_is_refused():
    a = luma_arena()
    a.bands[:] = abi.BANDS_ALL
    a.coef = np.zeros(a.mbx * a.mby * 256, np.int32)
    a.offsets[:] = np.arange(a.mbx * a.mby, dtype=np.uint32) * 256
    a.arena.coefficients = a.coef.ctypes.data_as(abi.C.POINTER(abi.C.c_int32))
    a.arena.coefficient_count = a.coef.size
    a.planes[0].coefficient_count = a.coef.size
    a.preds[2] = abi.PRED_FROM_TOP_LEFT  # which the reference rejects
    assert "top-left" in refused(luma_desc(a))


def test_duplicate_macroblock_coordinate():
    a = luma_arena()
    a.cx[1] = a.cx[0]
    a.cy[1] = a.cy[0]
    refused(luma_desc(a))


def test_window_that_is_not_a_rectangle():
    a = luma_arena()
    a.cx[5] = 7  # one macroblock far outside the others' window
    assert "rectangular" in refused(luma_desc(a))


def test_coefficient_span_beyond_the_arena():
    a = luma_arena()
    a.offsets[3] = a.coef.size  # a DC-only macroblock needs one coefficient there
    assert "span" in refused(luma_desc(a))


def test_band_code_out_of_range():
    a = luma_arena()
    a.bands[0] = 4
    assert "band" in refused(luma_desc(a))


def test_unsupported_component_geometry():
    a = ManualArena([np.zeros((2, 2), np.int64)], geometry=[(4, 2)])
    refused(luma_desc(a))


def test_overlap_mode_out_of_range():
    a = luma_arena(overlap=3)
    refused(luma_desc(a))


def test_hard_tiles_need_a_grid_that_covers_the_window():
    a = luma_arena(hard_tiles=True, tiles=([1, 1], [2]))  # columns cover 2 of 3 macroblocks
    refused(luma_desc(a))


def test_crop_outside_the_plane_is_refused_by_both():
    a = luma_arena()
    d = luma_desc(a, crop=(40, 0, 16, 16))
    refused(d)
    assert O.decode_status(luma_desc(a, crop=(40, 0, 16, 16)))[0] == abi.ERR_INVALID_PLAN


def test_alpha_layout_without_an_alpha_plane_is_refused_by_both():
    a = luma_arena()
    spec = OutputSpec(layout=abi.LAYOUT_RGBA)
    refused(luma_desc(a, spec=spec))
    assert O.decode_status(luma_desc(a, spec=spec))[0] == abi.ERR_INVALID_PLAN


def test_surface_geometry_and_strides():
    a = luma_arena()
    d = luma_desc(a)
    d.surface.planes[0].row_stride_bytes = 47  # below a row of 48 bytes
    refused(d)
    d = luma_desc(a)
    d.surface.byte_len -= 1  # last row runs past the surface
    refused(d)
    d = luma_desc(a)
    d.surface.planes[0].width = 47
    refused(d)


def test_channel_count_must_match_the_layout():
    a = luma_arena()
    d = luma_desc(a)
    d.output.channel_count = 3
    refused(d)


def test_abi_version_mismatch():
    a = luma_arena()
    d = luma_desc(a)
    d.abi_version += 1
    refused(d, abi.ERR_INVALID_ARGUMENT)


def test_planes_must_share_one_window():
    a = ManualArena([np.zeros((2, 3), np.int64)] * 3)
    a.cx[6:12] += 1  # second plane's window shifted one macroblock right
    d = desc_for(a, abi.COLOR_YUV444, 3, OutputSpec(), (0, 0, 48, 32))
    refused(d)


def test_coefficient_plane_range_that_wraps_is_refused_by_both():
    # check_range uses a checked addition (crates/jxr-core/src/plan.rs:193-207):
    # an offset so large that offset + count wraps past 2^64 is out of range
    a = luma_arena()
    a.planes[0].coefficient_offset = (1 << 64) - 2
    a.planes[0].coefficient_count = 4
    assert "coefficient plane range" in refused(luma_desc(a))
    assert O.decode_status(luma_desc(a))[0] == abi.ERR_INVALID_PLAN
    a = luma_arena()
    a.planes[0].coefficient_count = a.coef.size + 1  # past the arena, no wrap
    assert "coefficient plane range" in refused(luma_desc(a))
    assert O.decode_status(luma_desc(a))[0] == abi.ERR_INVALID_PLAN


def test_n_component_images_up_to_the_reference_limit():
    # DecodeLimits::max_components = 64 (crates/jxr-core/src/request.rs:85);
    # more than four colour planes run as plane groups on the device
    for n, ok in ((64, True), (65, False)):
        a = ManualArena([np.zeros((1, 2), np.int64)] * n)
        spec = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_NCOMPONENT,
                          output_color=abi.COLOR_NCOMPONENT, components=n)
        d = desc_for(a, abi.COLOR_NCOMPONENT, n, spec, (0, 0, 32, 16))
        if ok:
            plan = HipDecodePlan(d, keepalive=a)
            assert plan.output_bytes == 32 * 16 * n
            assert (O.decode(d) == 128).all()
        else:
            refused(d)
            assert O.decode_status(d)[0] == abi.ERR_INVALID_PLAN


def test_reduced_scales_validate_their_crop_and_policy():
    """A reduced decode's crop is in the coded plane at its scale (3 x 2
    macroblocks: 12 x 8 at a quarter, 3 x 2 at a sixteenth); a sixteenth of
    subsampled planar YUV output and an unknown scale are refused by both."""
    from jxr_amd import EncodedImage, image_desc, reduced_crop, synth
    a = luma_arena()
    for scale, (w, h) in [(abi.SCALE_QUARTER, (12, 8)), (abi.SCALE_SIXTEENTH, (3, 2))]:
        d = luma_desc(a, crop=(0, 0, w, h))
        d.scale = scale
        d.surface = surface_for(w, h)
        assert HipDecodePlan(d).output_bytes == w * h
        assert O.decode_status(d)[0] == 0
        d = luma_desc(a, crop=(0, 0, w + 1, h))
        d.scale = scale
        d.surface = surface_for(w + 1, h)
        refused(d)
        assert O.decode_status(d)[0] == abi.ERR_INVALID_PLAN
    d = luma_desc(a)
    d.scale = 3
    refused(d)
    assert O.decode_status(d)[0] == abi.ERR_INVALID_PLAN
    enc = EncodedImage(synth.gradient_noise(64, 48, 3, depth=8), color=abi.COLOR_YUV420, overlap=1)
    yuv = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_YUV, output_color=abi.COLOR_YUV420,
                     scale=abi.SCALE_SIXTEENTH)
    refused(image_desc(enc, yuv), abi.ERR_UNSUPPORTED_OUTPUT)
    assert O.decode_status(image_desc(enc, yuv))[0] == abi.ERR_UNSUPPORTED_OUTPUT
    yuv.scale = abi.SCALE_QUARTER
    assert HipDecodePlan.from_encoded(enc, yuv).output_bytes == 16 * 12 + 2 * 8 * 6
    # native_output_region: origin divided, far edges rounded up
    assert reduced_crop((37, 21, 77, 58), abi.SCALE_QUARTER) == (9, 5, 20, 15)
    assert reduced_crop((37, 21, 77, 58), abi.SCALE_SIXTEENTH) == (2, 1, 6, 4)


def surface_for(w, h):
    from jxr_amd import surface_layout
    return surface_layout(w, h, LUMA8)
