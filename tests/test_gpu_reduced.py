"""Reduced-resolution decode (quarter and sixteenth) on the GPU against the
CPU oracle, byte for byte, through the C ABI: the reference's known-answer
pictures, every parity case at both reduced scales on the per-image path and
at several tile shapes, and a persistent batch mixing reduced and
full-resolution images."""
import numpy as np
import pytest

import oracle_lib as O
from arena_util import BlockArena, dc_macroblock, desc_for
from cases import CASES, make
from jxr_amd import HipDecodePlan, HipDecoderSession, JxrHipError, OutputSpec, abi, image_desc

pytestmark = pytest.mark.gpu

SCALES = [abi.SCALE_QUARTER, abi.SCALE_SIXTEENTH]
I32_LUMA = OutputSpec(storage=abi.STORE_I32, depth=abi.DEPTH_I32, layout=abi.LAYOUT_LUMA,
                      output_color=abi.COLOR_LUMA)


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


def _kat(session, mb, overlap, scale, crop):
    a = BlockArena([[mb]], [(4, 4)], 1, 1, overlap=overlap, origin=(2, 3), tiles=([3], [4]))
    d = desc_for(a, abi.COLOR_LUMA, 1, I32_LUMA, crop)
    d.scale = scale
    got = session.decode_to_host(HipDecodePlan(d, keepalive=a)).view(np.int32)
    assert np.array_equal(got, O.decode(d).view(np.int32))
    return got.tolist()


def test_reference_known_answers(session):
    """pipeline_tests.rs:47-158: a quarter stops after the first inverse
    transform; a sixteenth applies both DC gain steps, and under overlap the
    scale the flat probes measure."""
    no_hp = dc_macroblock(256, bands=abi.BANDS_NO_HIGH_PASS)
    assert _kat(session, no_hp, 2, abi.SCALE_QUARTER, (8, 12, 4, 4)) == \
        [33, 33, 33, 33, 34, 34, 34, 34, 33, 33, 33, 33, 33, 34, 34, 34]
    assert _kat(session, dc_macroblock(256), 2, abi.SCALE_SIXTEENTH, (2, 3, 1, 1)) == [33]
    assert _kat(session, dc_macroblock(256), 0, abi.SCALE_SIXTEENTH, (2, 3, 1, 1)) == [16]


def _case(name, scale):
    enc, spec, _ = make(dict(CASES[name], scale=scale))
    return enc, spec


@pytest.mark.parametrize("scale", SCALES)
@pytest.mark.parametrize("name", sorted(CASES))
def test_reduced_matches_oracle(session, name, scale):
    enc, spec = _case(name, scale)
    code, _, want = O.decode_status(image_desc(enc, spec))
    if code:
        # the one combination both refuse: subsampled planar YUV at a sixteenth
        assert code == abi.ERR_UNSUPPORTED_OUTPUT and spec.layout in (abi.LAYOUT_YUV, abi.LAYOUT_YUVA)
        with pytest.raises(JxrHipError):
            HipDecodePlan.from_encoded(enc, spec)
        return
    got = session.decode_to_host(HipDecodePlan.from_encoded(enc, spec))
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{name} scale {scale}: {bad.size} bytes differ, first at {bad[:8]}"


@pytest.mark.parametrize("tile", [(1, 1), (3, 2), (5, 3)])
def test_reduced_tile_shapes_agree(session, tile):
    names = ["rgb8_444_ov2_hard", "rgb8_420_ov2_hard_1mb", "rgb8_422_ov2_hard", "gray16_ov2_hard",
             "rgbaf_separate_ov1", "rgb8_region_crop", "gray8_bits_black_crop"]
    session.set_tile(*tile)
    try:
        for name in names:
            for scale in SCALES:
                enc, spec = _case(name, scale)
                want = O.decode(image_desc(enc, spec))
                got = session.decode_to_host(HipDecodePlan.from_encoded(enc, spec))
                assert np.array_equal(got, want), f"{name} scale {scale} at tile {tile}"
    finally:
        session.set_tile(0, 0)


def test_persistent_batch_mixes_reduced_and_full(session):
    names = ["rgb8_420_ov1", "rgb8_444_ov2_hard", "gray16_ov2_q75", "rgbaf_integrated_ov2", "rgb8_420_nchw",
             "rgb16_420_ov2"]
    encs, plans, wants, labels = [], [], [], []
    for name in names:
        for scale in [abi.SCALE_FULL] + SCALES:
            enc, spec = _case(name, scale)
            encs.append(enc)
            plans.append(HipDecodePlan.from_encoded(enc, spec))
            wants.append(O.decode(image_desc(enc, spec)))
            labels.append(f"{name} scale {scale}")
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    for b in bufs:
        b.fill(0xA5)
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert not stages.any()
    for label, b, want in zip(labels, bufs, wants):
        assert np.array_equal(b.to_host(0, want.nbytes), want), label


def test_sample_scale_terms_are_computed_once_per_configuration():
    """Two images of one configuration (block shapes, doubling, overlap,
    scale, probe regions) share the session's probe terms: the second
    upload launches no probe decode, and both decodes match the oracle."""
    s = HipDecoderSession(0)
    try:
        launches = []
        for seed in (11, 12):
            enc, spec, _ = make(dict(CASES["rgb8_420_ov1"], seed=seed, scale=abi.SCALE_QUARTER))
            plan = HipDecodePlan.from_encoded(enc, spec)
            before = s.diagnostics().kernel_launches
            s.upload(plan)
            launches.append(s.diagnostics().kernel_launches - before)
            assert np.array_equal(s.decode_to_host(plan), O.decode(image_desc(enc, spec)))
        # luma and chroma probe sets: eight probe decodes and one fit each
        assert launches == [18, 0], launches
    finally:
        s.close()
