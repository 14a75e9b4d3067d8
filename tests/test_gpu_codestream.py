"""Codestreams decoded on the GPU: the host reader entropy-decodes the tile
packets into arenas, the HIP reconstruction turns them into pixels, and the
bytes must equal the reference's expected pixels where its fixtures give
them (the DC-only luma codestream, the T.835-encoded image), the source
pixels for lossless codestreams (the 512x512 RGB u8 lossless config among
them), and the oracle's decode of the same arenas otherwise."""
import copy

import numpy as np
import pytest

import oracle_lib as O
import test_codestream as T
from codestream_util import GOLDEN, dc_only_luma_codestream
from jxr_amd import HipDecodePlan, HipDecoderSession, abi, workloads
from jxr_amd.codestream import Codestream, encode_codestream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


def gpu_and_oracle(session, cs, spec=None):
    spec = spec if spec is not None else cs.default_output()
    dec = cs.decode(scale=spec.scale)
    plan = HipDecodePlan.from_codestream(cs, spec, decoded=dec)
    got = session.decode_to_host(plan)
    want = O.decode(cs.image_desc(dec, spec))
    return got, want


def test_dc_only_codestream_on_gpu(session):
    cs = Codestream(dc_only_luma_codestream(False, GOLDEN["zero_dc_macroblock"]["bytes"]))
    got, want = gpu_and_oracle(session, cs)
    assert np.all(got == GOLDEN["zero_dc_macroblock"]["expected_u8"]) and np.array_equal(got, want)
    for scaled in (False, True):
        cs = Codestream(dc_only_luma_codestream(scaled, GOLDEN["nonzero_dc_macroblock"]["bytes"]))
        got, want = gpu_and_oracle(session, cs)
        assert np.array_equal(got, want), scaled


def test_reference_encoder_image_on_gpu(session):
    g = GOLDEN["scaled_three_component"]
    cs = Codestream(bytes.fromhex(g["hex"]))
    got, want = gpu_and_oracle(session, cs)
    assert np.all(got.reshape(-1, 3) == np.array(g["expected_u8_pixel"]))
    assert np.array_equal(got, want)


def test_lossless_512_rgb_config_round_trips_on_gpu(session):
    # 512x512 RGB u8 gradients + noise (sigma 8), lossless 4:4:4, one tile,
    # overlap 1, spatial packets; decoded to interleaved RGB u8
    (enc, spec), = workloads.c1()
    px = workloads.c1_pixels()
    cs = Codestream(encode_codestream(enc))
    assert cs.info.frequency_mode == 0 and cs.tile_count == 1 and cs.info.overlap == 1
    got, want = gpu_and_oracle(session, cs, spec)
    assert np.array_equal(got, want)
    assert np.array_equal(got.reshape(px.shape), px)


@pytest.mark.parametrize("case", T.CASES, ids=[c[0] for c in T.CASES])
def test_writer_cases_on_gpu(session, case):
    px, enc = T.make(case)
    cs = Codestream(encode_codestream(enc, frequency=case[8], annex_a=case[9] == abi.ALPHA_SEPARATE))
    got, want = gpu_and_oracle(session, cs)
    assert np.array_equal(got, want), case[0]
    if case[5] == (0, 0, 0):
        assert np.array_equal(got.view(px.dtype).reshape(px.shape), px), case[0]


@pytest.mark.parametrize("scale", [abi.SCALE_QUARTER, abi.SCALE_SIXTEENTH])
def test_reduced_decodes_of_frequency_codestreams_on_gpu(session, scale):
    for case in (T.CASES[2], T.CASES[4], T.CASES[9]):
        px, enc = T.make(case)
        cs = Codestream(encode_codestream(enc, frequency=True))
        spec = copy.copy(cs.default_output())
        spec.scale = scale
        got, want = gpu_and_oracle(session, cs, spec)
        assert np.array_equal(got, want), (case[0], scale)


def test_region_of_a_tiled_codestream_on_gpu(session):
    case = T.CASES[3][:7] + (([3, 4, 3], [2, 4, 3]),) + T.CASES[3][8:]
    px, enc = T.make(case, width=160, height=144)
    cs = Codestream(encode_codestream(enc))
    assert cs.tile_count > 1
    whole = O.decode(cs.image_desc(cs.decode(), cs.default_output())).reshape(144, 160, 3)
    for region in ((0, 0, 48, 32), (37, 21, 64, 50), (100, 90, 60, 54)):
        spec = copy.copy(cs.default_output())
        spec.crop = region
        plan = HipDecodePlan.from_codestream(cs, spec)
        got = session.decode_to_host(plan)
        x, y, w, h = region
        assert np.array_equal(got.reshape(h, w, 3), whole[y:y + h, x:x + w]), region


def test_codestream_batch_on_gpu(session):
    plans, wants = [], []
    for case in T.CASES[:6]:
        px, enc = T.make(case)
        cs = Codestream(encode_codestream(enc, frequency=case[8]))
        dec = cs.decode()
        spec = cs.default_output(planar=True)
        plans.append(HipDecodePlan.from_codestream(cs, spec, decoded=dec))
        wants.append(O.decode(cs.image_desc(dec, spec)))
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert not stages.any()
    for b, want in zip(bufs, wants):
        assert np.array_equal(b.to_host(0, want.nbytes), want)


def test_decode_into_pinned_host_blocks_on_gpu(session):
    # coefficients entropy-decoded straight into page-locked host blocks of
    # the session, uploaded from there by DMA; the block is reused for a
    # second codestream once the first submission completed
    cases = [T.CASES[3], T.CASES[10]]
    streams = [Codestream(encode_codestream(T.make(c)[1], frequency=c[8], annex_a=c[9] == abi.ALPHA_SEPARATE))
               for c in cases]
    need = max(sum(cs.coefficient_count()) for cs in streams) * 4
    block = session.allocate_host(need)
    for cs in streams:
        dec = cs.decode(into=block)
        spec = cs.default_output()
        plan = HipDecodePlan.from_codestream(cs, spec, decoded=dec)
        got = session.decode_to_host(plan)
        want = O.decode(cs.image_desc(cs.decode(), spec))
        assert np.array_equal(got, want)
        del plan, dec
    block.close()
    from jxr_amd.abi import load
    import ctypes as C
    assert load().jxrhip_host_free(session.handle, C.c_void_p(16)) == abi.ERR_INVALID_ARGUMENT


def arena_path_and_codestream_path(session, enc, spec, **writer):
    """Device output of the encoder's own arenas and of the same arenas
    written to a codestream and entropy-decoded again, into pinned storage."""
    cs = Codestream(encode_codestream(enc, **writer))
    block = session.allocate_host(4 * sum(cs.coefficient_count(spec.scale)))
    dec = cs.decode(scale=spec.scale, into=block)
    got = session.decode_to_host(HipDecodePlan.from_codestream(cs, spec, decoded=dec))
    want = session.decode_to_host(HipDecodePlan.from_encoded(enc, spec))
    block.close()
    return got, want


@pytest.mark.parametrize("frequency", [False, True])
def test_4096_gray_u16_config_from_codestream_on_gpu(session, frequency):
    # 4096x4096 gray u16 pink noise, 8x8 soft tiles of 512 px, overlap 2,
    # quality 75, all bands -> u16: from its codestream (spatial or
    # frequency packets, 64 tiles on host threads) the device output equals
    # that of the encoder's arenas, byte for byte
    (enc, spec), = workloads.c2()
    got, want = arena_path_and_codestream_path(session, enc, spec, frequency=frequency)
    assert got.nbytes == 4096 * 4096 * 2 and np.array_equal(got, want)


def test_1024_rgba_f32_config_from_codestream_on_gpu(session):
    # 1024x1024 RGBA f32 lossless, separate alpha (Annex-A container),
    # overlap 1 -> planar f32 + alpha: two images of the batch of 128
    for enc, spec in workloads.c4(count=2):
        got, want = arena_path_and_codestream_path(session, enc, spec, annex_a=True)
        assert got.nbytes == 1024 * 1024 * 16 and np.array_equal(got, want)
