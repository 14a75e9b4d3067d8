"""The forward analysis that makes the synthetic arenas: with every step 1
and no chroma decimation, the oracle's reconstruction returns the input
exactly (the lossless invariant of reconstruct/overlap.rs:360-403), for every
overlap mode and tile layout; quantizer steps follow the T.832 QP map
(crates/jxr-native/src/tile_decode/quantizer.rs:357-403)."""
import numpy as np
import pytest

import oracle_lib as O
from jxr_amd import EncodedImage, OutputSpec, abi, image_desc, load, synth

GRAY16 = OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)
RGB16 = OutputSpec(storage=abi.STORE_U16, depth=abi.DEPTH_U16, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB)


@pytest.mark.parametrize("ov", [0, 1, 2])
@pytest.mark.parametrize("tiles,hard", [(None, False), (([2, 1, 3], [3, 2]), False), (([2, 1, 3], [3, 2]), True),
                                         (([1] * 6, [1] * 5), True)])
def test_lossless_rgb8_round_trip(ov, tiles, hard):
    px = synth.gradient_noise(96, 80, seed=11, depth=8, sigma=20)
    enc = EncodedImage(px, color=abi.COLOR_YUV444, overlap=ov, tiles=tiles, hard_tiles=hard)
    got = O.decode(image_desc(enc, OutputSpec())).reshape(80, 96, 3)
    assert np.array_equal(got, px)


@pytest.mark.parametrize("ov", [0, 1, 2])
def test_lossless_u16_round_trips(ov):
    g = synth.pink_noise(80, 64, seed=12)
    enc = EncodedImage(g, color=abi.COLOR_LUMA, overlap=ov, tiles=([2, 3], [2, 2]), hard_tiles=ov == 2)
    assert np.array_equal(O.decode(image_desc(enc, GRAY16)).view(np.uint16).reshape(64, 80), g)
    c = synth.gradient_noise(48, 32, seed=13, depth=16, sigma=2000)
    enc = EncodedImage(c, color=abi.COLOR_YUV444, overlap=ov)
    assert np.array_equal(O.decode(image_desc(enc, RGB16)).view(np.uint16).reshape(32, 48, 3), c)


def test_odd_sizes_crop_the_coded_margin():
    px = synth.gradient_noise(37, 23, seed=14, depth=8)
    enc = EncodedImage(px, color=abi.COLOR_YUV444, overlap=2)
    assert np.array_equal(O.decode(image_desc(enc, OutputSpec())).reshape(23, 37, 3), px)


def test_subsampled_chroma_error_comes_only_from_decimation():
    """4:2:0 and 4:2:2 at step 1: the transforms and ordered subsampled
    overlap invert exactly, so every overlap mode gives the same output."""
    px = synth.gradient_noise(64, 48, seed=15, depth=8)
    for color in (abi.COLOR_YUV420, abi.COLOR_YUV422):
        outs = []
        for ov in (0, 1, 2):
            for hard in (False, True):
                enc = EncodedImage(px, color=color, overlap=ov, tiles=([1, 2, 1], [1, 2]), hard_tiles=hard)
                outs.append(O.decode(image_desc(enc, OutputSpec())))
        assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_lossy_reconstruction_is_close():
    px = synth.gradient_noise(64, 64, seed=16, depth=8)
    enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=2, scaled=True, qp=(24, 24, 28))
    got = O.decode(image_desc(enc, OutputSpec())).reshape(64, 64, 3).astype(int)
    assert np.abs(got - px).mean() < 12


def test_quant_step_known_answers():
    lib = load()
    q = lib.jxrhip_quant_step
    assert q(0, 0, 0) == 1
    assert q(31, 0, 1) == 8
    assert q(32, 0, 1) == 8
    assert q(48, 0, 1) == 16
    assert q(15, 1, 1) == 30
    assert q(16, 1, 1) == 32
    # scaled streams: the first component's DC/LP and every HP step carry one
    # more factor of two (quantizer.rs:381-403)
    assert q(5, 1, 1) == 10 and q(5, 1, 0) == 5


def test_encode_rejects_inconsistent_parameters():
    from jxr_amd import JxrHipError
    px = np.zeros((16, 16, 3), np.uint8)
    with pytest.raises(JxrHipError):
        EncodedImage(px, color=abi.COLOR_LUMA, overlap=1)
    with pytest.raises(JxrHipError):
        EncodedImage(px, color=abi.COLOR_YUV444, overlap=1, tiles=([2], [1]))


def test_arena_layout_is_component_and_tile_major():
    px = synth.gradient_noise(64, 32, seed=17, depth=8)
    enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=1, tiles=([1, 3], [2]))
    a = enc.primary.contents
    assert a.plane_count == 3
    p0, p1, p2 = a.planes[0], a.planes[1], a.planes[2]
    assert (p0.block_columns, p0.block_rows) == (4, 4) and (p1.block_columns, p1.block_rows) == (2, 2)
    assert p0.macroblock_count == p1.macroblock_count == p2.macroblock_count == 8
    assert p1.macroblock_offset == 8 and p2.macroblock_offset == 16
    # first tile (1 MB wide, 2 tall) comes first: (0,0), (0,1), then the second tile
    xs = [a.coded_x[i] for i in range(8)]
    ys = [a.coded_y[i] for i in range(8)]
    assert list(zip(xs, ys))[:3] == [(0, 0), (0, 1), (1, 0)]
    # offsets advance by the macroblock's coefficient count (all bands: 256 / 64)
    assert a.coefficient_offsets[1] - a.coefficient_offsets[0] == 256
    assert a.coefficient_offsets[9] - a.coefficient_offsets[8] == 64
