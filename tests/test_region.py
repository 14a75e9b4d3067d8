"""Region decodes on the CPU oracle: reconstructing only the tiles a region
needs (the window `jxr_amd.prepare.RegionImage` builds, after
crates/jxr-native/src/prepare.rs:231-359) gives the same bytes as cropping a
whole-image decode, for soft and hard tiles, every overlap mode and chroma
sampling, aligned and unaligned regions."""
import copy

import numpy as np
import pytest

import oracle_lib as O
from cases import GRAY16, make
from jxr_amd import abi, image_desc
from jxr_amd.prepare import RegionImage, reconstruction_bounds, required_tiles

IMAGES = {
    "444_ov1_soft": dict(w=256, h=192, ov=1, quality=80, tiles=(4, 3)),
    "444_ov2_soft": dict(w=256, h=192, ov=2, quality=75, tiles=(4, 3), vary=True),
    "420_ov2_soft": dict(w=256, h=208, ov=2, quality=70, tiles=(3, 4), color=abi.COLOR_YUV420),
    "422_ov1_hard": dict(w=240, h=192, ov=1, quality=85, tiles=(4, 2), hard=True, color=abi.COLOR_YUV422),
    "420_ov0": dict(w=224, h=160, ov=0, quality=90, tiles=(3, 3), color=abi.COLOR_YUV420),
    "gray16_ov2": dict(w=256, h=256, ov=2, quality=75, tiles=(4, 4), image="gray16", color=abi.COLOR_LUMA,
                       out=GRAY16),
}

REGIONS = [(96, 64, 64, 48), (37, 21, 77, 58), (37, 20, 41, 33), (36, 21, 50, 40), (0, 0, 48, 32), (160, 112, 64, 48),
           (128, 80, 16, 16)]


def test_bounds_follow_the_reference_rule():
    # prepare.rs:322-359: covered macroblocks, one more each side with overlap
    # pixels 40..69 cover macroblocks 2..4
    assert reconstruction_bounds((40, 40, 30, 30), (16, 16), 1, False) == (1, 1, 6, 6)
    assert reconstruction_bounds((40, 40, 30, 30), (16, 16), 0, False) == (2, 2, 5, 5)
    assert reconstruction_bounds((40, 40, 30, 30), (16, 16), 0, True) == (1, 1, 6, 6)
    assert reconstruction_bounds((0, 0, 16, 16), (4, 4), 1, False) == (0, 0, 2, 2)
    assert reconstruction_bounds((0, 0, 256, 256), (16, 16), 2, False) == (0, 0, 16, 16)
    assert required_tiles([4, 4, 4, 4], 5, 9) == (4, 12)
    assert required_tiles([4, 4, 4, 4], 4, 8) == (4, 8)


@pytest.mark.parametrize("name", sorted(IMAGES))
@pytest.mark.parametrize("region", REGIONS)
def test_region_window_equals_cropped_whole_decode(name, region):
    case = IMAGES[name]
    enc, spec, _ = make(case)
    x, y, w, h = region
    if x + w > enc.width or y + h > enc.height:
        pytest.skip("region outside the image")
    spec = copy.copy(spec)
    spec.crop = region
    want = O.decode(image_desc(enc, spec))
    reg = RegionImage(enc, region)
    got = O.decode(image_desc(reg, spec))
    assert np.array_equal(got, want)
    mbx, mby = (enc.width + 15) // 16, (enc.height + 15) // 16
    x0, y0, x1, y1 = reg.window
    assert 0 <= x0 < x1 <= mbx and 0 <= y0 < y1 <= mby
    if region == (128, 80, 16, 16):
        assert (x1 - x0) * (y1 - y0) < mbx * mby  # a small region does not need the whole image
