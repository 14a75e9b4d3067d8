"""Region decodes on the GPU: the window of tiles a region needs
(`jxr_amd.prepare.RegionImage`, after crates/jxr-native/src/prepare.rs:
231-359) decodes to the same bytes as the oracle's decode of the whole
image cropped to the region, for soft and hard tiles, every overlap mode and
chroma sampling, aligned and unaligned regions (odd origins on subsampled
chroma included)."""
import copy

import numpy as np
import pytest

import oracle_lib as O
from cases import make
from jxr_amd import HipDecodePlan, HipDecoderSession, abi, image_desc
from jxr_amd.prepare import RegionImage
from test_region import IMAGES, REGIONS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


@pytest.mark.parametrize("name", sorted(IMAGES))
def test_region_windows_on_gpu(session, name):
    case = IMAGES[name]
    enc, spec, _ = make(case)
    ran = 0
    for region in REGIONS:
        x, y, w, h = region
        if x + w > enc.width or y + h > enc.height:
            continue
        sp = copy.copy(spec)
        sp.crop = region
        want = O.decode(image_desc(enc, sp))
        reg = RegionImage(enc, region)
        got = session.decode_to_host(HipDecodePlan(image_desc(reg, sp), keepalive=reg))
        assert np.array_equal(got, want), (name, region)
        ran += 1
    assert ran >= 3
