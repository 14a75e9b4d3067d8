"""Soft tiles do not change the samples (docs/support.md:825-828 of the
reference: the overlap filter runs across soft tile edges, so only the
arena's order changes): an image decodes to the same bytes with and without
a tile grid, for every chroma sampling and overlap mode. Hard tiles filter
each tile on its own, so with overlap they differ. Checked on the oracle
here and on the GPU in test_gpu_parity."""
import numpy as np
import pytest

import oracle_lib as O
from cases import make
from jxr_amd import abi, image_desc

SAMPLINGS = [abi.COLOR_YUV444, abi.COLOR_YUV420, abi.COLOR_YUV422]


@pytest.mark.parametrize("color", SAMPLINGS)
@pytest.mark.parametrize("ov", [0, 1, 2])
def test_soft_tiles_do_not_change_samples(color, ov):
    base = dict(w=176, h=144, ov=ov, quality=70, color=color)
    e1, s1, _ = make(base)
    e2, s2, _ = make(dict(base, tiles=(3, 2)))
    assert np.array_equal(O.decode(image_desc(e1, s1)), O.decode(image_desc(e2, s2)))


@pytest.mark.parametrize("ov", [1, 2])
def test_hard_tiles_do(ov):
    base = dict(w=176, h=144, ov=ov, quality=70)
    e1, s1, _ = make(dict(base, tiles=(3, 2)))
    e2, s2, _ = make(dict(base, tiles=(3, 2), hard=True))
    assert not np.array_equal(O.decode(image_desc(e1, s1)), O.decode(image_desc(e2, s2)))
