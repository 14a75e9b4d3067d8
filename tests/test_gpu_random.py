"""Seeded random parity sweep on the GPU: images of random size, chroma
sampling, overlap mode, tiling (soft or hard), quality, per-macroblock
quantizer variation and output format, decoded in one persistent batch and
compared byte for byte with the CPU oracle."""
import numpy as np
import pytest

import oracle_lib as O
from cases import quality_qp
from jxr_amd import EncodedImage, HipDecodePlan, HipDecoderSession, OutputSpec, abi, image_desc, synth

pytestmark = pytest.mark.gpu

U8 = dict(storage=abi.STORE_U8, depth=abi.DEPTH_U8)
U16 = dict(storage=abi.STORE_U16, depth=abi.DEPTH_U16)
RGB_OUTPUTS = [
    dict(U8, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB),
    dict(U8, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB, planar=True),
    dict(U8, layout=abi.LAYOUT_BGRX, output_color=abi.COLOR_RGB),
    dict(U16, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB, planar=True),
    dict(storage=abi.STORE_F16, depth=abi.DEPTH_F16, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB),
    dict(storage=abi.STORE_PACKED16, depth=abi.DEPTH_RGB565, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB),
]
GRAY_OUTPUTS = [
    dict(U8, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA),
    dict(U16, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA),
    dict(U8, layout=abi.LAYOUT_RGB, output_color=abi.COLOR_RGB),
]


def split(n, parts):
    parts = max(1, min(parts, n))
    base = [n // parts] * parts
    for i in range(n % parts):
        base[i] += 1
    return base


def random_case(rng):
    w, h = int(rng.integers(1, 401)), int(rng.integers(1, 321))
    gray = rng.random() < 0.25
    color = abi.COLOR_LUMA if gray else int(rng.choice([abi.COLOR_YUV444, abi.COLOR_YUV420, abi.COLOR_YUV422]))
    depth = 16 if rng.random() < 0.3 else 8
    px = synth.gradient_noise(w, h, int(rng.integers(1 << 30)), channels=1 if gray else 3, depth=depth,
                              sigma=float(rng.uniform(2, 20)))
    mbx, mby = (w + 15) // 16, (h + 15) // 16
    tiles = (split(mbx, int(rng.integers(1, 5))), split(mby, int(rng.integers(1, 5))))
    q = int(rng.integers(40, 101))
    qp = (quality_qp(q), quality_qp(q), quality_qp(q) + (4 if q < 100 else 0))
    enc = EncodedImage(px, color=color, overlap=int(rng.integers(0, 3)), tiles=tiles,
                       hard_tiles=bool(rng.random() < 0.4), scaled=q < 100, qp=qp,
                       seed=int(rng.integers(1 << 30)), vary_quantizers=bool(rng.random() < 0.5))
    outs = GRAY_OUTPUTS if gray else RGB_OUTPUTS
    spec = OutputSpec(**outs[int(rng.integers(len(outs)))])
    if rng.random() < 0.3 and w > 4 and h > 4:
        cx, cy = int(rng.integers(0, w // 2)), int(rng.integers(0, h // 2))
        if color in (abi.COLOR_YUV420, abi.COLOR_YUV422):
            cx -= cx % 2
        if color == abi.COLOR_YUV420:
            cy -= cy % 2
        spec.crop = (cx, cy, int(rng.integers(1, w - cx + 1)), int(rng.integers(1, h - cy + 1)))
    return enc, spec


@pytest.mark.parametrize("seed", [20261204, 1, 2, 3])
def test_random_sweep_in_one_batch(seed):
    rng = np.random.default_rng(seed)
    items = [random_case(rng) for _ in range(160)]
    s = HipDecoderSession(0)
    try:
        plans = [HipDecodePlan.from_encoded(e, sp) for e, sp in items]
        bufs = [s.allocate_destination(p.output_bytes) for p in plans]
        for b in bufs:
            b.fill(0xA5)
        stages = s.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
        assert not np.any(stages)
        wants = O.decode_many([image_desc(e, sp) for e, sp in items], 16)
        for i, (b, want) in enumerate(zip(bufs, wants)):
            e, sp = items[i]
            assert np.array_equal(b.to_host(), want), (i, e.width, e.height, e.color_format, e.overlap, sp)
    finally:
        s.close()


@pytest.mark.parametrize("seed", [7, 8])
def test_random_reduced_sweep_in_one_batch(seed):
    """The same random images decoded at a quarter or a sixteenth, mixed
    with full-resolution ones, in one persistent batch."""
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(120):
        enc, spec = random_case(rng)
        spec.scale = int(rng.choice([abi.SCALE_FULL, abi.SCALE_QUARTER, abi.SCALE_SIXTEENTH]))
        items.append((enc, spec))
    s = HipDecoderSession(0)
    try:
        plans = [HipDecodePlan.from_encoded(e, sp) for e, sp in items]
        bufs = [s.allocate_destination(p.output_bytes) for p in plans]
        for b in bufs:
            b.fill(0xA5)
        stages = s.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
        assert not np.any(stages)
        wants = O.decode_many([image_desc(e, sp) for e, sp in items], 16)
        for i, (b, want) in enumerate(zip(bufs, wants)):
            e, sp = items[i]
            assert np.array_equal(b.to_host(), want), (i, e.width, e.height, e.color_format, e.overlap, sp)
    finally:
        s.close()
