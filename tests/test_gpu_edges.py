"""Edge shapes on the GPU, bit-exact against the CPU oracle: empty batches in
both modes; images of one pixel, one macroblock, one macroblock wide or
tall, and sizes that are not whole macroblocks, in 4:4:4, 4:2:0 and 4:2:2
at every overlap mode, one at a time and all together in one persistent
batch (many one-tile images sharing a launch)."""
import numpy as np
import pytest

import oracle_lib as O
from jxr_amd import EncodedImage, HipDecodePlan, HipDecoderSession, abi, image_desc, synth, workloads

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (16, 16), (17, 9), (16, 272), (272, 16), (33, 47)]
COLORS = [abi.COLOR_YUV444, abi.COLOR_YUV420, abi.COLOR_YUV422]


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


@pytest.fixture(scope="module")
def corpus():
    out = []
    for w, h in SIZES:
        for color in COLORS:
            for ov in (0, 1, 2):
                px = synth.gradient_noise(w, h, w * 131 + h, depth=8, sigma=8.0)
                enc = EncodedImage(px, color=color, overlap=ov, scaled=True, qp=(20, 20, 24))
                spec = workloads.rgb_out(8, False)
                out.append(((w, h, color, ov), enc, spec, O.decode(image_desc(enc, spec))))
    return out


@pytest.mark.parametrize("mode", [abi.BATCH_PERSISTENT, abi.BATCH_PER_IMAGE])
def test_empty_batch_completes_with_no_images(session, mode):
    assert list(session.submit_batch([], [], mode).wait()) == []


def test_edge_shapes_one_at_a_time(session, corpus):
    for name, enc, spec, want in corpus:
        got = session.decode_to_host(HipDecodePlan.from_encoded(enc, spec))
        assert np.array_equal(got, want), name


def test_edge_shapes_in_one_batch(session, corpus):
    plans = [HipDecodePlan.from_encoded(enc, spec) for _, enc, spec, _ in corpus]
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    for b in bufs:
        b.fill(0x3C)
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert not np.any(stages)
    for (name, _, _, want), b in zip(corpus, bufs):
        assert np.array_equal(b.to_host(), want), name


def test_surface_planes_beyond_4_gib(session):
    """Output planes 2.5 and 5 GiB into the surface: every offset the output
    stage forms is 64-bit, and nothing lands outside the planes."""
    px = synth.gradient_noise(96, 64, 5, depth=16, sigma=8.0)
    enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=2, scaled=True, qp=(120, 120, 124))
    spec = workloads.rgb_out(16, True)
    want = O.decode(image_desc(enc, spec))
    plane = 96 * 64 * 2
    offsets = [0, (5 << 29) + 4096, (5 << 30) + 512]
    desc = image_desc(enc, spec)
    for c, off in enumerate(offsets):
        desc.surface.planes[c].byte_offset = off
    desc.surface.byte_len = offsets[-1] + plane + 4096
    plan = HipDecodePlan(desc, keepalive=enc)
    dst = session.allocate_destination(desc.surface.byte_len)
    try:
        dst.fill(0x3C)
        stage = session.submit_into(plan, dst.ptr, dst.nbytes).wait()
        assert not stage
        for c, off in enumerate(offsets):
            got = dst.to_host(off, plane)
            assert np.array_equal(got, want[c * plane:(c + 1) * plane]), f"plane {c}"
            after = dst.to_host(off + plane, 4096)
            assert np.all(after == 0x3C), f"bytes after plane {c}"
            if off:
                before = dst.to_host(off - 4096, 4096)
                assert np.all(before == 0x3C), f"bytes before plane {c}"
    finally:
        dst.close()
