"""The benchmark configurations at their full sizes on the GPU, bit-exact
against the CPU oracle:

  C1  512x512 RGB u8 lossless 4:4:4 ov1: also equal to the source pixels
      (the lossless round trip);
  C2  4096x4096 gray u16 pink noise, 8x8 soft tiles, ov2, q75, flexbits;
  C3  all 256 mixed 4:2:0 images as one persistent batch into one dense NCHW
      allocation, replayed from a captured graph;
  C4  two of the 1024x1024 RGBA f32 lossless images with separate alpha.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O
from jxr_amd import HipDecodePlan, HipDecoderSession, abi, image_desc, synth, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


def test_c1_full_size_lossless(session):
    (enc, spec), = workloads.c1()
    plan = HipDecodePlan.from_encoded(enc, spec)
    got = session.decode_to_host(plan)
    assert np.array_equal(got, O.decode(image_desc(enc, spec)))
    px = synth.gradient_noise(512, 512, workloads.SEED, depth=8, sigma=8.0)
    assert np.array_equal(got.reshape(512, 512, 3), px)


def test_c2_full_size(session):
    (enc, spec), = workloads.c2()
    plan = HipDecodePlan.from_encoded(enc, spec)
    got = session.decode_to_host(plan)
    want = O.decode(image_desc(enc, spec))
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8]}"


def test_c3_full_batch_in_one_dense_allocation(session):
    items = workloads.c3()
    plans = [HipDecodePlan.from_encoded(e, sp) for e, sp in items]
    offsets, total = [], 0
    for p in plans:
        offsets.append(total)
        total += (p.output_bytes + 255) // 256 * 256
    dense = session.allocate_destination(total)
    dense.fill(0x77)
    graph = session.capture_batch(plans, [dense.ptr + o for o in offsets])
    try:
        stages = graph.launch().wait()
    finally:
        graph.close()
    assert not np.any(stages)
    host = dense.to_host()
    with ThreadPoolExecutor(16) as ex:
        wants = list(ex.map(lambda it: O.decode(image_desc(*it)), items))
    for i, (o, p, want) in enumerate(zip(offsets, plans, wants)):
        got = host[o:o + p.output_bytes]
        assert np.array_equal(got, want), f"image {i}"


def test_c4_full_size_float_alpha(session):
    items = workloads.c4(count=2)
    plans = [HipDecodePlan.from_encoded(e, sp) for e, sp in items]
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert not np.any(stages)
    for (enc, spec), b in zip(items, bufs):
        want = O.decode(image_desc(enc, spec))
        assert np.array_equal(b.to_host(), want)
