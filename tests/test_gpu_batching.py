"""Captured batches and the tile autotuner, byte for byte against the CPU
oracle through the C ABI.

A captured batch is replayed several times, with eager work (destination
fills and an eager batch) between the replays, and every replay must equal
the oracle. The autotuner's chosen shapes must keep every image identical
on the per-image and the persistent path."""
import numpy as np
import pytest

import oracle_lib as O
from cases import CASES, make
from jxr_amd import HipDecodePlan, HipDecoderSession, JxrHipError, abi, image_desc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


@pytest.fixture(scope="module")
def corpus():
    names = sorted(CASES)
    encs, plans, wants = [], [], []
    for name in names:
        enc, spec, _ = make(CASES[name])
        encs.append(enc)
        plans.append(HipDecodePlan.from_encoded(enc, spec))
        wants.append(O.decode(image_desc(enc, spec)))
    return names, encs, plans, wants


def _check(names, bufs, wants, stages, what):
    assert not np.any(stages), f"{what}: stages {stages}"
    for name, b, want in zip(names, bufs, wants):
        got = b.to_host(0, want.nbytes)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, f"{what} {name}: {bad.size} bytes differ, first at {bad[:8]}"


def test_graph_replays_match_oracle(session, corpus):
    names, _, plans, wants = corpus
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    other = [session.allocate_destination(p.output_bytes) for p in plans]
    graph = session.capture_batch(plans, [b.ptr for b in bufs])
    try:
        for replay in range(4):
            for b in bufs:
                b.fill(0x11 * (replay + 1))
            stages = graph.launch().wait()
            _check(names, bufs, wants, stages, f"replay {replay}")
            # eager work between replays
            stages = session.submit_batch(plans, [b.ptr for b in other], abi.BATCH_PERSISTENT).wait()
            _check(names, other, wants, stages, f"eager after replay {replay}")
    finally:
        graph.close()


def test_graph_refuses_overlapping_replays(session, corpus):
    names, _, plans, wants = corpus
    bufs = [session.allocate_destination(p.output_bytes) for p in plans[:4]]
    graph = session.capture_batch(plans[:4], [b.ptr for b in bufs])
    try:
        first = graph.launch()
        with pytest.raises(JxrHipError) as err:
            graph.launch()
        assert err.value.code == abi.ERR_INVALID_ARGUMENT
        _check(names[:4], bufs, wants[:4], first.wait(), "first replay")
        _check(names[:4], bufs, wants[:4], graph.launch().wait(), "after the wait")
    finally:
        graph.close()


def test_graph_reports_indexed_errors(session):
    enc_ok, spec_ok, _ = make(CASES["rgb8_420_ov1"])
    enc_bad, spec_bad, _ = make(dict(w=32, h=32, ov=1, quality=100))
    arena = enc_bad.primary.contents
    arena.coefficients[arena.coefficient_offsets[0]] = 2**31 - 1  # leaves i32 after dequantization
    arena.quantizers[0] = 2
    plans = [HipDecodePlan.from_encoded(enc_ok, spec_ok), HipDecodePlan.from_encoded(enc_bad, spec_bad)]
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    graph = session.capture_batch(plans, [b.ptr for b in bufs])
    try:
        for _ in range(2):  # the status words are cleared inside the graph
            stages = graph.launch().wait(raise_on_error=False)
            assert list(stages) == [0, abi.STAGE_FIRST_TRANSFORM]
            assert np.array_equal(bufs[0].to_host(), O.decode(image_desc(enc_ok, spec_ok)))
    finally:
        graph.close()


def test_autotuned_shapes_keep_parity(corpus):
    names, _, plans, wants = corpus
    s = HipDecoderSession(0)
    try:
        bufs = [s.allocate_destination(p.output_bytes) for p in plans]
        choices = s.autotune(plans, [b.ptr for b in bufs], repeats=1)
        keys = {p.bucket_key for p in plans}
        assert {c["bucket_key"] for c in choices} == keys
        for c in choices:
            assert s.bucket_tile(c["bucket_key"]) == c["tile"]
            assert c["best_ms"] <= c["default_ms"] + 1e-9
        # the autotuner's own runs left the images in the destinations
        _check(names, bufs, wants, np.zeros(len(plans), np.uint32), "autotune output")
        for b in bufs:
            b.fill(0xC3)
        stages = s.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
        _check(names, bufs, wants, stages, "tuned persistent")
        for name, plan, want in zip(names, plans, wants):
            assert np.array_equal(s.decode_to_host(plan), want), f"tuned per-image {name}"
        # an installed shape is used, and removing it restores the default
        key = plans[0].bucket_key
        s.set_bucket_tile(key, 2, 8)
        assert s.bucket_tile(key) == (2, 8)
        assert np.array_equal(s.decode_to_host(plans[0]), wants[0])
        s.set_bucket_tile(key, 0, 0)
        assert s.bucket_tile(key) == (0, 0)
    finally:
        s.close()


def test_fresh_uploads_read_from_other_streams(session):
    """Coefficient and descriptor uploads are asynchronous on the stream of
    the submission that made them: the same large plans submitted again at
    once (the per-image path spreads submissions over the session's four
    streams, and a batch runs on its own) read the cached copies only after
    those copies have finished. Repeated after emptying the caches, each
    time with every output compared to the oracle."""
    from jxr_amd import synth, workloads
    from jxr_amd.decoder import EncodedImage
    from cases import quality_qp
    qp = quality_qp(60)
    items = []
    for i, (w, h) in enumerate([(2048, 1536), (1536, 2048), (1024, 1024)]):
        px = synth.gradient_noise(w, h, 300 + i, depth=8)
        enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=1 + i % 2, scaled=True, qp=(qp, qp, qp + 2))
        spec = workloads.rgb_out(8, True)
        items.append((enc, HipDecodePlan.from_encoded(enc, spec), O.decode(image_desc(enc, spec))))
    for rep in range(3):
        session.clear_cache()
        subs = []
        for _ in range(2):
            for enc, plan, want in items:
                subs.append((session.submit(plan), want))
        bufs = [session.allocate_destination(plan.output_bytes) for _, plan, _ in items]
        batch = session.submit_batch([p for _, p, _ in items], [b.ptr for b in bufs], abi.BATCH_PERSISTENT)
        for sub, want in subs:
            img = sub.wait()
            assert np.array_equal(img.readback(), want), rep
            img.close()
        assert not batch.wait().any()
        for b, (_, _, want) in zip(bufs, items):
            assert np.array_equal(b.to_host(0, want.nbytes), want), rep
