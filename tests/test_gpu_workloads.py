"""The benchmark configurations at their full sizes on the GPU, bit-exact
against the CPU oracle:

  C1  512x512 RGB u8 lossless 4:4:4 ov1: also equal to the source pixels
      (the lossless round trip);
  C2  4096x4096 gray u16 pink noise, 8x8 soft tiles, ov2, q75, flexbits;
  C3  all 256 mixed 4:2:0 images as one persistent batch into one dense NCHW
      allocation, replayed from a captured graph;
  C4  all 128 of the 1024x1024 RGBA f32 lossless images with separate alpha
      as one persistent batch (2 GiB of planar f32 output);
  C5  a Poisson burst of the serving mix (RGB u8 at 256..1024 px, 4096x4096
      gray u16, 512x512 windows of an 8192x8192 source) through the serving
      loop at its default HBM budget, every result's bytes equal to the
      oracle's decode of its pool entry.
"""
import time
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


def test_c4_full_batch_float_alpha(session):
    items = workloads.c4(count=128)
    plans = [HipDecodePlan.from_encoded(e, sp) for e, sp in items]
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    for b in bufs:
        b.fill(0x5C)
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert not np.any(stages)
    for k in range(0, len(items), 16):
        chunk = items[k:k + 16]
        wants = O.decode_many([image_desc(e, sp) for e, sp in chunk], 16)
        for i, want in enumerate(wants):
            assert np.array_equal(bufs[k + i].to_host(), want), f"image {k + i}"
    for b in bufs:
        b.close()


def test_c5_poisson_burst_through_the_serving_loop(session):
    from jxr_amd import HipServer
    from jxr_amd.serving import clock_ns
    pool = workloads.c5_pool()
    plans = {k: [HipDecodePlan.from_encoded(e, sp) for e, sp in pool[k]] for k in ("small", "gray", "region")}
    sched = workloads.c5_schedule(0.5)
    used = sorted({(kind, k) for _, kind, k in sched})
    with ThreadPoolExecutor(16) as ex:
        wants = dict(zip(used, ex.map(lambda u: O.decode(image_desc(*pool[u[0]][u[1]])), used)))
    print(f"[c5] {len(sched)} requests over {len(used)} pool entries", flush=True)
    server = HipServer(session)  # the default budget (8 GiB, or 3/4 of free HBM if less)
    start = clock_ns() + 10_000_000
    arrivals = [start + int(t * 1e9) for t, _, _ in sched]
    i = done = 0
    deadline = time.perf_counter() + 120.0
    try:
        while done < len(sched):
            assert time.perf_counter() < deadline, f"{len(sched) - done} requests without a result"
            now = clock_ns()
            while i < len(sched) and arrivals[i] <= now:
                _, kind, k = sched[i]
                server.submit(plans[kind][k], tag=i, arrival_ns=arrivals[i])
                i += 1
            for r in server.poll(1024, timeout_us=1000):
                _, kind, k = sched[r.tag]
                assert r.status == 0 and r.image is not None, (r.tag, r.status, r.stage)
                want = wants[(kind, k)]
                got = r.image.readback()
                r.image.close()
                assert np.array_equal(got[: want.nbytes], want), (r.tag, kind, k)
                done += 1
        st = server.stats()
        assert st["completed"] == len(sched) and st["failed"] == 0 and st["rejected"] == 0
    finally:
        server.close()
