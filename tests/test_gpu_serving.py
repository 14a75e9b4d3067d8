"""The serving loop on the GPU: every request's output is the same alone and
in a mix of concurrent requests (bit-exact, against the per-image path and
the CPU oracle), under an HBM budget small enough that admission has to wait
for earlier requests to finish; requests larger than the whole budget are
refused without disturbing the rest."""
import time

import numpy as np
import pytest

import oracle_lib as O
from cases import CASES, make, quality_qp
from jxr_amd import HipDecodePlan, HipDecoderSession, HipServer, abi, image_desc, synth, workloads
from jxr_amd.decoder import EncodedImage
from jxr_amd.prepare import RegionImage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


def _requests():
    """The parity cases plus larger images and region windows, so that waves
    hold real work: (name, plan, keepalive)."""
    out = []
    for name in sorted(CASES):
        enc, spec, _ = make(CASES[name])
        out.append((name, HipDecodePlan.from_encoded(enc, spec), enc))
    qp = quality_qp(75)
    for i, (w, h) in enumerate([(1024, 768), (640, 1024), (2048, 512)]):
        px = synth.gradient_noise(w, h, 90 + i, depth=8)
        enc = EncodedImage(px, color=abi.COLOR_YUV420 if i % 2 else abi.COLOR_YUV444, overlap=1 + i % 2,
                           scaled=True, qp=(qp, qp, qp + 4))
        out.append((f"large{i}", HipDecodePlan.from_encoded(enc, workloads.rgb_out(8, i == 2)), enc))
    px = synth.gradient_noise(2048, 2048, 99, depth=8)
    src = EncodedImage(px, color=abi.COLOR_YUV444, overlap=2, tiles=([16] * 8, [16] * 8), scaled=True,
                       qp=(qp, qp, qp + 4))
    for j, (x, y) in enumerate([(512, 512), (1040, 256), (0, 1536)]):
        reg = RegionImage(src, (x, y, 512, 512), cache=j != 1)  # one region uploads per request
        spec = workloads.rgb_out(8, False)
        spec.crop = (x, y, 512, 512)
        out.append((f"region{j}", HipDecodePlan(image_desc(reg, spec), keepalive=reg), reg))
    return out


@pytest.fixture(scope="module")
def requests(session):
    reqs = _requests()
    alone = [session.decode_to_host(plan) for _, plan, _ in reqs]
    return reqs, alone


def test_alone_matches_oracle(requests):
    reqs, alone = requests
    for (name, plan, keep), got in zip(reqs, alone):
        assert np.array_equal(got, O.decode(plan.desc)), name


@pytest.mark.parametrize("waves,budget_mb", [(4, 0), (2, 96), (1, 48)])
def test_mix_matches_alone(session, requests, waves, budget_mb):
    reqs, alone = requests
    rng = np.random.default_rng(waves)
    order = np.concatenate([rng.permutation(len(reqs)) for _ in range(3)])  # each request three times
    server = HipServer(session, hbm_budget_bytes=budget_mb << 20, max_wave_images=8, max_waves=waves)
    dsts = {}
    try:
        for n, i in enumerate(order):
            name, plan, _ = reqs[i]
            if n % 3 == 0:  # into a caller destination
                b = session.allocate_destination(plan.output_bytes)
                b.fill(0x5A)
                dsts[n] = b
                server.submit(plan, b.ptr, plan.output_bytes, tag=n)
            else:
                server.submit(plan, tag=n)
            if n % 7 == 0:
                time.sleep(0.0005)  # arrivals while earlier waves run
        server.drain()
        results = server.poll(capacity=len(order) + 4)
        assert sorted(r.tag for r in results) == list(range(len(order)))
        for r in results:
            i = order[r.tag]
            assert r.status == 0, (reqs[i][0], r.status, r.stage)
            got = dsts[r.tag].to_host() if r.tag in dsts else r.image.readback()
            assert np.array_equal(got[: alone[i].nbytes], alone[i]), reqs[i][0]
            assert r.latency_ms >= r.queue_ms >= 0
            if r.image:
                r.image.close()
        st = server.stats()
        assert st["completed"] == len(order) and st["failed"] == 0 and st["rejected"] == 0
        assert st["waves"] >= 2
        if budget_mb:
            assert st["inflight_high_water_bytes"] <= budget_mb << 20
        assert st["p99_ms"] >= st["p50_ms"] > 0
    finally:
        server.close()


def test_oversized_request_is_refused(session, requests):
    reqs, alone = requests
    big = max(range(len(reqs)), key=lambda i: reqs[i][1].output_bytes)
    small = min(range(len(reqs)), key=lambda i: reqs[i][1].output_bytes)
    budget = reqs[big][1].output_bytes // 2
    assert reqs[small][1].output_bytes * 4 < budget
    server = HipServer(session, hbm_budget_bytes=budget)
    try:
        server.submit(reqs[small][1], tag=0)
        server.submit(reqs[big][1], tag=1)
        server.submit(reqs[small][1], tag=2)
        server.drain()
        res = {r.tag: r for r in server.poll(8)}
        assert res[1].status == abi.ERR_BUDGET and res[1].image is None
        for t in (0, 2):
            assert res[t].status == 0
            assert np.array_equal(res[t].image.readback(), alone[small])
            res[t].image.close()
        assert server.stats()["rejected"] == 1
    finally:
        server.close()


def test_held_results_slow_admission_but_every_request_finishes(session, requests):
    """Result images the caller keeps count against the budget: with room
    for one and a half requests and a caller that polls only after draining,
    the kept outputs fill the budget after a few requests, the loop goes on
    one request at a time, and every request still finishes with the right
    bytes."""
    reqs, alone = requests
    i = next(k for k, r in enumerate(reqs) if r[0] == "large0")
    plan = reqs[i][1]
    assert plan.output_bytes * 4 < plan.device_footprint * 3 // 2 < plan.output_bytes * 10
    server = HipServer(session, hbm_budget_bytes=plan.device_footprint * 3 // 2, max_waves=1)
    try:
        for n in range(10):
            server.submit(plan, tag=n)
        server.drain()
        results = server.poll(capacity=16)
        assert sorted(r.tag for r in results) == list(range(10))
        for r in results:
            assert r.status == 0
            assert np.array_equal(r.image.readback()[: alone[i].nbytes], alone[i])
            r.image.close()
        st = server.stats()
        assert st["completed"] == 10 and st["failed"] == 0 and st["rejected"] == 0
    finally:
        server.close()


def test_codestream_requests_from_worker_threads_match_alone(session):
    """Requests arriving as files: four worker threads entropy-decode each
    one into a page-locked block of the session (tiles a region needs only),
    plan it and submit it while the main thread polls; every result equals
    the device decode of the encoder's own arenas for the same image, alone."""
    import queue
    from concurrent.futures import ThreadPoolExecutor
    from jxr_amd.codestream import Codestream, encode_codestream
    qp = quality_qp(75)
    items = []  # (codestream, spec, tiles, encoder arenas)
    for i, (w, h) in enumerate([(320, 256), (512, 384), (208, 528)]):
        px = synth.gradient_noise(w, h, 120 + i, depth=8)
        enc = EncodedImage(px, color=abi.COLOR_YUV420 if i % 2 else abi.COLOR_YUV444, overlap=1 + i % 2,
                           scaled=True, qp=(qp, qp, qp + 4), tiles=([8, w // 16 - 8], [h // 16]) if i == 2 else None)
        items.append((Codestream(encode_codestream(enc)), workloads.rgb_out(8, i == 1), None, enc))
    px = synth.gradient_noise(1024, 1024, 130, depth=8)
    src = EncodedImage(px, color=abi.COLOR_YUV444, overlap=2, tiles=([16] * 4, [16] * 4), scaled=True,
                       qp=(qp, qp, qp + 4))
    src_cs = Codestream(encode_codestream(src))
    for (x, y) in [(256, 256), (520, 128), (0, 768)]:
        spec = workloads.rgb_out(8, False)
        spec.crop = (x, y, 256, 256)
        items.append((src_cs, spec, src_cs.region_tiles(spec.crop), RegionImage(src, spec.crop)))
    alone = [session.decode_to_host(HipDecodePlan(image_desc(e, sp), keepalive=e)) for _, sp, _, e in items]
    need = max(sum(cs.coefficient_count(sp.scale, m)) for cs, sp, m, _ in items) * 4
    blocks = [session.allocate_host(need) for _ in range(6)]
    free = queue.SimpleQueue()
    for b in blocks:
        free.put(b)
    order = [k % len(items) for k in range(4 * len(items))]
    server = HipServer(session, hbm_budget_bytes=64 << 20, max_wave_images=4, max_waves=2)
    keep = {}

    def work(n):
        cs, sp, mask, _ = items[order[n]]
        blk = free.get()
        dec = cs.decode(scale=sp.scale, tiles=mask, threads=1, into=blk)
        plan = HipDecodePlan.from_codestream(cs, sp, decoded=dec)
        keep[n] = (plan, blk)
        server.submit(plan, tag=n)

    try:
        done = 0
        with ThreadPoolExecutor(4) as ex:
            futs = [ex.submit(work, n) for n in range(len(order))]
            deadline = time.time() + 60
            while done < len(order) and time.time() < deadline:
                for r in server.poll(64, timeout_us=2000):
                    i = order[r.tag]
                    assert r.status == 0, (r.tag, r.status, r.stage)
                    got = r.image.readback()
                    r.image.close()
                    assert np.array_equal(got[: alone[i].nbytes], alone[i]), i
                    _, blk = keep.pop(r.tag)
                    free.put(blk)
                    done += 1
            for f in futs:
                f.result()
        assert done == len(order)
        assert server.stats()["completed"] == len(order)
    finally:
        server.close()
        for b in blocks:
            b.close()
