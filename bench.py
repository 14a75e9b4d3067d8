#!/usr/bin/env python3
"""Benchmark: JPEG XR reconstruction throughput on one MI355X.

A step is one reconstruction pass over one batch of synthetic arenas that are
already resident in HBM (uploaded before timing; the upload cache keeps them).
Default workload (c3): 256 mixed-size RGB 4:2:0 images to device-resident
NCHW, through the persistent batched kernel (the reported value) and through
the plain per-image path (one launch per image, four streams) for the
batched/per-image ratio. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c2|c3|c4]
                  [--scale full|quarter|sixteenth]

--scale decodes the same arenas at a reduced resolution (the reference's
DecodeScale): pixels are the reduced output's, and the algorithmic bytes the
coefficients such a decode reads.
"""
import argparse
import json
import os
import sys
import time
from types import SimpleNamespace

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpixel/s reconstruction on 1 MI355X, batched vs per-image; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_figure(name, workload):
    """A per-launch figure of the reconstruction kernel for `workload` from the
    committed rocprofv3 counter summaries (a line `<name> <workload> <value>`),
    or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_counters.md"))):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == name and parts[1] == workload:
                    return float(parts[2])
    return None


def measured_traffic(workload):
    """HBM bytes per launch (FETCH_SIZE and WRITE_SIZE, each scaled by what the
    counter reads for a 256 MiB device copy), or None."""
    v = profile_figure("traffic_bytes_per_launch", workload)
    return None if v is None else int(v)


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = max(0, min(len(xs) - 1, int(np.ceil(p / 100.0 * len(xs))) - 1))
    return xs[k]


def time_path(session, submit, steps, warmup):
    """Times `steps` calls of submit().wait() after `warmup` untimed ones."""
    for _ in range(warmup):
        submit().wait()
    session.synchronize()
    walls, devs, launches = [], [], 0
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        submit().wait()
        walls.append((time.perf_counter() - t0) * 1e3)
        ms, launches = session.last_batch_timing()
        devs.append(ms)
    session.synchronize()
    total = (time.perf_counter() - t_all) * 1e3
    return dict(wall_ms=walls, device_ms=devs, launches=launches, total_ms=total)


def cpu_baseline(items, plans, dense, offsets, budget_s=15.0):
    """The oracle (a C restatement of the reference CPU decoder's
    reconstruction) timed on this host's cores over a bounded sample of the
    same arenas; its outputs are compared with the GPU's as a parity check."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from jxr_amd import image_desc
    # the host cores this process may use, at most the sixteen a one-GPU
    # share of the host has
    threads = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))
    descs = [image_desc(enc, spec) for enc, spec in items]
    # calibrate on the first image, then size the sample to the budget
    t0 = time.perf_counter()
    O.decode_many(descs[:1], 1)
    one = max(time.perf_counter() - t0, 1e-4)
    px_first = plans[0].output_pixels
    rate1 = px_first / one
    target_px = rate1 * threads * budget_s
    n, px = 0, 0
    while n < len(descs) and (px < target_px or n == 0):
        px += plans[n].output_pixels
        n += 1
    # one untimed pass (its outputs are the parity check), then timed passes:
    # ten, or as many as fit twenty seconds (at least three); the median pass
    # gives the rate and the nearest-rank 95th percentile its spread
    outs = O.decode_many(descs[:n], threads)
    runs = []
    scratch = [np.empty_like(o) for o in outs]
    for o in scratch:
        o.fill(0)  # touched before timing
    while len(runs) < 3 or (len(runs) < 10 and sum(runs) < 20.0):
        t0 = time.perf_counter()
        O.decode_many(descs[:n], threads, scratch)
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    slow = sorted(runs)[max(0, int(np.ceil(0.95 * len(runs))) - 1)]
    host = dense.to_host() if dense is not None else None
    checked = 0
    if host is not None:
        for i in range(n):
            o = offsets[i]
            if not np.array_equal(host[o:o + outs[i].nbytes], outs[i]):
                raise SystemExit(f"parity failure: image {i} differs between GPU and CPU oracle")
            checked += 1
    return dict(value=round(px / dt / 1e6, 3), unit="Mpixel/s", cores=threads, kind="port",
                sample=f"{n} of {len(descs)} images ({px / 1e6:.1f} Mpx) per pass, median of {len(runs)} passes "
                       f"{dt:.2f} s (p95 {slow:.2f} s), oracle on {threads} threads: a scalar C restatement "
                       f"of the reconstruction, one image per thread, capped at the sixteen host cores of a "
                       f"one-GPU share; the reference's CPU decoder also runs SIMD lanes, so this understates it",
                p95_value=round(px / slow / 1e6, 3), passes=len(runs), parity_checked_images=checked)


def end_to_end(session, items, plans, dense, offsets, cpu_recon, passes=3):
    """The CPU half of the decode beside the reconstruction: every image of
    the workload coded into a codestream by the repo's writer (untimed), then
    (i) the host entropy decoder alone, images in parallel on the host cores;
    (ii) the CPU baseline end to end, that entropy decode followed by the
    oracle's reconstruction on the same cores; (iii) the GPU path end to end
    from codestreams in host memory: entropy decode and planning on host
    threads overlapped with the uploads and persistent-kernel launches of
    groups of images already decoded, into the same NCHW destinations, whose
    bytes are then checked against the resident run's."""
    from concurrent.futures import ThreadPoolExecutor
    from jxr_amd import HipDecodePlan, abi
    from jxr_amd.codestream import Codestream, encode_codestream
    threads = max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16))
    t0 = time.perf_counter()
    # a reduced decode reads band packets: frequency-mode codestreams
    frequency = any(spec.scale != abi.SCALE_FULL for _, spec in items)
    with ThreadPoolExecutor(threads) as ex:
        streams = list(ex.map(lambda it: Codestream(encode_codestream(
            it[0], frequency=frequency, annex_a=it[0].alpha_mode == abi.ALPHA_SEPARATE)), items))
    coded = sum(cs.info.codestream_bytes + cs.info.alpha_bytes for cs in streams)
    log(f"[e2e] {len(streams)} codestreams, {coded / 1e6:.1f} MB, written in {time.perf_counter() - t0:.1f} s")
    pixels = sum(p.output_pixels for p in plans)
    scales = [spec.scale for _, spec in items]

    counts = [sum(streams[k].coefficient_count(scales[k])) for k in range(len(streams))]
    # largest images first, so that a pass ends on small ones
    order = sorted(range(len(streams)), key=lambda k: -counts[k])

    # coefficient storage allocated once before the passes: page-locked
    # blocks of the session, one slice per image
    block = session.allocate_host(4 * sum(counts))
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64) * 4
    slices = [SimpleNamespace(ptr=block.ptr + int(starts[k]), nbytes=4 * counts[k]) for k in range(len(streams))]

    def decode_all(ex, into):
        return list(ex.map(lambda k: streams[k].decode(scale=scales[k], threads=1,
                                                       into=slices[k] if into else None), order))

    def entropy_pass(into):
        runs = []
        with ThreadPoolExecutor(threads) as ex:
            decode_all(ex, into)
            for _ in range(passes):
                t0 = time.perf_counter()
                decs = decode_all(ex, into)
                runs.append(time.perf_counter() - t0)
                del decs
        return float(np.median(runs))

    t_ent = entropy_pass(False)
    t_ent_into = entropy_pass(True)
    ent_rate = pixels / min(t_ent, t_ent_into) / 1e6
    log(f"[e2e] entropy decode: {t_ent * 1e3:.1f} ms per pass into decoder-allocated arenas, "
        f"{t_ent_into * 1e3:.1f} ms into preallocated storage; best {ent_rate:.1f} Mpx/s on {threads} threads")

    # (iii) pipelined: decoding on the pool, groups of images submitted as
    # soon as all of them are decoded. Coefficients go either into arenas the
    # decoder allocates (pageable: each upload is a staged copy on the
    # submitting thread) or into the page-locked slices (each upload an
    # asynchronous DMA)
    host_resident = dense.to_host()
    group = 16

    def decode_and_plan(k, pinned):
        dec = streams[k].decode(scale=scales[k], threads=1, into=slices[k] if pinned else None)
        return HipDecodePlan.from_codestream(streams[k], items[k][1], decoded=dec)

    def pipelined(pinned):
        walls = []
        # one core stays with the thread that submits the groups
        with ThreadPoolExecutor(max(1, threads - 1)) as ex:
            for rep in range(passes + 1):
                dense.fill(0)
                session.clear_cache()
                session.synchronize()
                t0 = time.perf_counter()
                futs = {k: ex.submit(decode_and_plan, k, pinned) for k in order}
                batches, keep = [], []
                for g in range(0, len(order), group):
                    idx = order[g:g + group]
                    gp = [futs[k].result() for k in idx]
                    batches.append(session.submit_batch(gp, [dense.ptr + offsets[k] for k in idx],
                                                        abi.BATCH_PERSISTENT))
                    keep.append(gp)
                for b in batches:
                    if b.wait().any():
                        raise SystemExit("end-to-end decode: an image failed on the device")
                session.synchronize()
                if rep:
                    walls.append(time.perf_counter() - t0)
                del keep, batches, futs
        if not np.array_equal(dense.to_host(), host_resident):
            raise SystemExit("end-to-end decode differs from the resident arenas' decode")
        session.clear_cache()
        return float(np.median(walls))

    t_pageable = pipelined(False)
    t_e2e = pipelined(True)
    block.close()
    cpu_e2e = None
    if cpu_recon is not None:
        cpu_e2e = 1.0 / (1.0 / ent_rate + 1.0 / cpu_recon["value"])
    log(f"[e2e] GPU from codestreams: {t_e2e * 1e3:.1f} ms per pass, {pixels / t_e2e / 1e6:.1f} Mpx/s "
        f"(pinned staging; pageable arenas {t_pageable * 1e3:.1f} ms, {pixels / t_pageable / 1e6:.1f} Mpx/s)")
    return {
        "codestream_mb": round(coded / 1e6, 2),
        "bits_per_pixel": round(coded * 8 / pixels, 3),
        "entropy_decode": {"value": round(ent_rate, 1), "unit": "Mpixel/s", "cores": threads,
                           "ms_per_pass": round(min(t_ent, t_ent_into) * 1e3, 2),
                           "ms_per_pass_allocating": round(t_ent * 1e3, 2),
                           "ms_per_pass_preallocated": round(t_ent_into * 1e3, 2),
                           "what": "host adaptive-VLC decode of every tile packet into arenas, one image per "
                                   "thread, largest first, median of passes; the better of decoder-allocated "
                                   "arenas and preallocated storage"},
        "gpu_from_codestreams": {"value": round(pixels / t_e2e / 1e6, 1), "unit": "Mpixel/s",
                                 "ms_per_pass": round(t_e2e * 1e3, 2), "group_images": group,
                                 "pageable_value": round(pixels / t_pageable / 1e6, 1),
                                 "what": "codestreams in host memory -> host entropy decode into page-locked "
                                         "blocks, overlapped with DMA uploads and persistent launches -> device "
                                         "NCHW, bytes checked against the resident decode; pageable_value: the "
                                         "same with decoder-allocated arenas"},
        "cpu_baseline_end_to_end": None if cpu_e2e is None else {
            "value": round(cpu_e2e, 1), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "what": "entropy decode then the oracle's reconstruction on the same host cores"},
    }


_SERVE_POOLS = {}


def serve_codestreams(session, seconds, budget_gb):
    """The c5 mix served from codestreams: each request's file is
    entropy-decoded on a host worker (only the tiles a region needs) into a
    page-locked block of a per-kind pool, planned, and submitted to the
    serving loop with its arrival time, so the latency runs from arrival
    through the entropy decode, the upload and the reconstruction. First a
    closed-loop pass (every request of a one-second schedule submitted at
    once) measures the capacity; then Poisson arrivals at half of it give
    p50/p99. Returns both."""
    import queue
    from concurrent.futures import ThreadPoolExecutor
    from jxr_amd import HipDecodePlan, HipServer, workloads
    from jxr_amd.codestream import Codestream, encode_codestream
    from jxr_amd.serving import clock_ns
    t0 = time.perf_counter()
    if True not in _SERVE_POOLS and False not in _SERVE_POOLS:
        pool = workloads.c5_pool(cache=True)
        _SERVE_POOLS[True] = (pool, {k: [HipDecodePlan.from_encoded(e, sp) for e, sp in pool[k]]
                                     for k in ("small", "gray", "region")})
    pool = (_SERVE_POOLS.get(True) or _SERVE_POOLS[False])[0]
    workers = max(2, min(len(os.sched_getaffinity(0)), 16) - 1)
    with ThreadPoolExecutor(workers) as ex:
        small = list(ex.map(lambda es: Codestream(encode_codestream(es[0])), pool["small"]))
        gray = list(ex.map(lambda es: Codestream(encode_codestream(es[0])), pool["gray"]))
    source = Codestream(encode_codestream(pool["source"]))
    # per request kind and pool index: codestream, output, tiles to decode
    info = {"small": [(cs, sp, None) for cs, (_, sp) in zip(small, pool["small"])],
            "gray": [(cs, sp, None) for cs, (_, sp) in zip(gray, pool["gray"])],
            "region": [(source, sp, source.region_tiles(sp.crop)) for _, sp in pool["region"]]}
    coded = {k: sum(cs.info.codestream_bytes for cs, _, _ in v) / len(v) for k, v in info.items()}
    log(f"[serve-cs] codestreams written in {time.perf_counter() - t0:.1f} s; mean file bytes "
        + ", ".join(f"{k} {v / 1e3:.0f} kB" for k, v in coded.items()))
    # page-locked blocks, one per request in flight, pooled by kind
    blocks_per_kind = {"small": 48, "gray": 12, "region": 24}
    slabs, free = [], {}
    for kind, count in blocks_per_kind.items():
        need = max(sum(cs.coefficient_count(sp.scale, m)) for cs, sp, m in info[kind]) * 4
        slab = session.allocate_host(need * count)
        slabs.append(slab)
        free[kind] = queue.SimpleQueue()
        for j in range(count):
            free[kind].put(SimpleNamespace(ptr=slab.ptr + j * need, nbytes=need))

    # the bytes of the first requests of each kind are compared with the
    # device decode of the encoder's own arenas for the same image
    want = {}
    checked = {"small": 0, "gray": 0, "region": 0}
    plans_pool = (_SERVE_POOLS.get(True) or _SERVE_POOLS[False])[1]

    def check(kind, k, image):
        if checked[kind] >= 4:
            return
        if (kind, k) not in want:
            want[(kind, k)] = session.decode_to_host(plans_pool[kind][k])
        if not np.array_equal(image.readback(), want[(kind, k)]):
            raise SystemExit(f"serving from codestreams: {kind} image {k} differs from its arena's decode")
        checked[kind] += 1

    def run(sched, arrivals, label):
        n = len(sched)
        server = HipServer(session, hbm_budget_bytes=int(budget_gb * (1 << 30)), max_wave_images=256, max_waves=4)
        keep, errors = {}, []

        def work(i):
            try:
                _, kind, k = sched[i]
                cs, sp, mask = info[kind][k]
                blk = free[kind].get()
                # a 4096x4096 file's 64 tiles decode on four threads of its own
                dec = cs.decode(scale=sp.scale, tiles=mask, threads=4 if kind == "gray" else 1, into=blk)
                # each request uploads its own coefficients (no upload cache entry)
                dec.primary.contents.identity = 0
                plan = HipDecodePlan.from_codestream(cs, sp, decoded=dec)
                keep[i] = (plan, blk, kind)
                server.submit(plan, tag=i, arrival_ns=arrivals[i])
            except Exception as e:  # noqa: BLE001
                errors.append(e)

        px = 0
        done = failed = 0
        i = 0
        deadline = time.perf_counter() + max(seconds, 3.0) + 120.0
        start = clock_ns()
        with ThreadPoolExecutor(workers) as ex:
            while done < n:
                if errors:
                    server.close()
                    raise SystemExit(f"serving from codestreams ({label}): {errors[0]}")
                if time.perf_counter() > deadline:
                    server.close()
                    raise SystemExit(f"serving from codestreams ({label}): {n - done} of {n} requests without "
                                     "a result two minutes after the stream")
                now = clock_ns()
                while i < n and arrivals[i] <= now:
                    ex.submit(work, i)
                    i += 1
                wait_us = 2000 if i >= n else int(max(0, min(2000, (arrivals[i] - clock_ns()) // 1000)))
                for r in server.poll(4096, timeout_us=wait_us):
                    plan, blk, kind = keep.pop(r.tag)
                    px += plan.output_pixels
                    if r.image is not None:
                        if r.status == 0:
                            check(kind, sched[r.tag][2], r.image)
                        r.image.close()
                    free[kind].put(blk)
                    failed += r.status != 0
                    done += 1
        end = clock_ns()
        st = server.stats()
        server.close()
        wall = (end - start) / 1e9
        log(f"[serve-cs] {label}: {n} requests in {wall:.2f} s, {px / wall / 1e6:.1f} Mpx/s, "
            f"p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms")
        return dict(requests=n, completed=st["completed"], failed=failed, seconds=round(wall, 3),
                    requests_per_s=round(n / wall, 1), mpx_per_s=round(px / wall / 1e6, 1),
                    latency_ms={"p50": round(st["p50_ms"], 3), "p99": round(st["p99_ms"], 3)})

    # closed loop: a one-second schedule's requests all arrive now
    burst = workloads.c5_schedule(1.0)
    now = clock_ns() + 5_000_000
    cap = run(burst, [now] * len(burst), "capacity (closed loop)")
    rate = max(10.0, 0.5 * cap["requests_per_s"])
    sched = workloads.c5_schedule(seconds, rate=rate)
    start = clock_ns() + 20_000_000
    open_loop = run(sched, [start + int(t * 1e9) for t, _, _ in sched], f"Poisson {rate:.0f} req/s")
    for slab in slabs:
        slab.close()
    return {
        "workload": "c5 mix from codestreams: host entropy decode (tiles a region needs) into page-locked blocks "
                    f"on {workers} host workers, per-request DMA upload, serving loop; latency from arrival",
        "capacity": cap,
        "poisson_half_capacity": dict(open_loop, offered_requests_per_s=round(rate, 1)),
        "bytes_checked": dict(checked),
    }


def serve(session, seconds, cache, budget_gb, rate=None):
    """The serving stream (Poisson arrivals, 2000/s unless `rate` says
    otherwise, the c5 mix) through the native serving loop; returns its
    summary. With `cache` the pool's coefficients stay resident on the device
    (uploaded once before the stream); without, every request uploads its
    own. The pool is built once per process."""
    from jxr_amd import HipDecodePlan, HipServer, workloads
    from jxr_amd.serving import clock_ns
    t0 = time.perf_counter()
    if cache not in _SERVE_POOLS:
        pool = workloads.c5_pool(cache=cache)
        _SERVE_POOLS[cache] = (pool, {k: [HipDecodePlan.from_encoded(e, sp) for e, sp in pool[k]]
                                      for k in ("small", "gray", "region")})
    pool, plans = _SERVE_POOLS[cache]
    log(f"[serve] pool of {sum(len(v) for v in plans.values())} images in {time.perf_counter() - t0:.1f} s")
    if cache:
        for v in plans.values():
            for p in v:
                session.upload(p)
        session.synchronize()
    rate = rate or workloads.C5_RATE
    sched = workloads.c5_schedule(seconds, rate=rate)
    n = len(sched)
    offered_px = sum(plans[kind][k].output_pixels for _, kind, k in sched)
    server = HipServer(session, hbm_budget_bytes=int(budget_gb * (1 << 30)), max_wave_images=256, max_waves=4)
    start = clock_ns() + 20_000_000
    arrivals = [start + int(t * 1e9) for t, _, _ in sched]
    i = done = failed = 0
    next_log = time.perf_counter() + 5.0
    deadline = time.perf_counter() + seconds + 120.0
    while done < n:
        if time.perf_counter() > deadline:
            server.close()
            raise SystemExit(f"serving: {n - done} of {n} requests without a result two minutes after the stream")
        now = clock_ns()
        while i < n and arrivals[i] <= now:
            _, kind, k = sched[i]
            server.submit(plans[kind][k], tag=i, arrival_ns=arrivals[i])
            i += 1
        wait_us = 2000 if i >= n else int(max(0, min(2000, (arrivals[i] - clock_ns()) // 1000)))
        for r in server.poll(4096, timeout_us=wait_us):
            if r.image is not None:
                r.image.close()
            failed += r.status != 0
            done += 1
        if time.perf_counter() > next_log:
            log(f"[serve] {i}/{n} submitted, {done} done")
            next_log += 5.0
    end = clock_ns()
    st = server.stats()
    server.close()
    wall = (end - start) / 1e9
    return {
        "workload": f"c5: Poisson {rate:g} req/s; 70% RGB u8 256..1024 q75 4:4:4 ov1, 20% 4096x4096 gray u16 8x8 "
                    "tiles ov2 q75, 10% 512x512 regions of 8192x8192 RGB u8 (512-px tiles)",
        "seconds_offered": seconds,
        "uploads": "resident (pool uploaded once)" if cache else "per request (PCIe inside the latency)",
        "requests": n, "completed": st["completed"], "failed": failed, "rejected": st["rejected"],
        "offered_mpx_per_s": round(offered_px / seconds / 1e6, 1),
        "mpx_per_s": round(st["pixels_completed"] / wall / 1e6, 1),
        "latency_ms": {"p50": round(st["p50_ms"], 3), "p99": round(st["p99_ms"], 3)},
        "queue_ms": {"p50": round(st["queue_p50_ms"], 3), "p99": round(st["queue_p99_ms"], 3)},
        "waves": st["waves"], "mean_wave_images": round(st["mean_wave_images"], 2),
        "max_wave_images": st["max_wave_images"],
        "inflight_high_water_gb": round(st["inflight_high_water_bytes"] / 1e9, 3),
        "hbm_high_water_gb": round(st["hbm_high_water_bytes"] / 1e9, 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-image", action="store_true")
    ap.add_argument("--no-end-to-end", action="store_true",
                    help="skip the codestream figures (host entropy decode, decode from codestreams)")
    ap.add_argument("--tile", default=None, help="TXxTY macroblocks per work item")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--no-graph", action="store_true", help="submit the batch eagerly every step")
    ap.add_argument("--serve-rate", type=float, default=0.0, help="arrivals per second (0: the c5 rate, 2000)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="seconds of the c5 serving stream to run after the batch (0: none)")
    ap.add_argument("--serve-uploads", choices=["resident", "per-request"], default="resident")
    ap.add_argument("--serve-codestream-seconds", type=float, default=4.0,
                    help="Poisson run of the c5 mix served from codestreams (0: skip)")
    ap.add_argument("--serve-loaded-rate", type=float, default=40000.0,
                    help="arrivals per second of a second, loaded serving run (0: none)")
    ap.add_argument("--serve-budget-gb", type=float, default=0.0, help="admission budget (0: the server's default, 8 GiB or 3/4 of free HBM if less)")
    ap.add_argument("--scale", default="full", choices=["full", "quarter", "sixteenth"],
                    help="decode resolution (DecodeScale)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also copy 256 MiB device-to-device (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    # The library links the ROCm HIP runtime; torch bundles a second copy,
    # so this process allocates and synchronizes through the session.
    from jxr_amd import HipDecodePlan, HipDecoderSession, abi, workloads

    desc, make = workloads.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    items = make()
    log(f"[bench] built {len(items)} arenas for {args.workload} in {time.perf_counter() - t0:.1f} s")
    scale = {"full": abi.SCALE_FULL, "quarter": abi.SCALE_QUARTER, "sixteenth": abi.SCALE_SIXTEENTH}[args.scale]
    for _, spec in items:
        spec.scale = scale
    # counter summaries are of full-resolution decodes
    figures = args.workload if scale == abi.SCALE_FULL else None
    session = HipDecoderSession(0)
    if args.tile:
        tx, ty = (int(v) for v in args.tile.lower().split("x"))
        session.set_tile(tx, ty)
    plans = [HipDecodePlan.from_encoded(enc, spec) for enc, spec in items]
    offsets, total = [], 0
    for p in plans:
        offsets.append(total)
        total += (p.output_bytes + 255) // 256 * 256
    dense = session.allocate_destination(total)
    dense.fill(0)
    ptrs = [dense.ptr + o for o in offsets]
    t0 = time.perf_counter()
    for p in plans:
        session.upload(p)
    session.synchronize()
    log(f"[bench] uploaded {sum(e.coefficient_bytes() for e, _ in items) / 1e9:.2f} GB of coefficients "
        f"in {time.perf_counter() - t0:.1f} s")
    pixels = sum(p.output_pixels for p in plans)
    alg = sum(p.algorithmic_bytes for p in plans)
    if args.calibrate:
        a, b = session.allocate_destination(256 << 20), session.allocate_destination(256 << 20)
        a.fill(1)
        session.copy_device(b.ptr, a.ptr, 256 << 20)
        log("[bench] calibration: copied 268435456 bytes device to device")
        a.close()
        b.close()

    tuned = []
    if not args.no_autotune and not args.tile:
        t0 = time.perf_counter()
        tuned = session.autotune(plans, ptrs, repeats=3)
        for c in tuned:
            log(f"[bench] autotune bucket {c['bucket_key']:#x}: {c['images']} images, tile "
                f"{c['tile'][0]}x{c['tile'][1]} at {c['groups_per_cu']} workgroups/CU {c['best_ms']:.3f} ms "
                f"(default {c['default_ms']:.3f} ms, {c['candidates']} shapes)")
        log(f"[bench] autotune took {time.perf_counter() - t0:.1f} s")
    graph = None
    if args.no_graph:
        def submit_batched():
            return session.submit_batch(plans, ptrs, abi.BATCH_PERSISTENT)
    else:
        graph = session.capture_batch(plans, ptrs)
        submit_batched = graph.launch
    batched = time_path(session, submit_batched, args.steps, args.warmup)
    log(f"[bench] batched: median wall {np.median(batched['wall_ms']):.3f} ms, "
        f"device {np.median(batched['device_ms']):.3f} ms, launches/step {batched['launches']}")
    if graph is not None:
        graph.close()
    per_image = per_image_sync = None
    if not args.no_per_image:
        k = max(3, args.steps // 2)
        per_image = time_path(session, lambda: session.submit_batch(plans, ptrs, abi.BATCH_PER_IMAGE), k, 1)
        # the drop-in used as the reference's decoder uses its backend: one
        # image submitted and waited for at a time
        sizes = [p.output_bytes for p in plans]

        class _Each:
            def wait(self):
                for p, d, n in zip(plans, ptrs, sizes):
                    session.submit_into(p, d, n).wait()
        per_image_sync = time_path(session, _Each, max(2, k // 2), 1)
        log(f"[bench] per-image, one at a time: median wall {np.median(per_image_sync['wall_ms']):.3f} ms")
        log(f"[bench] per-image: median wall {np.median(per_image['wall_ms']):.3f} ms, "
            f"device {np.median(per_image['device_ms']):.3f} ms")

    ms_step = batched["total_ms"] / args.steps
    value = pixels / (ms_step / 1e3) / 1e6
    dev_ms = float(np.median(batched["device_ms"]))
    launches = max(batched["launches"], 1)
    per_launch_ms = dev_ms / launches
    achieved = alg / launches / (per_launch_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {desc}" + ("" if scale == abi.SCALE_FULL else f", decoded at a {args.scale}"),
            "images": len(items),
            "megapixels_per_step": round(pixels / 1e6, 3),
            "algorithmic_gb_per_step": round(alg / 1e9, 4),
            "path": "persistent batched kernel, one launch per shape bucket"
                    + ("" if args.no_graph else ", replayed from a captured HIP graph")
                    + (", autotuned tile shapes" if tuned else ""),
            "tiles": {f"{c['bucket_key']:#x}": f"{c['tile'][0]}x{c['tile'][1]}" for c in tuned},
            "workgroups_per_cu": {f"{c['bucket_key']:#x}": c["groups_per_cu"] for c in tuned},
            "parallelism": "single-gpu",
        },
        "latency_ms": {"p50": round(percentile(batched["wall_ms"], 50), 4),
                       "p99": round(percentile(batched["wall_ms"], 99), 4)},
        "device_ms_per_step": round(dev_ms, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(figures),
                     "kernel": "jxr_reconstruct_tiles", "avg_launch_ms": round(per_launch_ms, 4)},
    }
    valu = profile_figure("valu_instructions_per_launch", figures)
    if valu is not None:
        # the kernel's VALU instruction rate against the rate a plain chain
        # of independent integer VALU instructions reaches on this device
        rate = session.calibrate_valu()
        cycles = rate["simds"] * rate["clock_khz"] * 1e3
        peak = rate["wave_instructions"] / (rate["ms"] / 1e3) / cycles
        got = valu / (per_launch_ms / 1e3) / cycles
        out["valu_issue"] = {"achieved": round(got, 4), "peak": round(peak, 4),
                             "unit": "wave VALU instructions per SIMD cycle", "frac": round(got / peak, 4),
                             "instructions_per_launch": valu, "kernel": "jxr_reconstruct_tiles",
                             "peak_from": "jxrhip_calibrate_valu (16 independent integer VALU ops per "
                                          "iteration, 4 workgroups per CU)"}
        log(f"[bench] VALU issue: {got:.4f} of {peak:.4f} wave instructions per SIMD cycle ({got / peak:.1%})")
    if per_image is not None:
        pi_ms = per_image["total_ms"] / len(per_image["wall_ms"])
        pi_value = pixels / (pi_ms / 1e3) / 1e6
        out["per_image"] = {"value": round(pi_value, 2), "unit": "Mpixel/s", "ms_per_step": round(pi_ms, 4),
                            "launches_per_step": per_image["launches"],
                            "latency_ms": {"p50": round(percentile(per_image["wall_ms"], 50), 4),
                                           "p99": round(percentile(per_image["wall_ms"], 99), 4)}}
        out["batched_over_per_image"] = round(value / pi_value, 3)
        ps_ms = per_image_sync["total_ms"] / len(per_image_sync["wall_ms"])
        ps_value = pixels / (ps_ms / 1e3) / 1e6
        out["per_image_one_at_a_time"] = {"value": round(ps_value, 2), "unit": "Mpixel/s",
                                          "ms_per_step": round(ps_ms, 4)}
        out["batched_over_per_image_one_at_a_time"] = round(value / ps_value, 3)
    if args.serve_seconds > 0:
        out["serving"] = serve(session, args.serve_seconds, args.serve_uploads == "resident", args.serve_budget_gb,
                               args.serve_rate)
        if args.serve_loaded_rate > 0:
            # the same mix offered twenty times faster: continuous batching
            # with the device busy (waves of several images)
            out["serving_loaded"] = serve(session, min(args.serve_seconds, 5.0), args.serve_uploads == "resident",
                                          args.serve_budget_gb, args.serve_loaded_rate)
        if args.serve_uploads == "resident":
            # the same stream with every request uploading its own
            # coefficients from host memory: PCIe inside the latency
            out["serving_uploads"] = serve(session, min(args.serve_seconds, 3.0), False, args.serve_budget_gb,
                                           args.serve_rate)
        if args.serve_codestream_seconds > 0:
            out["serving_codestreams"] = serve_codestreams(session, args.serve_codestream_seconds,
                                                           args.serve_budget_gb)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(items, plans, dense, offsets)
    if not args.no_end_to_end:
        out["end_to_end"] = end_to_end(session, items, plans, dense, offsets, out.get("cpu_baseline"))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
