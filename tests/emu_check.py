"""Runs every parity case through the CPU build of the HIP path and compares
with the oracle (run as a script, so it can start under the AddressSanitizer
runtime). Exit status 0 when every output matches."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import emu_lib  # noqa: E402


def main():
    asan = "--asan" in sys.argv
    emu_lib.load(asan)
    import oracle_lib as O
    from cases import CASES, make
    from jxr_amd import HipDecodePlan, HipDecoderSession, abi, image_desc

    names = sorted(CASES)
    only = [a for a in sys.argv[1:] if not a.startswith("--")]
    if only:
        names = [n for n in names if n in only]
    session = HipDecoderSession(0)
    failures = 0
    tiles = [(0, 0), (1, 1), (3, 2), (5, 3)] if "--tiles" in sys.argv else [(0, 0)]
    for tile in tiles:
        session.set_tile(*tile)
        for name in names:
            enc, spec, _ = make(CASES[name])
            plan = HipDecodePlan.from_encoded(enc, spec)
            want = O.decode(image_desc(enc, spec))
            got = session.decode_to_host(plan)
            if not np.array_equal(got, want):
                bad = np.flatnonzero(got != want)
                print(f"MISMATCH {name} tile {tile}: {bad.size} bytes, first {bad[:6]}", flush=True)
                failures += 1
    session.set_tile(0, 0)
    # samples beyond the output stage's plain-arithmetic bound
    import large_cases
    for tile in tiles:
        session.set_tile(*tile)
        for name, desc, stage in large_cases.cases():
            if only and name not in only:
                continue
            plan = HipDecodePlan(desc, keepalive=desc._keep)
            if stage:
                try:
                    session.decode_to_host(plan)
                    print(f"MISSING ERROR {name} tile {tile}", flush=True)
                    failures += 1
                except Exception as e:  # noqa: BLE001
                    if getattr(e, "stage", None) != stage:
                        print(f"WRONG ERROR {name} tile {tile}: {e}", flush=True)
                        failures += 1
                continue
            if not np.array_equal(session.decode_to_host(plan), O.decode(desc)):
                print(f"MISMATCH {name} tile {tile}", flush=True)
                failures += 1
    session.set_tile(0, 0)
    # reduced resolution: every case at a quarter and a sixteenth (the one
    # combination the oracle refuses, the plan refuses too)
    for tile in tiles[:2]:
        session.set_tile(*tile)
        for name in names:
            for scale in (abi.SCALE_QUARTER, abi.SCALE_SIXTEENTH):
                enc, spec, _ = make(dict(CASES[name], scale=scale))
                code, _, want = O.decode_status(image_desc(enc, spec))
                try:
                    got = session.decode_to_host(HipDecodePlan.from_encoded(enc, spec))
                except Exception as e:  # noqa: BLE001
                    if not code:
                        print(f"ERROR {name} scale {scale} tile {tile}: {e}", flush=True)
                        failures += 1
                    continue
                if code or not np.array_equal(got, want):
                    print(f"MISMATCH {name} scale {scale} tile {tile}", flush=True)
                    failures += 1
    session.set_tile(0, 0)
    # the session keeps sample-scale terms per configuration: a second image
    # of the same configuration launches no probe decode
    counts = []
    for seed in (11, 12):
        enc, spec, _ = make(dict(CASES["rgb8_420_ov1"], seed=seed, scale=abi.SCALE_QUARTER))
        plan = HipDecodePlan.from_encoded(enc, spec)
        before = session.diagnostics().kernel_launches
        session.upload(plan)
        counts.append(session.diagnostics().kernel_launches - before)
        if not np.array_equal(session.decode_to_host(plan), O.decode(image_desc(enc, spec))):
            print(f"MISMATCH probe-cache seed {seed}", flush=True)
            failures += 1
    if counts[1] != 0:
        print(f"PROBE TERMS recomputed: {counts}", flush=True)
        failures += 1
    # persistent batch over every case at once, and every third case at a
    # reduced scale besides
    encs, plans, wants = [], [], []
    batch_names = []
    for k, name in enumerate(names):
        scales = [abi.SCALE_FULL] + ([abi.SCALE_QUARTER + k // 3 % 2] if k % 3 == 0 else [])
        for scale in scales:
            enc, spec, _ = make(dict(CASES[name], scale=scale))
            if O.decode_status(image_desc(enc, spec))[0]:
                continue
            encs.append(enc)
            plans.append(HipDecodePlan.from_encoded(enc, spec))
            wants.append(O.decode(image_desc(enc, spec)))
            batch_names.append(name if scale == abi.SCALE_FULL else f"{name}@{scale}")
    names = batch_names
    for mode in (abi.BATCH_PERSISTENT, abi.BATCH_PER_IMAGE):
        bufs = [session.allocate_destination(p.output_bytes) for p in plans]
        stages = session.submit_batch(plans, [b.ptr for b in bufs], mode).wait()
        for name, b, want, st in zip(names, bufs, wants, stages):
            if st or not np.array_equal(b.to_host(), want):
                print(f"MISMATCH batch mode {mode} {name} stage {st}", flush=True)
                failures += 1
    # captured batch, replayed with the destinations scribbled between replays
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    graph = session.capture_batch(plans, [b.ptr for b in bufs])
    for replay in range(2):
        for b in bufs:
            b.fill(0x5A + replay)
        stages = graph.launch().wait()
        for name, b, want, st in zip(names, bufs, wants, stages):
            if st or not np.array_equal(b.to_host(), want):
                print(f"MISMATCH graph replay {replay} {name} stage {st}", flush=True)
                failures += 1
    graph.close()
    # autotuned shapes, then a batch at those shapes
    choices = session.autotune(plans, [b.ptr for b in bufs], repeats=1)
    if len(choices) != len({p.bucket_key for p in plans}):
        print(f"AUTOTUNE returned {len(choices)} buckets", flush=True)
        failures += 1
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    for name, b, want, st in zip(names, bufs, wants, stages):
        if st or not np.array_equal(b.to_host(), want):
            print(f"MISMATCH tuned batch {name} stage {st}", flush=True)
            failures += 1
    # serving loop: every case in one mix, half into destinations, half into
    # server-allocated outputs, under a budget that admits a few at a time
    from jxr_amd import HipServer
    from jxr_amd.prepare import RegionImage
    # two region windows whose arenas have no identity: uploaded per request
    # through the server's pinned staging
    import copy
    extra_names, extra_plans, extra_keep = [], [], []
    for name, region in (("rgb8_444_ov2_hard", (32, 16, 64, 48)), ("rgb8_420_ov2_hard_1mb", (16, 16, 48, 32))):
        enc, spec, _ = make(CASES[name])
        reg = RegionImage(enc, region, cache=False)
        sp = copy.copy(spec)
        sp.crop = region
        d = image_desc(reg, sp)
        extra_names.append(name + "_region")
        extra_plans.append(HipDecodePlan(d, keepalive=(reg, enc)))
        wants.append(O.decode(d))
        extra_keep.append(reg)
    names = names + extra_names
    plans = plans + extra_plans
    budget = max(p.output_bytes for p in plans) * 6 + (64 << 20)
    server = HipServer(session, hbm_budget_bytes=budget, max_wave_images=5, max_waves=2)
    dsts = {}
    for i, plan in enumerate(plans):
        if i % 2 == 0:
            b = session.allocate_destination(plan.output_bytes)
            b.fill(0x3C)
            dsts[i] = b
            server.submit(plan, b.ptr, plan.output_bytes, tag=i)
        else:
            server.submit(plan, tag=i)
    server.drain()
    results = server.poll(capacity=len(plans) + 8)
    if sorted(r.tag for r in results) != list(range(len(plans))):
        print(f"SERVER returned tags {sorted(r.tag for r in results)}", flush=True)
        failures += 1
    for r in results:
        got = dsts[r.tag].to_host() if r.tag in dsts else (r.image.readback() if r.image else None)
        if r.status or got is None or not np.array_equal(got[: wants[r.tag].nbytes], wants[r.tag]):
            print(f"MISMATCH server {names[r.tag]} status {r.status} stage {r.stage}", flush=True)
            failures += 1
        if r.image:
            r.image.close()
    st = server.stats()
    if st["completed"] != len(plans) or st["inflight_high_water_bytes"] > budget or st["waves"] < 2:
        print(f"SERVER stats {st}", flush=True)
        failures += 1
    server.close()
    # pool accounting: reused buffers count as live while held
    hw = session.diagnostics().scratch_high_water_bytes
    if not 0 < hw < (1 << 36):
        print(f"POOL high water {hw}", flush=True)
        failures += 1
    print(f"emulated parity: {failures} failures over {len(names)} cases", flush=True)
    sys.exit(1 if failures else 0)


if __name__ == "__main__":
    main()
