"""Lifetime and upload-ordering checks of the runtime on its CPU build
(run as a script under AddressSanitizer by tests/test_emulated.py).

1. A session closed while a result image, an unwaited batch, a pending
   submission and a destination are alive stays valid until the last of
   them is released; each is then used and released without touching freed
   memory.
2. Under an upload budget that holds one arena at a time, a serving loop
   with two compute streams re-uploads arenas evicted between waves; every
   result equals the oracle's decode.
Exit status 0 when every check passes."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import emu_lib  # noqa: E402


def main():
    emu_lib.load("--asan" in sys.argv)
    import oracle_lib as O
    from cases import CASES, make
    from jxr_amd import HipDecodePlan, HipDecoderSession, HipServer, abi, image_desc

    failures = 0
    names = ["rgb8_420_ov1", "gray16_ov2_q75", "rgb8_444_ov2_hard"]
    encs = [make(CASES[n]) for n in names]
    plans = [HipDecodePlan.from_encoded(e, sp) for e, sp, _ in encs]
    wants = [O.decode(image_desc(e, sp)) for e, sp, _ in encs]

    # 1. owners outliving the session handle
    session = HipDecoderSession(0)
    image = session.submit(plans[0]).wait()
    dst = session.allocate_destination(plans[1].output_bytes)
    batch = session.submit_batch([plans[1]], [dst.ptr], abi.BATCH_PERSISTENT)
    pending = session.submit(plans[2])
    session.close()
    stages = batch.wait()
    if stages[0] != 0 or not np.array_equal(dst.to_host(), wants[1]):
        print("LIFETIME batch after close", flush=True)
        failures += 1
    img2 = pending.wait()
    for im, want in ((image, wants[0]), (img2, wants[2])):
        if im is None or not np.array_equal(im.readback(), want):
            print("LIFETIME image after close", flush=True)
            failures += 1
        if im is not None:
            im.close()
    dst.close()  # the last owner: the session is freed here

    # 2. re-uploads across two server streams under a one-arena budget
    session = HipDecoderSession(0)
    session.set_upload_budget(max(p.output_bytes for p in plans))
    server = HipServer(session, max_wave_images=1, max_waves=2)
    for rnd in range(3):
        for i, p in enumerate(plans):
            server.submit(p, tag=rnd * len(plans) + i)
    server.drain()
    results = server.poll(capacity=64)
    if len(results) != 3 * len(plans):
        print(f"REUPLOAD {len(results)} results", flush=True)
        failures += 1
    for r in results:
        want = wants[r.tag % len(plans)]
        got = r.image.readback() if r.image else None
        if r.status or got is None or not np.array_equal(got[: want.nbytes], want):
            print(f"REUPLOAD mismatch tag {r.tag} status {r.status}", flush=True)
            failures += 1
        if r.image:
            r.image.close()
    evictions = session.diagnostics().upload_evictions
    server.close()
    session.close()
    if evictions == 0:
        print("REUPLOAD budget caused no eviction", flush=True)
        failures += 1
    print(f"emulated lifetime: {failures} failures", flush=True)
    sys.exit(1 if failures else 0)


if __name__ == "__main__":
    main()
