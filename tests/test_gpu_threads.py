"""One session used from many threads at once (the reference's session is
shared across threads, crates/jxr-cuda/src/runtime.rs:35-43): concurrent
per-image submissions, batches and uploads from eight threads give every
image's bytes exactly."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as O
from cases import CASES, make
from jxr_amd import HipDecodePlan, HipDecoderSession, abi, image_desc

pytestmark = pytest.mark.gpu


def test_concurrent_submissions_from_eight_threads():
    names = sorted(CASES)
    items = [make(CASES[n])[:2] for n in names]
    wants = [O.decode(image_desc(e, sp)) for e, sp in items]
    s = HipDecoderSession(0)
    try:
        plans = [HipDecodePlan.from_encoded(e, sp) for e, sp in items]

        def worker(k):
            rng = np.random.default_rng(k)
            bad = []
            for _ in range(12):
                i = int(rng.integers(len(plans)))
                mode = int(rng.integers(3))
                if mode == 0:
                    got = s.decode_to_host(plans[i])
                    if not np.array_equal(got, wants[i]):
                        bad.append((names[i], "decode_to_host"))
                elif mode == 1:
                    buf = s.allocate_destination(plans[i].output_bytes)
                    s.submit_into(plans[i], buf.ptr, buf.nbytes).wait()
                    if not np.array_equal(buf.to_host(), wants[i]):
                        bad.append((names[i], "submit_into"))
                    buf.close()
                else:
                    js = [int(j) for j in rng.integers(len(plans), size=3)]
                    bufs = [s.allocate_destination(plans[j].output_bytes) for j in js]
                    b = s.submit_batch([plans[j] for j in js], [x.ptr for x in bufs],
                                       abi.BATCH_PERSISTENT if k % 2 else abi.BATCH_PER_IMAGE)
                    st = b.wait()
                    for j, x, code in zip(js, bufs, st):
                        if code or not np.array_equal(x.to_host(), wants[j]):
                            bad.append((names[j], "batch"))
                        x.close()
            return bad

        with ThreadPoolExecutor(8) as ex:
            bad = [b for r in ex.map(worker, range(8)) for b in r]
        assert not bad, bad[:5]
    finally:
        s.close()
