"""The allocation-free batch entry point (jxrhip_decode_batch_static), called
through the C ABI: its output equals jxrhip_submit_batch's persistent path,
byte for byte, on two calls in a row (the second with fewer, smaller
images, so the static scratch holds the first call's contents when the
second starts); it refuses plans whose uploads are not resident and
batches beyond its limits, before any device work."""
import ctypes as C

import numpy as np
import pytest

from cases import CASES, make
from jxr_amd import HipDecodePlan, HipDecoderSession, abi, workloads
from jxr_amd.abi import load

pytestmark = pytest.mark.gpu


def static_decode(session, plans, dsts):
    n = len(plans)
    parr = (C.c_void_p * max(n, 1))(*[p.handle.value for p in plans])
    darr = (C.c_void_p * max(n, 1))(*[int(d) for d in dsts])
    stages = np.zeros(max(n, 1), np.uint32)
    code = load().jxrhip_decode_batch_static(session.handle, parr, n, darr, stages.ctypes.data_as(abi.u32p))
    return code, stages[:n]


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


def corpus(names):
    out = []
    for name in names:
        enc, spec, _ = make(CASES[name])
        out.append((name, enc, HipDecodePlan.from_encoded(enc, spec)))
    return out


def persistent_reference(session, plans):
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert list(stages) == [0] * len(plans)
    return [b.to_host() for b in bufs]


def test_two_calls_in_a_row_equal_the_persistent_batch(session):
    first = corpus(sorted(CASES))
    second = corpus(["rgb8_420_ov1", "gray8_dc_only", "rgb8_444_ov0"])
    # make the second call's images smaller than the first's largest
    assert max(p.output_bytes for _, _, p in second) < max(p.output_bytes for _, _, p in first)
    for batch in (first, second):
        plans = [p for _, _, p in batch]
        want = persistent_reference(session, plans)  # also makes every upload resident
        bufs = [session.allocate_destination(p.output_bytes) for p in plans]
        for b in bufs:
            b.fill(0xA5)
        code, stages = static_decode(session, plans, [b.ptr for b in bufs])
        assert code == abi.OK, load().jxrhip_last_error()
        assert list(stages) == [0] * len(plans)
        for (name, _, _), b, w in zip(batch, bufs, want):
            assert np.array_equal(b.to_host(), w), name


def test_refuses_uploads_that_are_not_resident(session):
    session.clear_cache()
    enc, spec, _ = make(CASES["rgb8_420_ov1"])
    plan = HipDecodePlan.from_encoded(enc, spec)
    buf = session.allocate_destination(plan.output_bytes)
    code, _ = static_decode(session, [plan], [buf.ptr])
    assert code == abi.ERR_INVALID_ARGUMENT
    assert b"resident" in load().jxrhip_last_error()
    session.upload(plan)
    code, stages = static_decode(session, [plan], [buf.ptr])
    assert code == abi.OK and list(stages) == [0]


def test_refuses_batches_beyond_its_limits(session):
    enc, spec, _ = make(CASES["gray8_dc_only"])
    plan = HipDecodePlan.from_encoded(enc, spec)
    session.upload(plan)
    buf = session.allocate_destination(plan.output_bytes)
    code, _ = static_decode(session, [plan] * 4097, [buf.ptr] * 4097)
    assert code == abi.ERR_BUDGET
    assert b"4096" in load().jxrhip_last_error()
