"""HIP reconstruction vs the CPU oracle, byte for byte, through the C ABI.

Every case decodes the same arenas on the GPU (plain per-image path, several
tile shapes, and the persistent batch path) and on the CPU oracle; outputs
must be identical bytes (integer, fixed-point, half and float formats are
all exact bit patterns)."""
import ctypes as C

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


def _expected(enc, spec):
    return O.decode(image_desc(enc, spec))


@pytest.mark.parametrize("name", sorted(CASES))
def test_plain_path_matches_oracle(session, name):
    enc, spec, _ = make(CASES[name])
    plan = HipDecodePlan.from_encoded(enc, spec)
    want = _expected(enc, spec)
    got = session.decode_to_host(plan)
    assert got.nbytes == want.nbytes
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{name}: {bad.size} bytes differ, first at {bad[:8]}"


@pytest.mark.parametrize("tile", [(1, 1), (2, 1), (1, 3), (3, 2), (4, 4), (5, 3)])
def test_tile_shapes_agree(session, tile):
    names = ["rgb8_444_ov2_hard", "rgb8_420_ov2_hard_1mb", "rgb8_422_ov2_hard", "gray16_ov2_hard",
             "rgbaf_separate_ov1", "rgb8_region_crop"]
    session.set_tile(*tile)
    try:
        for name in names:
            enc, spec, _ = make(CASES[name])
            plan = HipDecodePlan.from_encoded(enc, spec)
            want = _expected(enc, spec)
            got = session.decode_to_host(plan)
            assert np.array_equal(got, want), f"{name} at tile {tile}"
    finally:
        session.set_tile(0, 0)




@pytest.mark.parametrize("mode", [abi.BATCH_PER_IMAGE, abi.BATCH_PERSISTENT])
def test_batch_matches_oracle(session, mode):
    names = sorted(CASES)
    encs, plans, wants = [], [], []
    for name in names:
        enc, spec, _ = make(CASES[name])
        encs.append(enc)
        plans.append(HipDecodePlan.from_encoded(enc, spec))
        wants.append(_expected(enc, spec))
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    for b in bufs:
        b.fill(0xA5)
    stages = session.submit_batch(plans, [b.ptr for b in bufs], mode).wait()
    assert not stages.any()
    for name, b, want in zip(names, bufs, wants):
        got = b.to_host(0, want.nbytes)
        assert np.array_equal(got, want), name


def test_dense_nchw_batch(session):
    """Many images into one device allocation (NCHW planes per image)."""
    names = ["rgb8_420_nchw", "rgb8_420_nchw", "rgb8_420_nchw"]
    encs, plans, wants, offs = [], [], [], []
    total = 0
    for i, name in enumerate(names):
        case = dict(CASES[name], seed=10 + i)
        enc, spec, _ = make(case)
        plan = HipDecodePlan.from_encoded(enc, spec)
        encs.append(enc)
        plans.append(plan)
        wants.append(_expected(enc, spec))
        offs.append(total)
        total += plan.output_bytes
    dense = session.allocate_destination(total)
    dense.fill(0)
    base = dense.ptr
    session.submit_batch(plans, [base + o for o in offs], abi.BATCH_PERSISTENT).wait()
    host = dense.to_host()
    for o, want in zip(offs, wants):
        assert np.array_equal(host[o:o + want.nbytes], want)


def test_repeat_submissions_hit_upload_cache(session):
    enc, spec, _ = make(CASES["rgb8_420_ov1"])
    plan = HipDecodePlan.from_encoded(enc, spec)
    want = _expected(enc, spec)
    before = session.diagnostics().upload_hits
    for _ in range(3):
        assert np.array_equal(session.decode_to_host(plan), want)
    assert session.diagnostics().upload_hits >= before + 2


def _overflowing_case():
    enc, spec, _ = make(dict(w=32, h=32, ov=1, quality=100))
    arena = enc.primary.contents
    # a DC coefficient whose dequantized value leaves i32 (plan/tests.rs:16-36)
    arena.coefficients[arena.coefficient_offsets[0]] = 2**31 - 1
    arena.quantizers[0] = 2
    return enc, spec


def test_overflow_reports_stage_and_does_not_poison_next(session):
    enc, spec = _overflowing_case()
    code, stage, _ = O.decode_status(image_desc(enc, spec))
    assert code == abi.ERR_KERNEL_ARITHMETIC and stage == abi.STAGE_FIRST_TRANSFORM
    plan = HipDecodePlan.from_encoded(enc, spec)
    with pytest.raises(JxrHipError) as err:
        session.decode_to_host(plan)
    assert err.value.code == abi.ERR_KERNEL_ARITHMETIC and err.value.stage == abi.STAGE_FIRST_TRANSFORM
    # the next valid submission still decodes
    enc2, spec2, _ = make(CASES["c1_rgb8_444_lossless_ov1"])
    plan2 = HipDecodePlan.from_encoded(enc2, spec2)
    assert np.array_equal(session.decode_to_host(plan2), _expected(enc2, spec2))


def test_batch_sibling_errors_are_indexed(session):
    enc_bad, spec_bad = _overflowing_case()
    enc_ok, spec_ok, _ = make(CASES["rgb8_420_ov1"])
    plans = [HipDecodePlan.from_encoded(enc_ok, spec_ok), HipDecodePlan.from_encoded(enc_bad, spec_bad),
             HipDecodePlan.from_encoded(enc_ok, spec_ok)]
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait(
        raise_on_error=False)
    assert list(stages) == [0, abi.STAGE_FIRST_TRANSFORM, 0]
    want = _expected(enc_ok, spec_ok)
    assert np.array_equal(bufs[0].to_host(), want)
    assert np.array_equal(bufs[2].to_host(), want)


@pytest.mark.gpu
@pytest.mark.parametrize("index", range(9))
def test_large_samples_take_checked_output_path(session, index):
    """Samples above 2^27 run the output stage with checked arithmetic;
    results and the output-stage overflow match the oracle."""
    import large_cases
    name, desc, stage = large_cases.cases()[index]
    plan = HipDecodePlan(desc, keepalive=desc._keep)
    if stage:
        with pytest.raises(JxrHipError) as err:
            session.decode_to_host(plan)
        assert err.value.code == abi.ERR_KERNEL_ARITHMETIC and err.value.stage == stage, name
    else:
        assert np.array_equal(session.decode_to_host(plan), O.decode(desc)), name


@pytest.mark.parametrize("color", [abi.COLOR_YUV444, abi.COLOR_YUV420, abi.COLOR_YUV422])
@pytest.mark.parametrize("ov", [0, 1, 2])
def test_soft_tiles_do_not_change_samples_on_gpu(session, color, ov):
    """The same image with and without a soft tile grid decodes to the same
    bytes on the GPU, and to the oracle's."""
    base = dict(w=176, h=144, ov=ov, quality=70, color=color)
    e1, s1, _ = make(base)
    e2, s2, _ = make(dict(base, tiles=(3, 2)))
    a = session.decode_to_host(HipDecodePlan.from_encoded(e1, s1))
    b = session.decode_to_host(HipDecodePlan.from_encoded(e2, s2))
    assert np.array_equal(a, b)
    assert np.array_equal(a, _expected(e1, s1))
