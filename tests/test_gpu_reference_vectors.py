"""The reference backend's own end-to-end vectors, run on the GPU against
their literal expected bytes (crates/jxr-cuda/src/plan/tests.rs:16-35): a
16x16 DC-only luma macroblock of zero coefficients decodes to U8 128
everywhere, and a DC coefficient of i32::MAX at DC step 2 fails the first
transform (stage 1) without poisoning the next submission, which again
decodes to 128s. Both through the per-image and the persistent path."""
import numpy as np
import pytest

from arena_util import ManualArena, desc_for
from jxr_amd import HipDecodePlan, HipDecoderSession, OutputSpec, abi
from jxr_amd.abi import JxrHipError

pytestmark = pytest.mark.gpu

LUMA8 = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


def luma_plan(dc=0, dc_step=1):
    arena = ManualArena([np.zeros((1, 1), np.int64)])
    arena.coef[0] = dc
    arena.quant[0] = dc_step
    return HipDecodePlan(desc_for(arena, abi.COLOR_LUMA, 1, LUMA8, (0, 0, 16, 16)), keepalive=arena)


def test_dc_only_zero_luma_decodes_to_128(session):
    got = session.decode_to_host(luma_plan())
    assert got.tolist() == [128] * 256


def test_overflow_reports_stage_one_and_does_not_poison_the_next_submission(session):
    with pytest.raises(JxrHipError) as e:
        session.decode_to_host(luma_plan(dc=2**31 - 1, dc_step=2))
    assert e.value.code == abi.ERR_KERNEL_ARITHMETIC and e.value.stage == 1
    assert session.decode_to_host(luma_plan()).tolist() == [128] * 256


def test_same_vectors_in_one_persistent_batch(session):
    plans = [luma_plan(), luma_plan(dc=2**31 - 1, dc_step=2), luma_plan()]
    bufs = [session.allocate_destination(256) for _ in plans]
    for b in bufs:
        b.fill(0x11)
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait(raise_on_error=False)
    assert list(stages) == [0, 1, 0]
    assert bufs[0].to_host().tolist() == [128] * 256
    assert bufs[2].to_host().tolist() == [128] * 256
