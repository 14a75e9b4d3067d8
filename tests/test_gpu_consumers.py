"""The consumer half of the boundary on the GPU (CudaSubmission::
{enqueue_consumer_wait, output_layout, pending_device_buffer},
crates/jxr-cuda/src/submission.rs:50-93): a stream the caller created
waits for a pending submission (or batch) on the device, a copy of the
pending output runs on it with no host wait on the decode, and the copied
bytes equal the oracle's. Also the same path into a torch tensor, in a
process where torch and this library share one HIP runtime."""
import ctypes as C
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import oracle_lib as O
from cases import CASES, make
from jxr_amd import HipDecodePlan, HipDecoderSession, abi, image_desc, synth, workloads, EncodedImage
from jxr_amd.abi import JxrHipError

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def hip_runtime():
    """The HIP runtime this library is bound to (already loaded by it)."""
    abi.load()
    hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return hip


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


@pytest.fixture(scope="module")
def stream():
    hip = hip_runtime()
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    yield hip, st.value
    hip.hipStreamDestroy(st)


def big_case():
    px = synth.gradient_noise(2048, 1536, 11, depth=8, sigma=8.0)
    enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=2, scaled=True, qp=(20, 20, 24))
    return enc, workloads.rgb_out(8, False)


def test_consumer_stream_reads_the_pending_output(session, stream):
    hip, st = stream
    enc, spec = big_case()
    plan = HipDecodePlan.from_encoded(enc, spec)
    want = O.decode(image_desc(enc, spec))
    sub = session.submit(plan)
    lay = sub.output_layout()
    assert (lay.width, lay.height, lay.byte_len) == (2048, 1536, plan.output_bytes)
    ptr, n = sub.pending_device_ptr()
    assert ptr and n == plan.output_bytes
    copy = session.allocate_destination(n)
    sub.enqueue_consumer_wait(st)
    assert hip.hipMemcpyAsync(copy.ptr, ptr, n, 3, st) == 0  # device to device, on the consumer stream
    assert hip.hipStreamSynchronize(st) == 0
    assert np.array_equal(copy.to_host(), want)
    img = sub.wait()
    assert np.array_equal(img.readback(), want)
    img.close()


def test_pending_pointer_is_refused_for_a_caller_destination(session):
    enc, spec, _ = make(CASES["rgb8_420_ov1"])
    plan = HipDecodePlan.from_encoded(enc, spec)
    dst = session.allocate_destination(plan.output_bytes)
    sub = session.submit_into(plan, dst.ptr, plan.output_bytes)
    with pytest.raises(JxrHipError):
        sub.pending_device_ptr()
    sub.wait()


@pytest.mark.parametrize("mode", [abi.BATCH_PERSISTENT, abi.BATCH_PER_IMAGE])
def test_consumer_stream_waits_for_a_batch(session, stream, mode):
    hip, st = stream
    names = ["rgb8_420_ov1", "gray16_ov2_q75", "rgb8_444_ov2_hard", "rgb16_420_ov2"]
    encs = [make(CASES[n]) for n in names]
    plans = [HipDecodePlan.from_encoded(e, sp) for e, sp, _ in encs]
    wants = [O.decode(image_desc(e, sp)) for e, sp, _ in encs]
    dsts = [session.allocate_destination(p.output_bytes) for p in plans]
    copies = [session.allocate_destination(p.output_bytes) for p in plans]
    batch = session.submit_batch(plans, [d.ptr for d in dsts], mode)
    batch.enqueue_consumer_wait(st)
    for d, c, p in zip(dsts, copies, plans):
        assert hip.hipMemcpyAsync(c.ptr, d.ptr, p.output_bytes, 3, st) == 0
    assert hip.hipStreamSynchronize(st) == 0
    for n, c, w in zip(names, copies, wants):
        assert np.array_equal(c.to_host(), w), n
    assert list(batch.wait()) == [0] * len(plans)


TORCH_CONSUMER = textwrap.dedent('''
    import sys
    sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
    import numpy as np
    import torch  # first: the library then binds to the HIP runtime torch loaded
    import oracle_lib as O
    from jxr_amd import EncodedImage, HipDecodePlan, HipDecoderSession, abi, image_desc, synth, workloads
    abi.load()
    maps = [l for l in open("/proc/self/maps") if "libamdhip64" in l]
    paths = sorted({{l.split()[-1] for l in maps}})
    print("[torch consumer] HIP runtimes loaded:", paths, flush=True)
    if len(paths) != 1:
        print("SKIP two HIP runtimes", flush=True)
        sys.exit(0)
    px = synth.gradient_noise(640, 480, 5, depth=8, sigma=8.0)
    enc = EncodedImage(px, color=abi.COLOR_YUV420, overlap=1, scaled=True, qp=(8, 8, 12))
    spec = workloads.rgb_out(8, True)  # planar NCHW
    plan = HipDecodePlan.from_encoded(enc, spec)
    out = torch.empty(plan.output_bytes, dtype=torch.uint8, device="cuda")
    session = HipDecoderSession(0)
    sub = session.submit_into(plan, out.data_ptr(), plan.output_bytes)
    ts = torch.cuda.current_stream()
    sub.enqueue_consumer_wait(ts.cuda_stream)
    chw = out.view(3, 480, 640).to(torch.int32) * 2  # torch work ordered after the decode
    got = (chw // 2).to(torch.uint8).cpu().numpy().reshape(-1)
    sub.wait()
    want = O.decode(image_desc(enc, spec))
    assert np.array_equal(got, want), int((got != want).sum())
    session.close()
    print("OK torch consumer", flush=True)
''')


def test_torch_tensor_as_destination_and_consumer_stream():
    code = TORCH_CONSUMER.format(tests=HERE, root=os.path.dirname(HERE))
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    print(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    if "SKIP" in res.stdout:
        pytest.skip("torch and this library load different HIP runtimes")
    assert "OK torch consumer" in res.stdout
