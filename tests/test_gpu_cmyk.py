"""Four-component and N-component outputs on the GPU, bit-exact against the
CPU oracle (output_format/color.rs:112-123, the YUVK inverse; direct CMYK
and N-component channels as coded, up to thirteen planes in plane groups):
hand-built arenas of full macroblocks with seeded coefficients, since the
forward analysis codes gray and RGB sources only. Also the reference's own
formatted-row vectors (crates/jxr-native/src/output_format/tests.rs:680-782)
against their literal bytes, and the N-component cases in one persistent
batch."""
import numpy as np
import pytest

import oracle_lib as O
from arena_util import FullArena, desc_for
from jxr_amd import HipDecodePlan, HipDecoderSession, OutputSpec, abi

pytestmark = pytest.mark.gpu


def coefficient_planes(n, mbx, mby, seed, dc=600, ac=40):
    """Seeded coefficients: each macroblock's DC at index 0 (the first
    level's DC), every other coefficient small."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        p = rng.integers(-ac, ac + 1, size=(mby, mbx, 256))
        p[:, :, 0] = rng.integers(-dc, dc + 1, size=(mby, mbx))
        out.append(p)
    return out


def spec(storage, depth, layout, color, planar=False, components=3):
    return OutputSpec(storage=storage, depth=depth, layout=layout, output_color=color, planar=planar,
                      components=components)


U8, U16 = (abi.STORE_U8, abi.DEPTH_U8), (abi.STORE_U16, abi.DEPTH_U16)
CASES = {
    "yuvk_to_cmyk_u8": dict(color=abi.COLOR_YUVK, n=4, ov=1, out=spec(*U8, abi.LAYOUT_CMYK, abi.COLOR_CMYK)),
    "yuvk_to_cmyk_u16_planar": dict(color=abi.COLOR_YUVK, n=4, ov=2,
                                    out=spec(*U16, abi.LAYOUT_CMYK, abi.COLOR_CMYK, planar=True)),
    "yuvk_to_cmyk_direct_u8": dict(color=abi.COLOR_YUVK, n=4, ov=0,
                                   out=spec(*U8, abi.LAYOUT_CMYK, abi.COLOR_CMYK_DIRECT)),
    "yuvk_to_cmyka_u8": dict(color=abi.COLOR_YUVK, n=4, ov=1, alpha=True,
                             out=spec(*U8, abi.LAYOUT_CMYKA, abi.COLOR_CMYK)),
    "ncomponent3_u8": dict(color=abi.COLOR_NCOMPONENT, n=3, ov=2,
                           out=spec(*U8, abi.LAYOUT_NCOMPONENT, abi.COLOR_NCOMPONENT, components=3)),
    "ncomponent4_alpha_u16": dict(color=abi.COLOR_NCOMPONENT, n=4, ov=1, alpha=True,
                                  out=spec(*U16, abi.LAYOUT_NCOMPONENT_ALPHA, abi.COLOR_NCOMPONENT,
                                           components=4)),
    # more than four colour planes: plane groups of four, each writing its
    # channels of the pixel
    "ncomponent5_u8": dict(color=abi.COLOR_NCOMPONENT, n=5, ov=1,
                           out=spec(*U8, abi.LAYOUT_NCOMPONENT, abi.COLOR_NCOMPONENT, components=5)),
    "ncomponent5_u8_planar": dict(color=abi.COLOR_NCOMPONENT, n=5, ov=2,
                                  out=spec(*U8, abi.LAYOUT_NCOMPONENT, abi.COLOR_NCOMPONENT, planar=True,
                                           components=5)),
    "ncomponent8_alpha_u16": dict(color=abi.COLOR_NCOMPONENT, n=8, ov=2, alpha=True,
                                  out=spec(*U16, abi.LAYOUT_NCOMPONENT_ALPHA, abi.COLOR_NCOMPONENT,
                                           components=8)),
    "ncomponent6_alpha_premultiplied_u8": dict(color=abi.COLOR_NCOMPONENT, n=6, ov=0, alpha=True,
                                               out=OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8,
                                                              layout=abi.LAYOUT_NCOMPONENT_ALPHA,
                                                              output_color=abi.COLOR_NCOMPONENT, components=6,
                                                              premultiply=True)),
    "ncomponent13_i16": dict(color=abi.COLOR_NCOMPONENT, n=13, ov=1,
                             out=OutputSpec(storage=abi.STORE_I16, depth=abi.DEPTH_I16, shift_bits=1,
                                            layout=abi.LAYOUT_NCOMPONENT, output_color=abi.COLOR_NCOMPONENT,
                                            components=13)),
}


@pytest.fixture(scope="module")
def session():
    s = HipDecoderSession(0)
    yield s
    s.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_four_and_n_component_outputs_match_oracle(session, name):
    c = CASES[name]
    mbx, mby = 5, 3
    arena = FullArena(coefficient_planes(c["n"], mbx, mby, seed=len(name)), overlap=c["ov"])
    alpha = FullArena(coefficient_planes(1, mbx, mby, seed=99), overlap=c["ov"]) if c.get("alpha") else None
    crop = (3, 5, mbx * 16 - 7, mby * 16 - 9)
    kw = dict(alpha_mode=abi.ALPHA_SEPARATE, separate_alpha=alpha) if alpha is not None else {}
    desc = desc_for(arena, c["color"], c["n"], c["out"], crop, **kw)
    want = O.decode(desc)
    got = session.decode_to_host(HipDecodePlan(desc, keepalive=(arena, alpha)))
    assert np.array_equal(got, want), name


def _case_desc(name):
    c = CASES[name]
    mbx, mby = 5, 3
    arena = FullArena(coefficient_planes(c["n"], mbx, mby, seed=len(name)), overlap=c["ov"])
    alpha = FullArena(coefficient_planes(1, mbx, mby, seed=99), overlap=c["ov"]) if c.get("alpha") else None
    crop = (3, 5, mbx * 16 - 7, mby * 16 - 9)
    kw = dict(alpha_mode=abi.ALPHA_SEPARATE, separate_alpha=alpha) if alpha is not None else {}
    return desc_for(arena, c["color"], c["n"], c["out"], crop, **kw), (arena, alpha)


def test_n_component_cases_in_one_persistent_batch(session):
    names = [n for n in sorted(CASES) if n.startswith("ncomponent")]
    descs = [_case_desc(n) for n in names]
    plans = [HipDecodePlan(d, keepalive=k) for d, k in descs]
    bufs = [session.allocate_destination(p.output_bytes) for p in plans]
    for b in bufs:
        b.fill(0x77)
    stages = session.submit_batch(plans, [b.ptr for b in bufs], abi.BATCH_PERSISTENT).wait()
    assert list(stages) == [0] * len(plans)
    for name, (d, _), b in zip(names, descs, bufs):
        assert np.array_equal(b.to_host(), O.decode(d)), name


def test_reference_formatted_rows_on_gpu(session):
    import test_golden
    for v in test_golden.GOLDEN["formatted_rows"]:
        d = test_golden.formatted_row_desc(v)
        got = session.decode_to_host(HipDecodePlan(d, keepalive=d._keep))
        assert got.tolist() == v["expected_u8"], v["source"]
