"""Four-component and N-component outputs on the GPU, bit-exact against the
CPU oracle (output_format/color.rs:112-123, the YUVK inverse; direct CMYK
and N-component channels as coded): hand-built arenas of full macroblocks
with seeded coefficients, since the forward analysis codes gray and RGB
sources only."""
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
