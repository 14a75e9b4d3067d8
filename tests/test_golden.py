"""The committed known-answer vectors of the reference's own unit tests
(tests/golden/reference_vectors.json): the oracle's operators, whole-plane
overlap, chroma upsampling, and the encoder's quality scale."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from jxr_amd import abi, quality_parameters

with open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")) as f:
    GOLDEN = json.load(f)


@pytest.mark.parametrize("v", GOLDEN["operators"], ids=lambda v: f"{v['op']}:{v['source'].split(':')[-1]}")
def test_operator_vectors(v):
    got, err = O.op(v["op"], v["input"])
    assert not err
    assert got.tolist() == v["expected"], v["source"]


@pytest.mark.parametrize("v", GOLDEN["planes"], ids=lambda v: v["source"].split("(")[-1].rstrip(")"))
def test_plane_vectors(v):
    s = np.ascontiguousarray(np.array(v["input"], np.int32))
    cb = np.array(v["column_boundaries"], np.uint32)
    rb = np.array(v["row_boundaries"], np.uint32)
    err = O.lib().jxro_overlap_plane(s.ctypes.data_as(C.POINTER(C.c_int32)), v["width"], v["height"],
                                     cb.ctypes.data_as(C.POINTER(C.c_uint32)), len(cb),
                                     rb.ctypes.data_as(C.POINTER(C.c_uint32)), len(rb), 1 if v["hard"] else 0)
    assert not err
    assert s.tolist() == v["expected"], v["source"]


@pytest.mark.parametrize("v", GOLDEN["upsampling"], ids=lambda v: f"centering{v['centering']}")
def test_upsampling_vectors(v):
    out = (C.c_int32 * 2)()
    assert O.lib().jxro_upsample_pair(v["previous"], v["current"], v["next"], v["centering"], out) == 0
    assert [out[0], out[1]] == v["expected"]


@pytest.mark.parametrize("v", GOLDEN["quality_anchors"]["entries"], ids=lambda v: f"q{v['quality']}")
def test_quality_anchor_vectors(v):
    luma, chroma, ov = quality_parameters(v["quality"], abi.COLOR_YUV444, True, 255)
    assert (list(luma), list(chroma), ov) == (v["luma"], v["chroma"], v["overlap"])


@pytest.mark.parametrize("sampling,color", [("444", abi.COLOR_YUV444), ("422", abi.COLOR_YUV422),
                                            ("420", abi.COLOR_YUV420)])
def test_chroma_offset_vectors(sampling, color):
    """Parameters 5, 16 and 50 are the difference levels of qualities 99, 96
    and 75, which take no DC relief, so their chroma high-pass parameter is
    the difference parameter; 128 lies beyond every quality's level and is
    not reachable through the scale."""
    g = GOLDEN["chroma_offsets"]
    want = dict(zip(g["parameters"], g["by_sampling"][sampling]))
    for level, q in ((5, 99), (16, 96), (50, 75)):
        assert quality_parameters(q, color, True, 255)[1][2] == want[level]


@pytest.mark.parametrize("v", GOLDEN["subsampled_planes"], ids=lambda v: v["source"].split("(")[-1].rstrip(")"))
def test_subsampled_first_level_overlap_vectors(v):
    s = np.ascontiguousarray(np.array(v["input"], np.int32))
    cb = np.array(v["column_boundaries"], np.uint32)
    rb = np.array(v["row_boundaries"], np.uint32)
    err = O.lib().jxro_subsampled_overlap_plane(s.ctypes.data_as(C.POINTER(C.c_int32)), v["width"], v["height"],
                                                cb.ctypes.data_as(C.POINTER(C.c_uint32)), len(cb),
                                                rb.ctypes.data_as(C.POINTER(C.c_uint32)), len(rb),
                                                1 if v["hard"] else 0)
    assert not err
    assert s.tolist() == v["expected"], v["source"]


def _color(name):
    return getattr(abi, "COLOR_" + name)


@pytest.mark.parametrize("v", GOLDEN["conversions"], ids=lambda v: v["source"].split("(")[-1].rstrip(")"))
def test_colour_conversion_vectors(v):
    got, code = O.convert(_color(v["internal"]), _color(v["output"]), v["input"], v["packed_rgb"],
                          v["red_blue_not_swapped"])
    assert code == 0
    assert got == v["expected"], v["source"]


def formatted_row_desc(v):
    """The reference's ComponentPlane rows as constant macroblocks: pixel k of
    the output row lies in macroblock k (the crop takes column 15 of the first
    macroblock and column 0 of the next), so it holds component values[k]."""
    from arena_util import ManualArena, desc_for
    from jxr_amd import OutputSpec
    comps = v["components"]
    n = len(comps[0])
    arena = ManualArena([np.array(c, np.int64).reshape(1, n) for c in comps])
    sep = ManualArena([np.array(v["alpha"], np.int64).reshape(1, n)]) if v["alpha"] is not None else None
    layout = getattr(abi, "LAYOUT_" + v["layout"])
    alpha = sep is not None
    spec = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=layout, output_color=_color(v["output"]),
                      components=len(comps))
    crop = (15, 0, 2, 1) if n == 2 else (0, 0, 1, 1)
    return desc_for(arena, _color(v["internal"]), len(comps), spec, crop,
                    alpha_mode=abi.ALPHA_SEPARATE if alpha else abi.ALPHA_NONE, separate_alpha=sep)


@pytest.mark.parametrize("v", GOLDEN["formatted_rows"], ids=lambda v: v["source"].split("(")[-1].rstrip(")"))
def test_formatted_row_vectors(v):
    assert O.decode(formatted_row_desc(v)).tolist() == v["expected_u8"], v["source"]
