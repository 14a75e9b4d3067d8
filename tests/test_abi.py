"""The C-ABI library loads without a GPU, exports every function the public
headers declare, and its ctypes structures have the C layout (gcc compiles
the headers and prints every size and offset)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from jxr_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("jxr_hip.h", "jxr_hip_encode.h", "jxr_hip_codestream.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(jxrhip_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared_functions() <= set(abi.SIGNATURES), sorted(declared_functions() - set(abi.SIGNATURES))


def test_library_reports_version_and_routes_without_a_device():
    lib = abi.load()
    assert lib.jxrhip_abi_version() == abi.ABI_VERSION
    use = C.c_int32()
    assert lib.jxrhip_route(100000, 0, 0, C.byref(use)) == abi.OK and use.value == 0
    assert lib.jxrhip_route(100000, 0, 1, C.byref(use)) == abi.OK and use.value == 1
    assert lib.jxrhip_route(100, 0, 1, C.byref(use)) == abi.OK and use.value == 0
    assert lib.jxrhip_route(100, 1, 0, C.byref(use)) == abi.ERR_UNAVAILABLE
    assert lib.jxrhip_route_threshold() == 16384


STRUCTS = {
    "jxrhip_coefficient_plane": abi.CoefficientPlane,
    "jxrhip_arena": abi.Arena,
    "jxrhip_output_format": abi.OutputFormat,
    "jxrhip_surface_plane": abi.SurfacePlane,
    "jxrhip_surface_layout": abi.SurfaceLayout,
    "jxrhip_image_desc": abi.ImageDesc,
    "jxrhip_diagnostics": abi.Diagnostics,
    "jxrhip_encode_params": abi.EncodeParams,
    "jxrhip_tile_choice": abi.TileChoice,
    "jxrhip_server_config": abi.ServerConfig,
    "jxrhip_server_result": abi.ServerResult,
    "jxrhip_server_stats": abi.ServerStats,
    "jxrhip_valu_rate": abi.ValuRate,
    "jxrhip_stage_report": abi.StageReport,
    "jxrhip_decode_report": abi.DecodeReport,
    "jxrhip_codestream_info": abi.CodestreamInfo,
    "jxrhip_codestream_options": abi.CodestreamOptions,
}


def test_ctypes_layouts_match_the_c_headers(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADERS[2]}"', "int main(void){"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(py, field).offset, f"{cname}.{field}"


def test_session_creation_fails_loudly_without_a_device():
    if abi.load().jxrhip_is_available():
        pytest.skip("a device is present")
    from jxr_amd import HipDecoderSession, JxrHipError
    with pytest.raises(JxrHipError) as e:
        HipDecoderSession(0)
    assert e.value.code == abi.ERR_UNAVAILABLE


def test_decode_report_names_cpu_and_device_stages():
    # CudaDecodePlan::decode_report (crates/jxr-cuda/src/plan.rs:214-253):
    # five CPU stages, eight device stages, then the readback when asked
    import numpy as np
    from arena_util import ManualArena, desc_for
    from jxr_amd import HipDecodePlan, OutputSpec
    a = ManualArena([np.zeros((1, 1), np.int64)])
    spec = OutputSpec(storage=abi.STORE_U8, depth=abi.DEPTH_U8, layout=abi.LAYOUT_LUMA, output_color=abi.COLOR_LUMA)
    plan = HipDecodePlan(desc_for(a, abi.COLOR_LUMA, 1, spec, (0, 0, 16, 16)), keepalive=a)
    for readback in (False, True):
        r = plan.decode_report(abi.BACKEND_AUTO, host_readback=readback)
        assert r["requested"] == "Auto" and r["selected"] == "Hip" and r["fallback"] is None
        want = [(st, "CpuScalar") for st in abi.DECODE_STAGES[:5]] + [(st, "Hip") for st in abi.DECODE_STAGES[5:13]]
        if readback:
            want.append(("HostReadback", "CpuScalar"))
        assert r["stages"] == want
