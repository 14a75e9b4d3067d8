"""jxr_amd — MI355X (gfx950) JPEG XR reconstruction backend.

The reconstruction itself is hand-written HIP in csrc/ behind the C ABI of
include/jxr_hip.h; this package is the Python view of that ABI (ctypes),
named after the reference's CUDA backend API. Importing it does not need a
GPU; creating a session does, and fails loudly without one.
"""
from . import abi
from .abi import JxrHipError, load
from .decoder import (
    EncodedImage,
    HipBatch,
    HipDecodePlan,
    HipDecoderSession,
    HipSubmission,
    OutputSpec,
    ResidentHipImage,
    image_desc,
    plan_hip_route,
    quality_parameters,
    reduced_crop,
    sample_scale,
    surface_layout,
)
from .serving import HipServer, ServerResult

__all__ = [
    "abi", "JxrHipError", "load", "EncodedImage", "HipBatch", "HipDecodePlan", "HipDecoderSession",
    "HipSubmission", "OutputSpec", "ResidentHipImage", "image_desc", "plan_hip_route", "quality_parameters",
    "surface_layout", "reduced_crop", "sample_scale", "HipServer", "ServerResult",
]
