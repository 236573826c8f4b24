"""hl2sdk_tools_sharp_amd -- vrad's lighting hot path on one MI355X (gfx950).

The compute lives in ``libvradhip.so`` (hand-written HIP, C ABI in
``include/vradhip.h``); this package is the host side that mirrors the
reference's C# seams (``IRayTracer``/``IGpuTracerFactory``, and the bounce
stage) so the library can be driven and tested from Python.
"""
from .tracer import (  # noqa: F401
    HIT_DTYPE,
    TRACE_ID_OPAQUE,
    TRACE_ID_SKY,
    GpuTracerOffer,
    HipRayTracer,
    HipTracerFactory,
    NotSupportedError,
    ObjectDisposedError,
    OperationCanceledError,
    RayTraceOptions,
    ShadowCasterSet,
    ray_segment,
    unpack_bits,
)

__all__ = [
    "HIT_DTYPE",
    "TRACE_ID_OPAQUE",
    "TRACE_ID_SKY",
    "GpuTracerOffer",
    "HipRayTracer",
    "HipTracerFactory",
    "NotSupportedError",
    "ObjectDisposedError",
    "OperationCanceledError",
    "RayTraceOptions",
    "ShadowCasterSet",
    "ray_segment",
    "unpack_bits",
]
