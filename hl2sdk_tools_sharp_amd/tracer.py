"""The ray seam, mirrored from the reference: ``IRayTracer`` and ``IGpuTracerFactory``.

Reference interfaces (paths under the reference's ``src/``):

* ``IRayTracer`` -- SourceSharp.MapTools/Tracing/IRayTracer.cs:128-230
  (``TraceVisibilityAsync``, ``TraceClosestAsync``, ``Supports``, ``TracerIdentity``);
* ``Ray`` -- Tracing/IRayTracer.cs:32-88 (seven floats, ``Ray.Segment``);
* ``HitId`` -- Tracing/HitId.cs:70-90 (8 bytes; a miss is (-1, 1.0f));
* ``RayTraceOptions`` -- Tracing/RayTraceOptions.cs:65-145;
* ``IGpuTracerFactory`` / ``GpuTracerOffer`` -- SourceSharp.MapTools/Rad/GpuTracing.cs:28-67;
* ``VulkanRayTracer`` (the backend this one replaces) -- SourceSharp.MapTools.Gpu/VulkanRayTracer.cs.

Names are the reference's in Python spelling; argument meaning and error
behaviour follow it: a short output buffer raises ``ValueError`` (the
reference's ``ArgumentException``), sky pass-through on a closest-hit batch
raises ``ValueError``, options the kernel cannot honour raise
``NotSupportedError`` (``NotSupportedException``), a disposed tracer raises
``ObjectDisposedError``, and a device that cannot prove itself makes the
factory DECLINE with a reason rather than raise.

Every trace runs in libvradhip.so on the GPU.  There is no CPU path here.
"""
from __future__ import annotations

import asyncio
import ctypes
from dataclasses import dataclass
from typing import NamedTuple, Optional

import numpy as np

from . import _lib
from ._lib import VH_DECLINED, VH_ECANCELED, VH_EINVAL, VH_ENOTSUP, VH_ETIMEOUT, VH_OK, VradHipError

TRACE_ID_SKY = 0x01000000  # Rad/TraceId.cs:31
TRACE_ID_OPAQUE = 0x02000000  # Rad/TraceId.cs:34
TRACE_ID_STATIC_PROP = 0x04000000  # Rad/TraceId.cs:37

# Ray wire format: origin xyz, direction xyz, reach (Tracing/IRayTracer.cs:32-39).
RAY_FLOATS = 7
# HitId: 8 bytes, (surface, fraction) (Tracing/HitId.cs:70).
HIT_DTYPE = np.dtype([("surface", "<i4"), ("fraction", "<f4")])
HIT_MISS = -1


class NotSupportedError(Exception):
    """Options this tracer cannot honour (the reference's NotSupportedException)."""


class OperationCanceledError(Exception):
    """A call given up at a cancel request (the reference's
    OperationCanceledException, raised when the CancellationToken fires)."""


class ObjectDisposedError(RuntimeError):
    """The tracer was disposed (the reference's ObjectDisposedException)."""


@dataclass(frozen=True)
class RayTraceOptions:
    """Tracing/RayTraceOptions.cs:65-145."""

    min_distance: float = 0.0
    skip_id: Optional[int] = None
    sky_does_not_block: bool = False
    isolated_rays: bool = False

    @property
    def is_plain(self) -> bool:
        return self.skip_id is None and not self.sky_does_not_block

    @staticmethod
    def stock_exact() -> "RayTraceOptions":
        return RayTraceOptions(0.0)

    @staticmethod
    def self_intersection_safe() -> "RayTraceOptions":
        return RayTraceOptions(1.0e-3)

    @staticmethod
    def test_line(skip_id: int = -1, sky_does_not_block: bool = False) -> "RayTraceOptions":
        return RayTraceOptions(
            0.0, None if skip_id < 0 else skip_id, sky_does_not_block, isolated_rays=True
        )


def ray_segment(start: np.ndarray, end: np.ndarray) -> np.ndarray:
    """``Ray.Segment`` with the exact reciprocal (Tracing/IRayTracer.cs:71-88):
    direction = (end - start) * (1 / len), reach = len, len = sqrt(x*x + y*y + z*z)
    summed left to right in float32."""
    start = np.asarray(start, dtype=np.float32).reshape(-1, 3)
    end = np.asarray(end, dtype=np.float32).reshape(-1, 3)
    d = end - start
    sq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    length = np.sqrt(sq, dtype=np.float32)
    out = np.empty((start.shape[0], RAY_FLOATS), dtype=np.float32)
    out[:, :3] = start
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.float32(1.0) / length
        out[:, 3:6] = d * inv[:, None]
    out[:, 6] = length
    return out


@dataclass
class ShadowCasterSet:
    """The scene handed to the factory (Rad/ShadowCasterSet.cs:66-193):
    ``vertices`` (n, 9) float32, the caller's ``ids`` per triangle
    (Rad/TraceId.cs) and ``flags`` (FCACHETRI_*, Tracing/KdTree.cs:33-37)."""

    vertices: np.ndarray
    ids: np.ndarray
    flags: Optional[np.ndarray] = None

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float32).reshape(-1, 9)
        self.ids = np.ascontiguousarray(self.ids, dtype=np.int32).reshape(-1)
        if self.flags is not None:
            self.flags = np.ascontiguousarray(self.flags, dtype=np.uint8).reshape(-1)
        if self.ids.shape[0] != self.vertices.shape[0]:
            raise ValueError("one id per triangle")

    @property
    def count(self) -> int:
        return self.vertices.shape[0]


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _as_rays(rays) -> np.ndarray:
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    if rays.ndim != 2 or rays.shape[1] not in (6, 7):
        raise ValueError("rays must be an (n, 7) float32 array (origin, direction, reach)")
    return rays


class HipRayTracer:
    """``IRayTracer`` on one MI355X: a BVH2 traversal with the reference's exact
    triangle test.  Created by :class:`HipTracerFactory`; owned by the caller,
    who disposes it once (Rad/GpuTracing.cs:51-58)."""

    def __init__(self, handle: int, scene: ShadowCasterSet):
        self._h = ctypes.c_void_p(handle)
        self._disposed = False
        self._triangle_count = scene.count

    # -- identity and capability ------------------------------------------
    @property
    def tracer_identity(self) -> str:
        self._alive()
        return _lib.lib().vh_tracer_identity(self._h).decode()

    def supports(self, options: RayTraceOptions) -> bool:
        """Skip ids and sky pass-through are honoured in-kernel; only a NaN
        minimum distance is not (``IRayTracer.Supports``)."""
        self._alive()
        skip = -1 if options.skip_id is None else int(options.skip_id)
        return bool(
            _lib.lib().vh_supports(self._h, float(options.min_distance), skip, int(options.sky_does_not_block))
        )

    def load_scene(self, casters: ShadowCasterSet) -> None:
        """Replaces the scene on this context (a new map on the same device
        state: the warm-start network and its data set stay)."""
        self._alive()
        st = _lib.lib().vh_scene_load(self._h, _ptr(casters.vertices), _ptr(casters.ids), _ptr(casters.flags),
                                      casters.count)
        self._raise(st)
        self._triangle_count = casters.count

    @property
    def handle(self) -> ctypes.c_void_p:
        self._alive()
        return self._h

    def scene_info(self):
        t, n, d = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.lib().vh_scene_info(self._h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(d)))
        return {"triangles": t.value, "nodes": n.value, "depth": d.value}

    def stats(self):
        s = _lib.VhStats()
        _lib.check(_lib.lib().vh_get_stats(self._h, ctypes.byref(s)))
        return {"rays_visibility": s.rays_visibility, "rays_closest": s.rays_closest, "calls": s.calls,
                "rays_gather": s.rays_gather}

    # -- the two batch operations ----------------------------------------
    def trace_visibility(self, rays, hit_bits: np.ndarray, options: RayTraceOptions = RayTraceOptions()) -> None:
        """``TraceVisibilityAsync``: bit i of ``hit_bits`` (uint64, LSB first) is
        set when ray i is blocked strictly short of its end."""
        self._alive()
        rays = _as_rays(rays)
        n = rays.shape[0]
        words = (n + 63) // 64
        if hit_bits.dtype != np.uint64 or not hit_bits.flags.c_contiguous:
            raise ValueError("hit_bits must be a contiguous uint64 array")
        if hit_bits.shape[0] < words:
            raise ValueError(f"{n} rays need {words} words of hit bits, and {hit_bits.shape[0]} were given")
        self._require_supported(options)
        if n == 0:
            return
        skip = -1 if options.skip_id is None else int(options.skip_id)
        st = _lib.lib().vh_trace_visibility(
            self._h, _ptr(rays), rays.shape[1], 1.0, n, float(options.min_distance), skip,
            int(options.sky_does_not_block), _ptr(hit_bits), hit_bits.shape[0],
        )
        self._raise(st)

    def trace_closest(self, rays, hits: np.ndarray, options: RayTraceOptions = RayTraceOptions()) -> None:
        """``TraceClosestAsync``: per ray the caller's id of the nearest surface
        and its fraction; a miss writes (-1, 1.0)."""
        self._alive()
        rays = _as_rays(rays)
        n = rays.shape[0]
        self._require_supported(options)
        if options.sky_does_not_block:
            raise ValueError(
                "sky pass-through is a visibility option; a closest-hit trace reports the sky hit itself"
            )
        if hits.dtype != HIT_DTYPE or not hits.flags.c_contiguous:
            raise ValueError("hits must be a contiguous HIT_DTYPE array")
        if hits.shape[0] < n:
            raise ValueError(f"{n} rays need {n} hit slots, and {hits.shape[0]} were given")
        if n == 0:
            return
        surface = np.empty(n, dtype=np.int32)
        fraction = np.empty(n, dtype=np.float32)
        skip = -1 if options.skip_id is None else int(options.skip_id)
        st = _lib.lib().vh_trace_closest(
            self._h, _ptr(rays), rays.shape[1], 1.0, n, float(options.min_distance), skip,
            _ptr(surface), _ptr(fraction), n,
        )
        self._raise(st)
        hits["surface"][:n] = surface
        hits["fraction"][:n] = fraction

    async def trace_visibility_async(self, rays, hit_bits, options=RayTraceOptions()) -> None:
        await asyncio.to_thread(self.trace_visibility, rays, hit_bits, options)

    async def trace_closest_async(self, rays, hits, options=RayTraceOptions()) -> None:
        await asyncio.to_thread(self.trace_closest, rays, hits, options)

    def trace_point_lights(self, points, lights, options: RayTraceOptions = RayTraceOptions()) -> np.ndarray:
        """Shadow rays from every point to every light, generated on the GPU
        (direction = light - point, reach 1, LightRayLog.MakeRay's correct
        mode, Rad/Light/LightRayLog.cs:686-690).  Bit (p * nlights + l)."""
        self._alive()
        points = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
        n = points.shape[0] * lights.shape[0]
        bits = np.zeros((n + 63) // 64, dtype=np.uint64)
        if n == 0:
            return bits
        st = _lib.lib().vh_trace_point_lights(
            self._h, _ptr(points), points.shape[0], _ptr(lights), lights.shape[0],
            float(options.min_distance), _ptr(bits), bits.shape[0],
        )
        self._raise(st)
        return bits

    def set_point_light_method(self, method: int, cells: int = 0) -> None:
        """How :meth:`trace_point_lights` finds candidate triangles:
        ``_lib.VH_POINT_LIGHTS_AUTO`` (default), ``..._BVH`` (wave packet walks)
        or ``..._LIGHT_BUFFER`` (per-light cubes of ``cells`` x ``cells``
        direction cells a face, 0 = 64).  The bits are the same either way."""
        self._alive()
        self._raise(_lib.lib().vh_set_point_light_method(self._h, int(method), int(cells)))

    # -- lifetime ---------------------------------------------------------
    def set_timeout(self, ms: int) -> None:
        """The bound on each wait on the device (VulkanRayTracer's 120 s fence
        bound by default); a call past it raises TimeoutError."""
        self._alive()
        self._raise(_lib.lib().vh_set_timeout(self._h, int(ms)))

    def cancel(self) -> None:
        """A CancellationToken firing: calls in flight raise
        OperationCanceledError at their next wait, later ones at once, until
        :meth:`clear_cancel`.  Safe from any thread while a call runs."""
        self._alive()
        self._raise(_lib.lib().vh_cancel(self._h))

    def clear_cancel(self) -> None:
        self._alive()
        self._raise(_lib.lib().vh_cancel_clear(self._h))

    def dispose(self) -> None:
        """Waits for in-flight work, then releases the device (VulkanRayTracer.Dispose)."""
        if self._disposed:
            return
        self._disposed = True
        _lib.lib().vh_close(self._h)
        self._h = ctypes.c_void_p(None)

    close = dispose

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.dispose()

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass

    # -- helpers ----------------------------------------------------------
    def _alive(self):
        if self._disposed:
            raise ObjectDisposedError("the tracer was disposed")

    def _require_supported(self, options: RayTraceOptions):
        if not self.supports(options):
            raise NotSupportedError(
                f"the HIP kernel cannot trace with minimum distance {options.min_distance}; "
                "trace this batch on the CPU tracer"
            )

    @staticmethod
    def _raise(status: int):
        if status == VH_OK:
            return
        msg = _lib.last_error()
        if status == VH_EINVAL:
            raise ValueError(msg)
        if status == VH_ENOTSUP:
            raise NotSupportedError(msg)
        if status == VH_ETIMEOUT:
            raise TimeoutError(msg)  # the reference's fence bound: TimeoutException
        if status == VH_ECANCELED:
            raise OperationCanceledError(msg)
        raise VradHipError(status, msg)


class GpuTracerOffer(NamedTuple):
    """Rad/GpuTracing.cs:28: the tracer, or why there is none."""

    tracer: Optional[HipRayTracer]
    decline_reason: Optional[str]


class HipTracerFactory:
    """``IGpuTracerFactory`` (Rad/GpuTracing.cs:60-67): builds the GPU scene for
    the casters, or declines with the reason.  The factory keeps no reference to
    what it offers."""

    def __init__(self, device: int = -1, self_test: bool = True, fail_stage: int = 0):
        self.device = device
        self.self_test = self_test
        self.fail_stage = fail_stage

    def try_create(self, casters: ShadowCasterSet) -> GpuTracerOffer:
        try:
            L = _lib.lib()
        except (ImportError, OSError) as e:  # the backend is not built: decline, as a missing package does
            return GpuTracerOffer(None, str(e))
        opts = _lib.VhOpenOpts(self.device, 1 if self.self_test else 0, self.fail_stage, 0)
        h = ctypes.c_void_p()
        st = L.vh_open(ctypes.byref(opts), ctypes.byref(h))
        if st == VH_DECLINED:
            return GpuTracerOffer(None, _lib.last_error())
        if st != VH_OK:
            return GpuTracerOffer(None, f"{_lib.STATUS_NAMES.get(st, st)}: {_lib.last_error()}")
        st = L.vh_scene_load(
            h, _ptr(casters.vertices), _ptr(casters.ids), _ptr(casters.flags), casters.count
        )
        if st != VH_OK:
            reason = _lib.last_error()
            L.vh_close(h)
            if st == VH_EINVAL:
                raise ValueError(reason)
            return GpuTracerOffer(None, reason)
        return GpuTracerOffer(HipRayTracer(h.value, casters), None)

    async def try_create_async(self, casters: ShadowCasterSet) -> GpuTracerOffer:
        return await asyncio.to_thread(self.try_create, casters)


def unpack_bits(bits: np.ndarray, n: int) -> np.ndarray:
    """LSB-first uint64 words -> one bool per ray (Tracing/IRayTracer.cs:150-153)."""
    b = np.unpackbits(np.ascontiguousarray(bits, dtype="<u8").view(np.uint8), bitorder="little")
    return b[:n].astype(bool)

