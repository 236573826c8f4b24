"""CPU oracle for the vrad lighting hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package, and only as the checker (or the timed CPU
restatement of the reference's method).  The product package
``hl2sdk_tools_sharp_amd`` never imports it.

The ray functions live in ``libvradcpu.so`` (``vradcpu.c``, plain C built with
``-ffp-contract=off``); the bounce, gather and encode restatements are numpy
modules beside this file.  Each function cites the reference file:line it
restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvradcpu.so")

TRI_DTYPE = np.dtype(
    [
        ("n", "<f4", 3),
        ("d", "<f4"),
        ("id", "<i4"),
        ("e", "<f4", 6),
        ("cs0", "u1"),
        ("cs1", "u1"),
        ("flags", "u1"),
        ("unused", "u1"),
    ]
)
assert TRI_DTYPE.itemsize == 48

_lib = None


def build() -> str:
    """Compile libvradcpu.so with make (gcc only; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, f32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    L.vc_tri_convert.argtypes = [P, P, P, i32, P]
    L.vc_tri_convert.restype = ctypes.c_int
    L.vc_trace_closest_brute.argtypes = [P, i32, P, i32, f32, i64, f32, i32, i32, P, P]
    L.vc_trace_visibility_brute.argtypes = [P, i32, P, i32, f32, i64, f32, i32, i32, P]
    L.vc_bvh_build.argtypes = [P, P, i32]
    L.vc_bvh_build.restype = P
    L.vc_bvh_free.argtypes = [P]
    L.vc_bvh_node_count.argtypes = [P]
    L.vc_bvh_node_count.restype = i32
    L.vc_bvh_visibility.argtypes = [P, P, i32, f32, i64, f32, i32, i32, P, i32]
    L.vc_bvh_closest.argtypes = [P, P, i32, f32, i64, f32, i32, P, P, i32]
    L.vc_bvh_point_lights.argtypes = [P, P, i64, P, i32, f32, P, i32]
    _lib = L
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _rays(rays: np.ndarray):
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    if rays.ndim != 2 or rays.shape[1] not in (6, 7):
        raise ValueError("rays must be (n, 7) or (n, 6) float32")
    return rays, rays.shape[1]


def convert_triangles(v9: np.ndarray, ids=None, flags=None) -> np.ndarray:
    """KdTreeBuilder.ToIntersectionFormat, Correct side
    (src/SourceSharp.MapTools/Tracing/KdTreeBuilder.cs:1568-1612)."""
    v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
    n = v9.shape[0]
    out = np.zeros(n, dtype=TRI_DTYPE)
    ids_a = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
    fl_a = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
    lib().vc_tri_convert(
        _p(v9), None if ids_a is None else _p(ids_a), None if fl_a is None else _p(fl_a), n, _p(out)
    )
    return out


def closest_brute(tris: np.ndarray, rays: np.ndarray, tmin=0.0, skip_id=-1, clip=True, shared_reach=1.0):
    """Nearest hit per ray by brute force (triangle INDEX, t); -1 = miss."""
    rays, rw = _rays(rays)
    n = rays.shape[0]
    idx = np.empty(n, dtype=np.int32)
    t = np.empty(n, dtype=np.float32)
    lib().vc_trace_closest_brute(
        _p(tris), len(tris), _p(rays), rw, shared_reach, n, tmin, skip_id, 1 if clip else 0, _p(idx), _p(t)
    )
    return idx, t


def visibility_brute(tris, rays, tmin=0.0, skip_id=-1, sky_passes=False, shared_reach=1.0):
    rays, rw = _rays(rays)
    n = rays.shape[0]
    bits = np.zeros((n + 63) // 64, dtype=np.uint64)
    lib().vc_trace_visibility_brute(
        _p(tris), len(tris), _p(rays), rw, shared_reach, n, tmin, skip_id, 1 if sky_passes else 0, _p(bits)
    )
    return bits


class CpuBvh:
    """The oracle's CPU BVH: the exact triangle test behind conservative boxes."""

    def __init__(self, v9: np.ndarray, tris: np.ndarray):
        self._v9 = np.ascontiguousarray(v9, dtype=np.float32).reshape(-1, 9)
        self._tris = np.ascontiguousarray(tris)
        self._h = lib().vc_bvh_build(_p(self._v9), _p(self._tris), len(self._tris))

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.vc_bvh_free(h)

    @property
    def node_count(self) -> int:
        return lib().vc_bvh_node_count(self._h)

    def visibility(self, rays, tmin=0.0, skip_id=-1, sky_passes=False, shared_reach=1.0, threads=None):
        rays, rw = _rays(rays)
        n = rays.shape[0]
        bits = np.zeros((n + 63) // 64, dtype=np.uint64)
        lib().vc_bvh_visibility(
            self._h, _p(rays), rw, shared_reach, n, tmin, skip_id, 1 if sky_passes else 0, _p(bits),
            threads or os.cpu_count() or 1,
        )
        return bits

    def closest(self, rays, tmin=0.0, skip_id=-1, shared_reach=1.0, threads=None):
        rays, rw = _rays(rays)
        n = rays.shape[0]
        idx = np.empty(n, dtype=np.int32)
        t = np.empty(n, dtype=np.float32)
        lib().vc_bvh_closest(
            self._h, _p(rays), rw, shared_reach, n, tmin, skip_id, _p(idx), _p(t), threads or os.cpu_count() or 1
        )
        return idx, t

    def point_lights(self, points, lights, tmin=0.0, threads=None):
        points = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 3)
        n = points.shape[0] * lights.shape[0]
        bits = np.zeros((n + 63) // 64, dtype=np.uint64)
        lib().vc_bvh_point_lights(
            self._h, _p(points), points.shape[0], _p(lights), lights.shape[0], tmin, _p(bits),
            threads or os.cpu_count() or 1,
        )
        return bits


def unpack_bits(bits: np.ndarray, n: int) -> np.ndarray:
    """LSB-first uint64 words -> bool per ray (Tracing/IRayTracer.cs:150-153)."""
    b = np.unpackbits(np.ascontiguousarray(bits, dtype="<u8").view(np.uint8), bitorder="little")
    return b[:n].astype(bool)
