"""The leaf-ambient ray fan on the GPU.

AmbientSampler.ComputeRayCube (SourceSharp.MapTools/Rad/Ambient/AmbientSampler.cs)
for one light style: 162 rays from each sample point along the vertex
normals, each ray's light from the lightmap lump where it lands
(RayAmbientLighting.Accumulate, Rad/Ambient/RayAmbientLighting.cs), projected
onto the six box directions (AmbientCube.Project, Rad/Ambient/AmbientCube.cs).
The lump is the final stage's, read where it lies on the device, or one the
caller passes.  The world lights LeafAmbientSurfaceLights.Classify flags
(dim surface lights of style 0) are added where the segment to them is open
(AmbientCube.AddEmitSurfaceLights); leaf sampling and the lump of cubes stay
with the caller.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .tracer import HipRayTracer

# vh_ambient_face
FACE_DTYPE = np.dtype([
    ("s", "<f4", 4), ("t", "<f4", 4), ("min_s", "<f4"), ("min_t", "<f4"), ("smax", "<i4"),
    ("tmax", "<i4"), ("light_ofs", "<i4"), ("sky", "<i4"), ("reflectivity", "<f4", 3),
])


# vh_ambient_light
LIGHT_DTYPE = np.dtype([
    ("origin", "<f4", 3), ("normal", "<f4", 3), ("intensity", "<f4", 3), ("radius", "<f4"),
    ("type", "<i4"), ("style", "<i4"),
])


class VhAmbientDesc(ctypes.Structure):
    _fields_ = [("nfaces", ctypes.c_int32), ("faces", ctypes.c_void_p), ("has_sky_ambient", ctypes.c_int32),
                ("sky_ambient", ctypes.c_float * 3), ("lump", ctypes.c_void_p), ("lump_bytes", ctypes.c_int64),
                ("nlights", ctypes.c_int32), ("lights", ctypes.c_void_p)]


def directions() -> np.ndarray:
    """The fan's 162 directions as the kernels use them."""
    out = np.zeros((162, 3), np.float32)
    HipRayTracer._raise(_lib.lib().vh_ambient_directions(out.ctypes.data_as(ctypes.c_void_p)))
    return out


class HipLeafAmbient:
    """ComputeRayCube for batches of sample points against the tracer's scene."""

    def __init__(self, tracer: HipRayTracer):
        self.tracer = tracer
        self._keep = None

    def load(self, faces: np.ndarray, sky_ambient=None, lump: Optional[np.ndarray] = None,
             lights: Optional[np.ndarray] = None) -> None:
        """faces: FACE_DTYPE per face (the scene's ids carry the face index);
        lump None reads the last final-light lump on the device; lights: the
        pass's world lights (LIGHT_DTYPE), classified here."""
        fa = np.ascontiguousarray(faces, FACE_DTYPE)
        li = None if lights is None else np.ascontiguousarray(lights, LIGHT_DTYPE)
        lp = None if lump is None else np.ascontiguousarray(lump, np.uint8)
        d = VhAmbientDesc()
        d.nfaces = len(fa)
        d.faces = fa.ctypes.data if len(fa) else None
        d.has_sky_ambient = 0 if sky_ambient is None else 1
        if sky_ambient is not None:
            d.sky_ambient[:] = [float(x) for x in sky_ambient]
        d.lump = None if lp is None else lp.ctypes.data
        d.lump_bytes = 0 if lp is None else lp.shape[0]
        d.nlights = 0 if li is None else len(li)
        d.lights = None if li is None or len(li) == 0 else li.ctypes.data
        self._keep = (fa, lp, li, d)
        HipRayTracer._raise(_lib.lib().vh_ambient_load(self.tracer.handle, ctypes.byref(d)))

    def cubes(self, points: np.ndarray) -> np.ndarray:
        """(n, 6, 3): +x, -x, +y, -y, +z, -z."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.zeros((p.shape[0], 6, 3), np.float32)
        HipRayTracer._raise(_lib.lib().vh_ambient_cubes(
            self.tracer.handle, p.ctypes.data_as(ctypes.c_void_p), p.shape[0], out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def timed_cubes(self, points: np.ndarray):
        """cubes() with the fan kernel timed by HIP events: (cubes, ms)."""
        L = _lib.lib()
        h = self.tracer.handle
        _lib.check(L.vh_timing_enable(h, 1))
        try:
            out = self.cubes(points)
            ms, n = ctypes.c_double(), ctypes.c_int64()
            _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_AMBIENT, ctypes.byref(ms), ctypes.byref(n)))
        finally:
            _lib.check(L.vh_timing_enable(h, 0))
        return out, ms.value
