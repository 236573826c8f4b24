"""FinalLighting on the GPU: the luxel radial filter and the lighting lump.

The reference's final stage (FinalLighting.RunAsync over FinalLightFace.Run,
SourceSharp.MapTools/Rad/Final/FinalLightFace.cs:93-299, with LuxelRadial,
Rad/Final/LuxelRadial.cs) filters each face's light samples onto its luxel
grid, adds the bounced light of the patches around it, clamps to minlight and
writes the lump with each face's median colour before its luxels; the layout
is LightmapOffsets.Compute's (Rad/Light/LightmapOffsets.cs).  Here it runs in
libvradhip.so on the device-resident gather results and, with a bounce, the
device-resident bounced light (``vh_final_load`` / ``vh_final_light``).

Scope: one light style, faces that are not displacements, no macro textures,
the Correct sides of SampleRadialEdgeOffByOne and
PatchRadialNeighbourBumpFromSelf.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .light import HipFaceLighter
from .tracer import HipRayTracer

# vh_final_face
FACE_DTYPE = np.dtype([
    ("origin", "<f4", 3), ("w2l_s", "<f4", 3), ("w2l_t", "<f4", 3), ("l2w_s", "<f4", 3),
    ("l2w_t", "<f4", 3), ("min_s", "<i4"), ("min_t", "<i4"), ("width", "<i4"), ("height", "<i4"),
    ("minlight", "<f4"), ("lit", "<i4"),
])

VH_FINAL_NO_BOUNCE = 0
VH_FINAL_BOUNCE_DEVICE = 1
VH_FINAL_BOUNCE_HOST = 2


class VhFinalDesc(ctypes.Structure):
    _fields_ = [
        ("nfaces", ctypes.c_int32), ("faces", ctypes.c_void_p), ("luxel_pos", ctypes.c_void_p),
        ("sample_bounds", ctypes.c_void_p), ("nbr_ptr", ctypes.c_void_p), ("nbr", ctypes.c_void_p),
        ("npatch", ctypes.c_int32), ("patch_origin", ctypes.c_void_p), ("patch_face", ctypes.c_void_p),
        ("patch_child1", ctypes.c_void_p), ("winding_ptr", ctypes.c_void_p), ("winding", ctypes.c_void_p),
    ]


class VhFinalOpts(ctypes.Structure):
    _fields_ = [("which", ctypes.c_int32), ("bounce", ctypes.c_int32), ("red_errors", ctypes.c_int32)]


@dataclass
class FinalFaces:
    """Per face its luxel frame (FaceLightInfo) and final-light inputs
    (``faces``, FACE_DTYPE), the luxels' world points (width * height per
    face, face after face; None for BuildFaceLuxels' LuxelToWorld(s, t) of
    every grid point, made on the device), per gather sample its bounds in
    its face's luxel space (min s, min t, max s, max t) and the neighbour
    lists as CSR."""

    faces: np.ndarray
    luxel_pos: Optional[np.ndarray]
    sample_bounds: np.ndarray
    nbr_ptr: np.ndarray
    nbr: np.ndarray

    @property
    def nfaces(self) -> int:
        return len(self.faces)


@dataclass
class FinalPatches:
    """The patch radial's patches: origin, face, child1 (-1 for a leaf) and
    winding points as CSR."""

    origin: np.ndarray
    face: np.ndarray
    child1: np.ndarray
    winding_ptr: np.ndarray
    winding: np.ndarray

    @property
    def npatch(self) -> int:
        return len(self.origin)


class HipFinalLighter:
    """FinalLighting.RunAsync for one style against a loaded gather."""

    def __init__(self, lighter: HipFaceLighter):
        self.lighter = lighter
        self.tracer: HipRayTracer = lighter.tracer
        self._keep = None
        self.npatch = 0

    def load(self, faces: FinalFaces, patches: Optional[FinalPatches] = None) -> None:
        keep = {}

        def c(name, a, dt):
            if a is None:
                return None
            keep[name] = np.ascontiguousarray(a, dtype=dt)
            return keep[name].ctypes.data if keep[name].size else None

        fa = np.ascontiguousarray(faces.faces, FACE_DTYPE)
        keep["faces"] = fa
        d = VhFinalDesc(
            len(fa), fa.ctypes.data if len(fa) else None, c("lux", faces.luxel_pos, np.float32),
            c("bounds", faces.sample_bounds, np.float32), c("nbr_ptr", faces.nbr_ptr, np.int32),
            c("nbr", faces.nbr, np.int32), 0, None, None, None, None, None)
        if patches is not None and patches.npatch > 0:
            d.npatch = patches.npatch
            d.patch_origin = c("po", patches.origin, np.float32)
            d.patch_face = c("pf", patches.face, np.int32)
            d.patch_child1 = c("pc", patches.child1, np.int32)
            d.winding_ptr = c("wp", patches.winding_ptr, np.int32)
            d.winding = c("w", patches.winding, np.float32)
        self.npatch = int(d.npatch)
        self._keep = (d, keep)
        HipRayTracer._raise(_lib.lib().vh_final_load(self.tracer.handle, ctypes.byref(d)))

    def layout(self) -> Tuple[np.ndarray, int]:
        """(each face's light offset, -1 if not lit; the lump's size)."""
        n = len(self._keep[1]["faces"])
        offs = np.zeros(max(n, 1), np.int32)
        size = ctypes.c_int64()
        HipRayTracer._raise(_lib.lib().vh_final_layout(
            self.tracer.handle, offs.ctypes.data_as(ctypes.c_void_p), ctypes.byref(size)))
        return offs[:n], int(size.value)

    def run(self, hdr: bool = False, bounce: str = "none", patch_light=None,
            red_errors: bool = False) -> np.ndarray:
        """The lump bytes.  ``bounce``: 'none' (no patch radial), 'device'
        (the last bounce's light, flat) or 'host' (``patch_light``, TotalLight
        per normal: npatch x 4 x 3)."""
        mode = {"none": VH_FINAL_NO_BOUNCE, "device": VH_FINAL_BOUNCE_DEVICE,
                "host": VH_FINAL_BOUNCE_HOST}[bounce]
        _, size = self.layout()
        out = np.zeros(max(size, 1), np.uint8)
        pl = None
        if mode == VH_FINAL_BOUNCE_HOST:
            pl = np.ascontiguousarray(patch_light, np.float32).reshape(self.npatch, 4, 3)
        o = VhFinalOpts(1 if hdr else 0, mode, 1 if red_errors else 0)
        HipRayTracer._raise(_lib.lib().vh_final_light(
            self.tracer.handle, ctypes.byref(o), None if pl is None else pl.ctypes.data_as(ctypes.c_void_p),
            out.ctypes.data_as(ctypes.c_void_p), out.shape[0]))
        return out[:size]

    def timed_run(self, **kw) -> Tuple[np.ndarray, float]:
        """run() with the stage's launches timed by HIP events on the
        library's stream: (lump, milliseconds)."""
        L = _lib.lib()
        h = self.tracer.handle
        _lib.check(L.vh_timing_enable(h, 1))
        try:
            lump = self.run(**kw)
            ms, n = ctypes.c_double(), ctypes.c_int64()
            _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_FINAL, ctypes.byref(ms), ctypes.byref(n)))
        finally:
            _lib.check(L.vh_timing_enable(h, 0))
        return lump, ms.value
