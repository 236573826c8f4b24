"""The direct-light gather: lights, sampling tables and the face-lighting stage.

Reference (paths under the reference's ``src/SourceSharp.MapTools/``):

* ``DirectLight`` -- Rad/Light/DirectLight.cs (type, origin, normal, intensity,
  attenuation, fade, spot cone, style, PVS);
* ``DirectionalSampler`` / ``HaltonSequence`` -- Rad/Light/DirectionalSampler.cs
  (the sky directions and the sun's jitter);
* ``RadWorld.LightFacesAsync`` -> ``FaceLightJob`` -> ``DirectLightGatherer`` --
  Rad/Light/RadWorld.Facelights.cs:194, FaceLightJob.cs:329-463,
  DirectLightGatherer.cs:441-1306: the per-sample gather into LightingValue;
* ``StockLightColor.Encode`` -- Rad/Ambient/StockLightColor.cs:136-185.

``HipFaceLighter`` runs the gather and the encode in libvradhip.so on the GPU.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .tracer import HipRayTracer

EMIT_SURFACE = 0
EMIT_POINT = 1
EMIT_SPOTLIGHT = 2
EMIT_SKYLIGHT = 3
EMIT_SKYAMBIENT = 5

VERTEX_NORMAL_COUNT = 162  # LightConstants.VertexNormalCount: the sky directions
SUN_AREA_LIGHT_SAMPLES = 30  # LightConstants.SunAreaLightSamples

LIGHT_DTYPE = np.dtype(
    [
        ("type", "<i4"), ("style", "<i4"),
        ("origin", "<f4", 3), ("normal", "<f4", 3), ("intensity", "<f4", 3), ("intensity_hdr", "<f4", 3),
        ("constant_attn", "<f4"), ("linear_attn", "<f4"), ("quadratic_attn", "<f4"), ("cap_dist", "<f4"),
        ("start_fade", "<f4"), ("end_fade", "<f4"), ("stopdot", "<f4"), ("stopdot2", "<f4"),
        ("exponent", "<f4"),
    ]
)
assert LIGHT_DTYPE.itemsize == 92  # sizeof(vh_light)


def make_lights(n: int) -> np.ndarray:
    """n lights with DirectLight's defaults (CapDist 1e22, EndFadeDistance -1)."""
    L = np.zeros(n, LIGHT_DTYPE)
    L["cap_dist"] = 1.0e22
    L["end_fade"] = -1.0
    return L


def _halton(base: int, count: int) -> np.ndarray:
    """HaltonSequence.NextValue, from seed 2 (the sequence's first value is its
    second element), in float32 arithmetic as the reference writes it."""
    out = np.empty(count, np.float32)
    fbase = np.float32(base)
    for i in range(count):
        seed = i + 2
        ret = np.float32(0.0)
        base_inv = np.float32(1.0 / float(fbase))
        while seed:
            dig = seed % base
            ret = np.float32(ret + np.float32(np.float32(dig) * base_inv))
            base_inv = np.float32(base_inv / fbase)
            seed //= base
        out[i] = ret
    return out


def directional_sampler(count: int) -> np.ndarray:
    """The first ``count`` directions of DirectionalSampler: z from Halton base 2,
    the azimuth from base 3; DetMathF's correctly rounded acos/sin/cos as a
    double evaluation rounded once to float."""
    z2 = _halton(2, count)
    h3 = _halton(3, count)
    out = np.empty((count, 3), np.float32)
    for i in range(count):
        z = np.float32(float(np.float32(np.float32(2.0) * z2[i])) - 1.0)  # (float)((2*z) - 1.0)
        phi = np.float32(math.acos(float(z)))
        theta = np.float32(2.0 * math.pi * float(h3[i]))
        sinp = np.float32(math.sin(float(phi)))
        out[i, 0] = np.float32(np.float32(math.cos(float(theta))) * sinp)
        out[i, 1] = np.float32(np.float32(math.sin(float(theta))) * sinp)
        out[i, 2] = z
    return out


def sky_directions(scale: float = 1.0) -> np.ndarray:
    """The sky-ambient directions: int(162 * SkySampleScale) of them."""
    return directional_sampler(int(VERTEX_NORMAL_COUNT * scale))


def sun_jitter(samples: int = SUN_AREA_LIGHT_SAMPLES) -> np.ndarray:
    """The sun's jitter table: ray d >= 1 uses entry d - 1."""
    return directional_sampler(max(samples - 1, 0))


def ao_directions(face_normals: np.ndarray, count: int) -> np.ndarray:
    """Cosine-weighted hemisphere directions about each face normal (the
    build's own ambient-occlusion term): a fixed Hammersley set (i + 0.5)/n
    and the base-2 radical inverse, r = sqrt(u), phi = 2 pi v, in a tangent
    frame of the normal.  Returns (F, count, 3) float32."""
    n = np.asarray(face_normals, np.float64).reshape(-1, 3)
    i = np.arange(count)
    u = (i + 0.5) / count
    v = np.array([int(format(k, "032b")[::-1], 2) / 2.0**32 for k in i])
    r = np.sqrt(u)
    phi = 2 * np.pi * v
    local = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1 - u))], axis=1)
    helper = np.where(np.abs(n[:, 0:1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    t = np.cross(helper, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    out = (local[None, :, 0:1] * t[:, None, :] + local[None, :, 1:2] * b[:, None, :]
           + local[None, :, 2:3] * n[:, None, :])
    return out.astype(np.float32)


def bump_basis(face_normal: np.ndarray) -> np.ndarray:
    """Three bump normals around a flat face normal, from the reference's local
    basis (Rad/Light/BumpBasis.cs:43-58) in a tangent frame of the normal."""
    n = np.asarray(face_normal, np.float64)
    local = np.array([[0.81649661064147949, 0.0, 0.57735025882720947],
                      [-0.40824821591377258, 0.70710676908493042, 0.57735025882720947],
                      [-0.40824821591377258, -0.70710676908493042, 0.57735025882720947]])
    helper = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(helper, n)
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    return (local[:, 0:1] * t + local[:, 1:2] * b + local[:, 2:3] * n).astype(np.float32)


@dataclass
class FaceSet:
    """Faces to light: per face its sample range, normal count and normals
    (the face normal, then three bump normals), per sample its position on the
    face and its cluster."""

    face_sample_ptr: np.ndarray
    face_normal_count: np.ndarray
    face_normals: np.ndarray  # (F, 4, 3)
    sample_pos: np.ndarray  # (S, 3)
    sample_cluster: np.ndarray  # (S,)
    sample_area: Optional[np.ndarray] = None  # (S,) luxel areas, for the patch-light hand-off

    @property
    def nfaces(self) -> int:
        return len(self.face_normal_count)

    @property
    def nsamples(self) -> int:
        return self.sample_pos.shape[0]

    def subset(self, faces: np.ndarray) -> "FaceSet":
        """The listed faces (in the given order) with their samples."""
        faces = np.asarray(faces, np.int64)
        lo = self.face_sample_ptr[faces].astype(np.int64)
        hi = self.face_sample_ptr[faces + 1].astype(np.int64)
        idx = np.concatenate([np.arange(x, y) for x, y in zip(lo, hi)]) if len(faces) else np.zeros(0, np.int64)
        ptr = np.zeros(len(faces) + 1, np.int32)
        np.cumsum(hi - lo, out=ptr[1:])
        return FaceSet(ptr, self.face_normal_count[faces].copy(), self.face_normals[faces].copy(),
                       self.sample_pos[idx].copy(), self.sample_cluster[idx].copy(),
                       None if self.sample_area is None else self.sample_area[idx].copy())


class VhGatherDesc(ctypes.Structure):
    _fields_ = [
        ("nfaces", ctypes.c_int32), ("nsamples", ctypes.c_int32), ("nlights", ctypes.c_int32),
        ("pvs_bytes", ctypes.c_int32),
        ("face_sample_ptr", ctypes.c_void_p), ("face_normal_count", ctypes.c_void_p),
        ("face_normals", ctypes.c_void_p), ("sample_pos", ctypes.c_void_p),
        ("sample_cluster", ctypes.c_void_p), ("lights", ctypes.c_void_p), ("light_pvs", ctypes.c_void_p),
        ("sun_extent", ctypes.c_float), ("sun_samples", ctypes.c_int32), ("sun_jitter", ctypes.c_void_p),
        ("sky_dirs_count", ctypes.c_int32), ("sky_dirs", ctypes.c_void_p),
        ("ao_count", ctypes.c_int32), ("ao_radius", ctypes.c_float), ("ao_dirs", ctypes.c_void_p),
    ]


def pack_desc(faces: FaceSet, lights: np.ndarray, light_pvs: Optional[np.ndarray], sun_extent: float,
              sun_samples: int, sky_dirs: Optional[np.ndarray], ao_count: int = 0,
              ao_radius: float = 40.0, skybox=None):
    """Builds the descriptor (shared by the product and the oracle, whose
    struct has the same layout) and the arrays it points into; ``skybox``
    (skybox.Skybox or None) rides along as keep["skybox_desc"], the
    vh_skybox_desc the device loads with vh_gather_skybox."""
    keep = {}
    if skybox is not None and skybox.ncameras > 0:
        keep["skybox"] = skybox
        keep["skybox_desc"] = skybox.pack()

    def c(name, a, dt):
        keep[name] = None if a is None else np.ascontiguousarray(a, dtype=dt)
        return None if keep[name] is None else keep[name].ctypes.data

    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    keep["lights"] = lights
    pvs_bytes = 0 if light_pvs is None else int(np.asarray(light_pvs).shape[1])
    jit = sun_jitter(sun_samples) if sun_samples > 1 else None
    ao = ao_directions(faces.face_normals[:, 0], ao_count) if ao_count > 0 else None
    d = VhGatherDesc(
        faces.nfaces, faces.nsamples, len(lights), pvs_bytes,
        c("fsp", faces.face_sample_ptr, np.int32), c("fnc", faces.face_normal_count, np.int32),
        c("fn", faces.face_normals, np.float32), c("pos", faces.sample_pos, np.float32),
        c("cl", faces.sample_cluster, np.int32), lights.ctypes.data,
        c("pvs", light_pvs, np.uint8), float(sun_extent), int(max(sun_samples, 1)),
        c("jit", jit, np.float32), 0 if sky_dirs is None else len(sky_dirs), c("sky", sky_dirs, np.float32),
        int(ao_count), float(ao_radius), c("ao", ao, np.float32),
    )
    return d, keep


@dataclass
class GatherResult:
    ldr: np.ndarray  # (4, S, 4): per normal slot, per sample, RGB + direct sun amount
    hdr: np.ndarray
    ao: np.ndarray  # (S,)


class VhSupersampleDesc(ctypes.Structure):
    _fields_ = [("faces", ctypes.c_void_p), ("sample_st", ctypes.c_void_p), ("which", ctypes.c_int32),
                ("extra_passes", ctypes.c_int32), ("unset_is_garbage", ctypes.c_int32)]


class HipFaceLighter:
    """The per-sample direct-light gather on the GPU, against the tracer's scene."""

    def __init__(self, tracer: HipRayTracer):
        self.tracer = tracer
        self._keep = None
        self.faces: Optional[FaceSet] = None

    def load(self, faces: FaceSet, lights: np.ndarray, light_pvs=None, sun_extent: float = 0.0,
             sun_samples: int = SUN_AREA_LIGHT_SAMPLES, sky_dirs=None, ao_count: int = 0,
             ao_radius: float = 40.0, skybox=None) -> None:
        """Uploads the faces, lights and tables; ``skybox`` (skybox.Skybox)
        gives the sky rays their 3D-skybox stage, None removes it."""
        d, keep = pack_desc(faces, lights, light_pvs, sun_extent, sun_samples, sky_dirs, ao_count, ao_radius,
                            skybox)
        self._keep = (d, keep)
        self.faces = faces
        L = _lib.lib()
        HipRayTracer._raise(L.vh_gather_load(self.tracer.handle, ctypes.byref(d)))
        sb = keep.get("skybox_desc")
        HipRayTracer._raise(L.vh_gather_skybox(self.tracer.handle, None if sb is None else ctypes.byref(sb)))

    def run(self, schedule: str = "queue", waves_per_cu: int = 0, chunks_per_ticket: int = 1) -> None:
        code = {"queue": 0, "per_face": 1, "grid": 2}[schedule]
        HipRayTracer._raise(_lib.lib().vh_gather_run_shaped(self.tracer.handle, code, int(waves_per_cu),
                                                            int(chunks_per_ticket)))

    def read(self) -> GatherResult:
        S = self.faces.nsamples
        ldr = np.zeros((4, S, 4), np.float32)
        hdr = np.zeros((4, S, 4), np.float32)
        ao = np.zeros(S, np.float32)
        HipRayTracer._raise(_lib.lib().vh_gather_read(
            self.tracer.handle, ldr.ctypes.data_as(ctypes.c_void_p), hdr.ctypes.data_as(ctypes.c_void_p),
            ao.ctypes.data_as(ctypes.c_void_p)))
        return GatherResult(ldr, hdr, ao)

    def read_styles(self):
        """The last run's light styles: (face_styles (F, 4) uint8 style numbers
        per slot, 255 empty; ldr and hdr (4 slots, 4 normals, S, 4); the
        (face, style) pairs dropped past the fourth slot)."""
        L = _lib.lib()
        h = self.tracer.handle
        F, S = self.faces.nfaces, self.faces.nsamples
        fs = np.zeros((max(F, 1), 4), np.uint8)
        slots, dropped = ctypes.c_int32(), ctypes.c_int64()
        HipRayTracer._raise(L.vh_gather_styles(h, fs.ctypes.data_as(ctypes.c_void_p), ctypes.byref(slots),
                                               ctypes.byref(dropped)))
        ldr = np.zeros((4, 4, S, 4), np.float32)
        hdr = np.zeros((4, 4, S, 4), np.float32)
        for k in range(4):
            HipRayTracer._raise(L.vh_gather_read_slot(h, k, ldr[k].ctypes.data_as(ctypes.c_void_p),
                                                      hdr[k].ctypes.data_as(ctypes.c_void_p)))
        return fs[:F], ldr, hdr, int(dropped.value)

    def supersample(self, frames: np.ndarray, which: int = 0, extra_passes: int = 4, sample_st=None,
                    unset_is_garbage: bool = False) -> list:
        """FaceLightJob's supersampling passes (Rad/Light/FaceLightJob.cs:698-1207)
        over the last run's LDR (``which`` 0) or HDR (1) light, refined in
        place on the device.  ``frames``: each face's luxel frame and grid
        (final.FACE_DTYPE); ``sample_st`` (S, 2) each sample's luxel, or None
        for sample k of a face on (k % width, k // width).  Returns the number
        of samples each pass re-lit."""
        from .final import FACE_DTYPE

        fr = np.ascontiguousarray(frames, FACE_DTYPE)
        st = None if sample_st is None else np.ascontiguousarray(sample_st, np.int32).reshape(-1, 2)
        d = VhSupersampleDesc(fr.ctypes.data, None if st is None else st.ctypes.data, int(which),
                              int(extra_passes), 1 if unset_is_garbage else 0)
        counts = np.zeros(max(int(extra_passes), 1), np.int32)
        HipRayTracer._raise(_lib.lib().vh_gather_supersample(self.tracer.handle, ctypes.byref(d),
                                                             counts.ctypes.data_as(ctypes.c_void_p)))
        return [int(x) for x in counts[: int(extra_passes)] if x > 0]

    def timed_supersample(self, frames, **kw) -> tuple:
        """supersample() with its launches timed by HIP events on the
        library's stream: (counts, kernel ms, rays traced)."""
        L = _lib.lib()
        h = self.tracer.handle
        before = self.tracer.stats()["rays_gather"]
        _lib.check(L.vh_timing_enable(h, 1))
        try:
            counts = self.supersample(frames, **kw)
            ms, n = ctypes.c_double(), ctypes.c_int64()
            _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_SUPERSAMPLE, ctypes.byref(ms), ctypes.byref(n)))
        finally:
            _lib.check(L.vh_timing_enable(h, 0))
        return counts, ms.value, self.tracer.stats()["rays_gather"] - before

    def timed_run(self, schedule: str = "queue", waves_per_cu: int = 0, chunks_per_ticket: int = 1) -> tuple:
        """One run with the gather kernel timed by HIP events on the library's
        stream: (kernel ms summed over its launches, launches, rays traced)."""
        L = _lib.lib()
        h = self.tracer.handle
        before = self.tracer.stats()["rays_gather"]
        _lib.check(L.vh_timing_enable(h, 1))
        try:
            self.run(schedule, waves_per_cu, chunks_per_ticket)
            ms, n = ctypes.c_double(), ctypes.c_int64()
            _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_FACELIGHT, ctypes.byref(ms), ctypes.byref(n)))
        finally:
            _lib.check(L.vh_timing_enable(h, 0))
        return ms.value, n.value, self.tracer.stats()["rays_gather"] - before

    def light_faces(self, **kw) -> GatherResult:
        self.run(**kw)
        return self.read()

    def lightmap(self, hdr: bool = False) -> np.ndarray:
        """The lighting lump bytes (ColorRGBExp32 per face, normal, sample)."""
        f = self.faces
        n = int(np.sum(f.face_normal_count.astype(np.int64) * np.diff(f.face_sample_ptr))) * 4
        out = np.zeros(max(n, 1), np.uint8)
        HipRayTracer._raise(_lib.lib().vh_lightmap_encode(
            self.tracer.handle, 1 if hdr else 0, out.ctypes.data_as(ctypes.c_void_p), out.shape[0]))
        return out[:n]


def encode_rgbe(tracer: HipRayTracer, rgb: np.ndarray) -> np.ndarray:
    """StockLightColor.Encode on the GPU: (n, >=3) float32 -> (n, 4) bytes."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    n, stride = rgb.shape
    out = np.zeros((max(n, 1), 4), np.uint8)
    HipRayTracer._raise(_lib.lib().vh_encode_rgbe(
        tracer.handle, rgb.ctypes.data_as(ctypes.c_void_p), n, stride, out.ctypes.data_as(ctypes.c_void_p)))
    return out[:n]
