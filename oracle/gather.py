"""Python side of the oracle's gather, sampler and encode -- TEST INFRASTRUCTURE ONLY.

Wraps vradcpu_gather.c (DirectLightGatherer, FaceLightJob, DirectionalSampler
and StockLightColor restated in C; citations there).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import CpuBvh, _p, convert_triangles, lib


def _bind():
    L = lib()
    if getattr(L, "_gather_bound", False):
        return L
    P = ctypes.c_void_p
    L.vc_directional_sampler.argtypes = [ctypes.c_int32, P]
    L.vc_halton.argtypes = [ctypes.c_int32, ctypes.c_int32]
    L.vc_halton.restype = ctypes.c_float
    L.vc_encode_rgbe.argtypes = [P, ctypes.c_int64, ctypes.c_int32, P]
    L.vc_gather.argtypes = [P, P, P, P, P, ctypes.c_int32]
    L.vc_gather.restype = ctypes.c_int64
    L.vc_pow_f32_array.argtypes = [P, P, ctypes.c_int64, P]
    L.vc_pow_to_single_array.argtypes = [P, P, ctypes.c_int64, P]
    L.vc_pow_needs_exact.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double]
    L.vc_pow_needs_exact.restype = ctypes.c_int32
    L._gather_bound = True
    return L


def halton(base: int, index: int) -> float:
    """HaltonSequence's value after `index` increments from seed 1."""
    return _bind().vc_halton(base, index)


def pow_f32(x, y) -> np.ndarray:
    """DetMathF.Pow: x^y for float x and y, rounded once to float."""
    x = np.ascontiguousarray(x, np.float32).reshape(-1)
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float32), x.shape), np.float32)
    out = np.zeros(max(x.size, 1), np.float32)
    _bind().vc_pow_f32_array(_p(x), _p(y), x.size, _p(out))
    return out[: x.size]


def pow_to_single(x, y) -> np.ndarray:
    """DetMath.PowToSingle: x^y for double x and y, rounded to double, then float."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float64), x.shape), np.float64)
    out = np.zeros(max(x.size, 1), np.float32)
    _bind().vc_pow_to_single_array(_p(x), _p(y), x.size, _p(out))
    return out[: x.size]


def pow_needs_exact(kind: int, x: float, y: float) -> bool:
    return bool(_bind().vc_pow_needs_exact(kind, x, y))


def directional_sampler(n: int) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    _bind().vc_directional_sampler(n, _p(out))
    return out


def encode_rgbe(rgb: np.ndarray) -> np.ndarray:
    rgb = np.ascontiguousarray(rgb, np.float32)
    n, stride = rgb.shape
    out = np.zeros((max(n, 1), 4), np.uint8)
    _bind().vc_encode_rgbe(_p(rgb), n, stride, _p(out))
    return out[:n]


def _desc_type(base):
    class Desc(ctypes.Structure):
        _fields_ = list(base._fields_) + [("light_filter", ctypes.c_int32), ("no_skip", ctypes.c_int32),
                                          ("face_style_filter", ctypes.c_void_p), ("skybox", ctypes.c_void_p)]

    return Desc


def bsp_leaf_from_point(skybox, p) -> int:
    """The oracle's CompiledBspTree.LeafFromPoint over a skybox.Skybox's BSP."""
    L = _bind()
    L.vc_bsp_leaf_from_point.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.vc_bsp_leaf_from_point.restype = ctypes.c_int32
    desc = skybox.pack()
    q = np.ascontiguousarray(p, np.float32).reshape(3)
    return int(L.vc_bsp_leaf_from_point(ctypes.byref(desc), _p(q)))


def sky_resolve(count: int, first_occ, inner, mask: int) -> np.ndarray:
    """ResolveSkyTest's arithmetic over given answers: first_occ 4 first-ray
    occlusions, inner (cameras, 4) skybox-ray occlusions, mask the lanes that
    recurse; the 4 visible fractions."""
    L = _bind()
    L.vc_sky_resolve.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                 ctypes.c_int32, ctypes.c_void_p]
    f = np.ascontiguousarray(first_occ, np.float32).reshape(4)
    i = np.ascontiguousarray(inner, np.float32).reshape(-1, 4)
    out = np.zeros(4, np.float32)
    L.vc_sky_resolve(int(count), _p(f), len(i), _p(i), int(mask), _p(out))
    return out


def set_mutation(which: int) -> None:
    """Plants one mistake in the oracle's gather (0: none); for the tests that
    show each closed form catches its mistake."""
    L = _bind()
    L.vc_gather_set_mutation.argtypes = [ctypes.c_int32]
    L.vc_gather_set_mutation(int(which))


class GatherOut:
    """The oracle's gather: ``ldr``/``hdr`` (4 slots, 4 normals, S, 4), ``ao``
    (S,), ``face_styles`` (F, 4) style numbers per slot (255 empty),
    ``rays`` traced and ``dropped`` (group, light) records lost to a fifth
    style."""

    def __init__(self, ldr, hdr, ao, face_styles, rays, dropped):
        self.ldr, self.hdr, self.ao = ldr, hdr, ao
        self.face_styles, self.rays, self.dropped = face_styles, rays, dropped


def _prepare(casters, desc_and_keep, light_filter, no_skip, face_style_filter):
    base, _keep = desc_and_keep
    d = _desc_type(type(base))()
    for name, _ in type(base)._fields_:
        setattr(d, name, getattr(base, name))
    d.light_filter = int(light_filter)
    d.no_skip = 1 if no_skip else 0
    fsf = None
    if face_style_filter is not None:
        fsf = np.ascontiguousarray(face_style_filter, np.int32)
        d.face_style_filter = fsf.ctypes.data
    sb = _keep.get("skybox_desc")
    if sb is not None:  # the 3D skybox: vc_skybox has vh_skybox_desc's layout
        d.skybox = ctypes.addressof(sb)
    tris = convert_triangles(casters.vertices, casters.ids)
    bvh = CpuBvh(casters.vertices, tris)
    return d, bvh, fsf


def gather(casters, desc_and_keep, nsamples, threads=None, light_filter: int = 0, no_skip: bool = False,
           face_style_filter=None):
    """Runs the oracle's gather over a descriptor built by
    hl2sdk_tools_sharp_amd.light.pack_desc (the oracle's struct is that layout
    plus the resample's light filter, no-skip flag and per-face style filter,
    set here), every passed light into one slot.
    Returns (ldr (4,S,4), hdr (4,S,4), ao (S,))."""
    d, bvh, _fsf = _prepare(casters, desc_and_keep, light_filter, no_skip, face_style_filter)
    ldr = np.zeros((4, nsamples, 4), np.float32)
    hdr = np.zeros((4, nsamples, 4), np.float32)
    ao = np.zeros(max(nsamples, 1), np.float32)
    _bind().vc_gather(ctypes.byref(d), bvh._h, _p(ldr), _p(hdr), _p(ao), threads or os.cpu_count() or 1)
    return ldr, hdr, ao[:nsamples]


def gather_styles(casters, desc_and_keep, nsamples, threads=None) -> GatherOut:
    """The gather with light-style slots (FaceLight.FindOrAllocateStyle per
    face as lights first touch it), plus the rays it traced."""
    L = _bind()
    L.vc_gather_styles.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_void_p, ctypes.c_int32]
    L.vc_gather_styles.restype = ctypes.c_int64
    d, bvh, _fsf = _prepare(casters, desc_and_keep, 0, False, None)
    ldr = np.zeros((4, 4, nsamples, 4), np.float32)
    hdr = np.zeros((4, 4, nsamples, 4), np.float32)
    ao = np.zeros(max(nsamples, 1), np.float32)
    fs = np.zeros((max(d.nfaces, 1), 4), np.uint8)
    dropped = ctypes.c_int64(0)
    rays = L.vc_gather_styles(ctypes.byref(d), bvh._h, _p(ldr), _p(hdr), _p(ao), _p(fs), ctypes.byref(dropped),
                              threads or os.cpu_count() or 1)
    return GatherOut(ldr, hdr, ao[:nsamples], fs[: d.nfaces], int(rays), int(dropped.value))
