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
    L._gather_bound = True
    return L


def halton(base: int, index: int) -> float:
    """HaltonSequence's value after `index` increments from seed 1."""
    return _bind().vc_halton(base, index)


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
        _fields_ = list(base._fields_) + [("light_filter", ctypes.c_int32), ("no_skip", ctypes.c_int32)]

    return Desc


def gather(casters, desc_and_keep, nsamples, threads=None, light_filter: int = 0, no_skip: bool = False):
    """Runs the oracle's gather over a descriptor built by
    hl2sdk_tools_sharp_amd.light.pack_desc (the oracle's struct is that layout
    plus the resample's light filter and no-skip flag, set here).
    Returns (ldr (4,S,4), hdr (4,S,4), ao (S,))."""
    base, _keep = desc_and_keep
    d = _desc_type(type(base))()
    for name, _ in type(base)._fields_:
        setattr(d, name, getattr(base, name))
    d.light_filter = int(light_filter)
    d.no_skip = 1 if no_skip else 0
    tris = convert_triangles(casters.vertices, casters.ids)
    bvh = CpuBvh(casters.vertices, tris)
    ldr = np.zeros((4, nsamples, 4), np.float32)
    hdr = np.zeros((4, nsamples, 4), np.float32)
    ao = np.zeros(max(nsamples, 1), np.float32)
    _bind().vc_gather(ctypes.byref(d), bvh._h, _p(ldr), _p(hdr), _p(ao), threads or os.cpu_count() or 1)
    return ldr, hdr, ao[:nsamples]
