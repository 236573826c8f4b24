"""Seeded synthetic workloads for the lighting hot path.

The reference compiles real maps (vbsp/vvis) before vrad runs; here the
geometry is generated, with the same shapes the reference's tracer sees:
axis-aligned box faces split into two triangles that share one id (vrad's
"one id per quad", Tracing/KdTree.cs:17-22), tagged opaque
(Rad/TraceId.cs:34), and sample points one unit off the face along its
normal, as FaceLightJob places them (Rad/Light/FaceLightJob.cs:244-251).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .tracer import TRACE_ID_OPAQUE, ShadowCasterSet

# Box face k: axis, side (0 = min face, 1 = max face); outward normal sign.
_FACES = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)]


def box_triangles(lo: np.ndarray, hi: np.ndarray, first_id: int = 0):
    """Twelve triangles per box (two per face).  Returns (v9 (12n, 9), ids (12n,)).
    A face's two triangles share the id ``TRACE_ID_OPAQUE | (first_id + 6*box + face)``."""
    lo = np.asarray(lo, dtype=np.float32).reshape(-1, 3)
    hi = np.asarray(hi, dtype=np.float32).reshape(-1, 3)
    n = lo.shape[0]
    v9 = np.empty((n, 6, 2, 9), dtype=np.float32)
    ids = np.empty((n, 6, 2), dtype=np.int32)
    for f, (ax, side) in enumerate(_FACES):
        u, v = (ax + 1) % 3, (ax + 2) % 3
        c = hi[:, ax] if side else lo[:, ax]
        quad = np.empty((n, 4, 3), dtype=np.float32)
        for q, (su, sv) in enumerate([(0, 0), (1, 0), (1, 1), (0, 1)]):
            quad[:, q, ax] = c
            quad[:, q, u] = hi[:, u] if su else lo[:, u]
            quad[:, q, v] = hi[:, v] if sv else lo[:, v]
        if not side:  # keep the winding outward
            quad = quad[:, [0, 3, 2, 1], :]
        v9[:, f, 0] = np.concatenate([quad[:, 0], quad[:, 1], quad[:, 2]], axis=1)
        v9[:, f, 1] = np.concatenate([quad[:, 0], quad[:, 2], quad[:, 3]], axis=1)
        fid = TRACE_ID_OPAQUE | (first_id + np.arange(n, dtype=np.int64) * 6 + f)
        ids[:, f, 0] = fid
        ids[:, f, 1] = fid
    return v9.reshape(-1, 9), ids.reshape(-1)


@dataclass
class VisibilityWorkload:
    """Config "visibility kernel": boxes, sample points on their faces and point lights."""

    lo: np.ndarray
    hi: np.ndarray
    casters: ShadowCasterSet
    points: np.ndarray  # (P, 3) float32
    lights: np.ndarray  # (L, 3) float32

    @property
    def rays(self) -> int:
        return self.points.shape[0] * self.lights.shape[0]


def random_boxes(nboxes: int, rng: np.random.Generator, volume: float = 8192.0,
                 edge_min: float = 16.0, edge_max: float = 512.0):
    """Axis-aligned boxes: edges log-uniform in [edge_min, edge_max] per axis,
    centres uniform in [0, volume]^3."""
    edges = np.exp(rng.uniform(np.log(edge_min), np.log(edge_max), size=(nboxes, 3)))
    centres = rng.uniform(0.0, volume, size=(nboxes, 3))
    lo = (centres - edges / 2).astype(np.float32)
    hi = (centres + edges / 2).astype(np.float32)
    return lo, hi


def points_on_boxes(lo: np.ndarray, hi: np.ndarray, npoints: int, rng: np.random.Generator,
                    offset: float = 1.0) -> np.ndarray:
    """Area-weighted uniform points on box faces, moved ``offset`` units along
    the outward normal (FaceLightJob's sample + normal, FaceLightJob.cs:244-251)."""
    ext = (hi - lo).astype(np.float64)
    areas = np.empty((lo.shape[0], 6))
    for f, (ax, _side) in enumerate(_FACES):
        u, v = (ax + 1) % 3, (ax + 2) % 3
        areas[:, f] = ext[:, u] * ext[:, v]
    flat = areas.reshape(-1)
    pick = rng.choice(flat.shape[0], size=npoints, p=flat / flat.sum())
    box, face = pick // 6, pick % 6
    s = rng.uniform(0.0, 1.0, size=(npoints, 2))
    pts = np.empty((npoints, 3), dtype=np.float64)
    for f, (ax, side) in enumerate(_FACES):
        m = face == f
        if not m.any():
            continue
        u, v = (ax + 1) % 3, (ax + 2) % 3
        b = box[m]
        pts[m, u] = lo[b, u] + s[m, 0] * ext[b, u]
        pts[m, v] = lo[b, v] + s[m, 1] * ext[b, v]
        pts[m, ax] = (hi[b, ax] + offset) if side else (lo[b, ax] - offset)
    return pts.astype(np.float32)


def visibility_workload(nboxes: int = 2000, npoints: int = 2_000_000, nlights: int = 64,
                        seed: int = 2, volume: float = 8192.0) -> VisibilityWorkload:
    """2,000 boxes -> 24,000 triangles; 2M points x 64 lights = 128M shadow rays at
    the defaults (seed 2)."""
    rng = np.random.default_rng(seed)
    lo, hi = random_boxes(nboxes, rng, volume)
    v9, ids = box_triangles(lo, hi)
    points = points_on_boxes(lo, hi, npoints, rng)
    lights = rng.uniform(0.0, volume, size=(nlights, 3)).astype(np.float32)
    return VisibilityWorkload(lo, hi, ShadowCasterSet(v9, ids), points, lights)


def point_light_rays(points: np.ndarray, lights: np.ndarray) -> np.ndarray:
    """The materialised rays of a point x light batch, point-major, in the
    correct-mode parameterisation (direction = light - point, reach 1)."""
    p = np.asarray(points, dtype=np.float32).reshape(-1, 1, 3)
    l = np.asarray(lights, dtype=np.float32).reshape(1, -1, 3)
    d = (l - p).astype(np.float32)
    out = np.empty((p.shape[0], l.shape[1], 7), dtype=np.float32)
    out[..., :3] = np.broadcast_to(p, d.shape)
    out[..., 3:6] = d
    out[..., 6] = 1.0
    return out.reshape(-1, 7)


def batching_workload(nfaces: int = 4096, min_luxels: int = 4, max_luxels: int = 16384,
                      nboxes: int = 2000, seed: int = 4):
    """The batching config: ``nfaces`` box faces of the visibility config's
    scene with luxel counts log-uniform over [min_luxels, max_luxels] (sample
    positions uniform on the face), lit by a sun and the sky.  Returns
    (casters, FaceSet)."""
    from .light import FaceSet

    rng = np.random.default_rng(seed)
    lo, hi = random_boxes(nboxes, np.random.default_rng(2))  # the visibility scene's boxes
    v9, ids = box_triangles(lo, hi)
    pick = rng.choice(nboxes * 6, size=nfaces, replace=False)
    counts = np.exp(rng.uniform(np.log(min_luxels), np.log(max_luxels), size=nfaces)).astype(np.int64)
    ptr = np.zeros(nfaces + 1, np.int64)
    np.cumsum(counts, out=ptr[1:])
    S = int(ptr[-1])
    pos = np.empty((S, 3), np.float32)
    normals = np.zeros((nfaces, 4, 3), np.float32)
    for k, q in enumerate(pick):
        box, face = q // 6, q % 6
        ax, side = _FACES[face]
        u, v = (ax + 1) % 3, (ax + 2) % 3
        n = int(counts[k])
        s = rng.uniform(0.0, 1.0, size=(n, 2))
        p = np.empty((n, 3))
        p[:, u] = lo[box, u] + s[:, 0] * (hi[box, u] - lo[box, u])
        p[:, v] = lo[box, v] + s[:, 1] * (hi[box, v] - lo[box, v])
        p[:, ax] = hi[box, ax] if side else lo[box, ax]
        pos[ptr[k]:ptr[k + 1]] = p
        normals[k, 0, ax] = 1.0 if side else -1.0
    faces = FaceSet(ptr.astype(np.int32), np.ones(nfaces, np.int32), normals, pos,
                    np.zeros(S, np.int32))
    return ShadowCasterSet(v9, ids), faces
