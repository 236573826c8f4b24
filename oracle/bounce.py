"""Python side of the oracle's transfer build and bounce -- TEST INFRASTRUCTURE ONLY.

Wraps vradcpu_bounce.c (VisMatrix, FormFactors, MakeScales and Radiosity
restated in C; citations there).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import CpuBvh, _p, convert_triangles, lib

P_ = ctypes.c_void_p


class PatchSetC(ctypes.Structure):
    _fields_ = [
        ("npatch", ctypes.c_int32), ("nface", ctypes.c_int32), ("ncluster", ctypes.c_int32),
        ("nreceivers", ctypes.c_int32),
        ("origin", P_), ("normal", P_), ("plane_dist", P_), ("area", P_),
        ("child1", P_), ("child2", P_), ("next_parent", P_), ("face", P_), ("cluster", P_),
        ("winding", P_), ("face_sky", P_), ("face_parent", P_), ("row_faces_ptr", P_),
        ("row_faces", P_), ("receivers", P_),
    ]


def _bind():
    L = lib()
    if getattr(L, "_bounce_bound", False):
        return L
    f32, i32, i64, f64 = ctypes.c_float, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.vc_diff_to_diff.argtypes = [P_, P_, P_, P_]
    L.vc_diff_to_diff.restype = f32
    L.vc_poly_to_diff.argtypes = [P_, i32, f32, P_, P_]
    L.vc_poly_to_diff.restype = f32
    L.vc_make_scales.argtypes = [P_, i32]
    L.vc_candidates.argtypes = [ctypes.POINTER(PatchSetC), i32, P_, i64]
    L.vc_candidates.restype = i64
    L.vc_make_transfer.argtypes = [ctypes.POINTER(PatchSetC), i32, i32, ctypes.POINTER(f32)]
    L.vc_make_transfer.restype = ctypes.c_int
    L.vc_transfers_build.argtypes = [ctypes.POINTER(PatchSetC), P_, P_, P_, P_, P_, i32]
    L.vc_bounce.argtypes = [i32, P_, P_, P_, P_, P_, P_, P_, P_, P_, P_, i32, i32, f64, i32, P_, P_, P_, P_,
                            P_, P_, P_, P_]
    L._bounce_bound = True
    return L


def _v3(x):
    return np.ascontiguousarray(x, dtype=np.float32).reshape(3)


def diff_to_diff(o1, n1, o2, n2) -> float:
    a, b, c, d = _v3(o1), _v3(n1), _v3(o2), _v3(n2)
    return _bind().vc_diff_to_diff(_p(a), _p(b), _p(c), _p(d))


def poly_to_diff(poly, area, o, n) -> float:
    poly = np.ascontiguousarray(poly, dtype=np.float32).reshape(-1, 3)
    oo, nn = _v3(o), _v3(n)
    return _bind().vc_poly_to_diff(_p(poly), poly.shape[0], float(area), _p(oo), _p(nn))


def make_scales(w) -> np.ndarray:
    w = np.ascontiguousarray(w, dtype=np.float32).copy()
    _bind().vc_make_scales(_p(w), w.shape[0])
    return w


class PatchSet:
    """The oracle's view of a rooms.RoomGrid (keeps the arrays alive)."""

    def __init__(self, g):
        self.g = g
        self._arrays = dict(
            origin=np.ascontiguousarray(g.origin, np.float32),
            normal=np.ascontiguousarray(g.normal, np.float32),
            plane_dist=np.ascontiguousarray(g.plane_dist, np.float32),
            area=np.ascontiguousarray(g.area, np.float32),
            child1=np.ascontiguousarray(g.child1, np.int32),
            child2=np.ascontiguousarray(g.child2, np.int32),
            next_parent=np.ascontiguousarray(g.next_parent, np.int32),
            face=np.ascontiguousarray(g.face, np.int32),
            cluster=np.ascontiguousarray(g.cluster, np.int32),
            winding=np.ascontiguousarray(g.winding, np.float32),
            face_sky=np.ascontiguousarray(g.face_sky, np.uint8),
            face_parent=np.ascontiguousarray(g.face_parent, np.int32),
            row_faces_ptr=np.ascontiguousarray(g.row_faces_ptr, np.int32),
            row_faces=np.ascontiguousarray(g.row_faces, np.int32),
            receivers=np.ascontiguousarray(g.receivers, np.int32),
        )
        a = self._arrays
        self.c = PatchSetC(
            g.npatch, g.nface, g.ncluster, len(g.receivers),
            *[a[k].ctypes.data for k in (
                "origin", "normal", "plane_dist", "area", "child1", "child2", "next_parent",
                "face", "cluster", "winding", "face_sky", "face_parent", "row_faces_ptr",
                "row_faces", "receivers")]
        )

    def candidates(self, receiver: int) -> np.ndarray:
        L = _bind()
        n = L.vc_candidates(ctypes.byref(self.c), int(receiver), None, 0)
        out = np.empty(n, np.int32)
        L.vc_candidates(ctypes.byref(self.c), int(receiver), _p(out), n)
        return out

    def make_transfer(self, r: int, s: int):
        t = ctypes.c_float()
        ok = _bind().vc_make_transfer(ctypes.byref(self.c), int(r), int(s), ctypes.byref(t))
        return bool(ok), t.value


def build_transfers(g, threads=None):
    """VisMatrix.BuildAsync restated: returns CSR (row_ptr int64 (P+1), src int32,
    w float32) with one row per patch (empty for non-receivers), rows in build
    order, each scaled by MakeScales."""
    L = _bind()
    ps = PatchSet(g)
    counts = np.zeros(g.npatch, np.int64)
    for r in g.receivers:
        counts[r] = L.vc_candidates(ctypes.byref(ps.c), int(r), None, 0)
    cand_ptr = np.zeros(g.npatch + 1, np.int64)
    np.cumsum(counts, out=cand_ptr[1:])
    total = int(cand_ptr[-1])
    src = np.zeros(max(total, 1), np.int32)
    w = np.zeros(max(total, 1), np.float32)
    made = np.zeros(g.npatch, np.int64)
    tris = convert_triangles(g.casters.vertices, g.casters.ids)
    bvh = CpuBvh(g.casters.vertices, tris)
    L.vc_transfers_build(ctypes.byref(ps.c), bvh._h, _p(cand_ptr), _p(made), _p(src), _p(w),
                         threads or os.cpu_count() or 1)
    if (made < 0).any():
        raise RuntimeError("candidate count and fill passes disagree")
    row_ptr = np.zeros(g.npatch + 1, np.int64)
    np.cumsum(made, out=row_ptr[1:])
    keep = np.zeros(total, bool)
    for r in np.nonzero(made)[0]:
        keep[cand_ptr[r]:cand_ptr[r] + made[r]] = True
    return row_ptr, src[:total][keep].copy(), w[:total][keep].copy(), int(counts.sum())


def bounce(g, row_ptr, src, w, emit0, start=None, mode=0, max_iter=100, tol=0.0, use_double=False,
           bumped=None, bump_normals=None, bump_out=None):
    """Radiosity.BounceAsync restated (see vradcpu_bounce.c for the modes).
    ``bumped`` (a flag per patch) with ``bump_normals`` ((P, 3, 3)) gathers
    those receivers per normal (GatherBumpedScalar); ``bump_out`` ((P, 3, 3)
    float64) then receives TotalLight's bump directions.
    Returns (total (P,3) float64, residuals, energies (iters,3), iters)."""
    L = _bind()
    P = g.npatch
    refl = np.ascontiguousarray(g.refl, np.float32)
    e0 = np.ascontiguousarray(emit0, np.float32)
    sky = np.ascontiguousarray(g.sky, np.uint8)
    st = None if start is None else np.ascontiguousarray(start, np.float32)
    out = np.zeros((P, 3), np.float64)
    res = np.zeros(max_iter, np.float64)
    en = np.zeros((max_iter, 3), np.float64)
    it = ctypes.c_int32()
    rp = np.ascontiguousarray(row_ptr, np.int64)
    s = np.ascontiguousarray(src, np.int32)
    ww = np.ascontiguousarray(w, np.float32)
    bump_args = (None, None, None, None)
    if bumped is not None:
        org = np.ascontiguousarray(g.origin, np.float32)
        n4 = np.zeros((P, 4, 3), np.float32)
        n4[:, 0] = g.normal
        n4[:, 1:] = np.asarray(bump_normals, np.float32).reshape(P, 3, 3)
        bf = np.ascontiguousarray(bumped, np.uint8)
        tb = np.zeros((P, 3, 3), np.float64)
        keep = (org, n4, bf, tb)
        bump_args = (_p(org), _p(n4), _p(bf), _p(tb))
    L.vc_bounce(P, _p(rp), _p(s), _p(ww), _p(refl), _p(e0), _p(sky),
                _p(np.ascontiguousarray(g.child1, np.int32)), _p(np.ascontiguousarray(g.child2, np.int32)),
                _p(np.ascontiguousarray(g.area, np.float32)), None if st is None else _p(st),
                int(mode), int(max_iter), float(tol), 1 if use_double else 0, _p(out), _p(res), _p(en),
                ctypes.byref(it), *bump_args)
    n = it.value
    if bumped is not None and bump_out is not None:
        bump_out[...] = keep[3]
    return out, res[:n], en[:n], n


def gather_light(row_ptr, src, w, shoot, threads=None) -> np.ndarray:
    """Radiosity.GatherLight over every row (FP32, transfer order), rows
    shared over threads: add (P, 3) = per row sum of shoot[src] * w."""
    L = _bind()
    L.vc_gather_light.argtypes = [ctypes.c_int32] + [ctypes.c_void_p] * 5 + [ctypes.c_int32]
    rp = np.ascontiguousarray(row_ptr, np.int64)
    P = len(rp) - 1
    sh = np.ascontiguousarray(shoot, np.float32).reshape(P, 3)
    out = np.zeros((P, 3), np.float32)
    L.vc_gather_light(P, _p(rp), _p(np.ascontiguousarray(src, np.int32)), _p(np.ascontiguousarray(w, np.float32)),
                      _p(sh), _p(out), int(threads or os.cpu_count() or 1))
    return out


def patch_lights(g, faces, light, sample_area):
    """PatchLighting.BuildPatchLights with AddSampleToPatch
    (Rad/Light/PatchLighting.cs:65-249), numpy, in the reference's float order.

    ``light`` is each sample's flat style-0 lighting (S, >=3); ``faces`` has the
    patch system's face order.  Every leaf patch that is not sky takes, in
    sample order, each sample of its face whose average light is at least 1 and
    whose position grown by sqrt(area)/2 (a float sqrt, a double halving
    narrowed) overlaps the patch winding's bounds: SampleArea += area,
    SampleLight += area * light.  A leaf's direct light is SampleLight *
    (float)(1.0 / SampleArea) (0 with no sample); a parent's is the
    area-weighted blend of its children's, children first (:133-148).
    Returns (P, 3) float32."""
    f32 = np.float32
    light = np.asarray(light, f32)[:, :3]
    area = np.asarray(sample_area, f32)
    pos = np.asarray(faces.sample_pos, f32)
    mins = g.winding.astype(f32).min(axis=1)
    maxs = g.winding.astype(f32).max(axis=1)
    avg_ok = ~(((light[:, 0] + light[:, 1]) + light[:, 2]) / f32(3) < f32(1))
    radius = (np.sqrt(area).astype(np.float64) / 2.0).astype(f32)
    direct = np.zeros((g.npatch, 3), f32)
    sky = g.face_sky[g.face].astype(bool)
    for p in np.nonzero((g.child1 < 0) & ~sky)[0]:
        f = g.face[p]
        a0, a1 = faces.face_sample_ptr[f], faces.face_sample_ptr[f + 1]
        if a1 == a0:
            continue
        ps, rs = pos[a0:a1], radius[a0:a1, None]
        inside = ~np.any((mins[p] > ps + rs) | (maxs[p] < ps - rs), axis=1)
        m = inside & avg_ok[a0:a1]
        if not m.any():
            continue
        sa = np.cumsum(np.where(m, area[a0:a1], f32(0)), dtype=f32)[-1]
        contrib = np.where(m[:, None], area[a0:a1, None] * light[a0:a1], f32(0)).astype(f32)
        sl = np.cumsum(contrib, axis=0, dtype=f32)[-1]
        direct[p] = sl * f32(1.0 / np.float64(sa))
    return g.pull_up(direct)
