"""The bounce stage, mirrored from the reference's ``IVradBounceStage``.

Reference (paths under the reference's ``src/``):

* ``IVradBounceStage.BounceAsync(RadPass, CancellationToken)`` --
  SourceSharp.MapTools/Rad/VradStages.cs:112-122, the stage seam a host may
  replace with "a GPU implementation";
* its built-in body ``RadWorld.BounceAsync`` -- Rad/Light/RadWorld.Bounce.cs:321-397:
  transfers from ``VisMatrix.BuildAsync`` (Rad/Bounce/VisMatrix.cs:219-620),
  then ``Radiosity.BounceAsync`` (Rad/Bounce/Radiosity.cs:132-316).

``HipBounceStage`` runs both halves in libvradhip.so on the GPU: the transfer
build against the tracer's scene, and the bounce with FP32 or FP16 weights,
FP32 accumulation and FP64 convergence sums.  There is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .tracer import HipRayTracer


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class BounceResult:
    """What one bounce produced: the bounced light per patch (parents pulled
    up), and per iteration the residual and the summed leaf emission -- the
    stock log's ``Bounce #%i added RGB(...)`` (Radiosity.BounceAsync returns
    the same list)."""

    total_light: np.ndarray  # (P, 3) float32
    residuals: np.ndarray  # (iters,)
    added: np.ndarray  # (iters, 3)
    iterations: int
    bump_light: Optional[np.ndarray] = None  # (P, 3, 3): TotalLight's bump directions, when bumped


class HipBounceStage:
    """``IVradBounceStage`` on the GPU.  ``patches`` is any object with the
    fields of :class:`hl2sdk_tools_sharp_amd.rooms.RoomGrid` (the reference's
    PatchSet as the bounce reads it)."""

    def __init__(self, tracer: HipRayTracer, patches):
        self.tracer = tracer
        self.patches = patches
        self._keep = {}
        self._loaded = False
        self.transfer_count = 0
        self.candidate_count = 0
        self.blocked_count = 0

    @property
    def _h(self):
        return self.tracer.handle

    def load_patches(self, bumped=None, bump_normals=None) -> None:
        """Uploads the patch system.  ``bumped`` (one flag per patch,
        Patch.NeedsBumpmap) with ``bump_normals`` ((P, 3, 3), each patch's bump
        basis) makes those receivers gather per normal."""
        g = self.patches
        k = self._keep
        k["origin"] = np.ascontiguousarray(g.origin, np.float32)
        k["normal"] = np.ascontiguousarray(g.normal, np.float32)
        k["plane_dist"] = np.ascontiguousarray(g.plane_dist, np.float32)
        k["area"] = np.ascontiguousarray(g.area, np.float32)
        k["child1"] = np.ascontiguousarray(g.child1, np.int32)
        k["child2"] = np.ascontiguousarray(g.child2, np.int32)
        k["next_parent"] = np.ascontiguousarray(g.next_parent, np.int32)
        k["face"] = np.ascontiguousarray(g.face, np.int32)
        k["cluster"] = np.ascontiguousarray(g.cluster, np.int32)
        k["winding"] = np.ascontiguousarray(g.winding, np.float32)
        k["reflectivity"] = np.ascontiguousarray(g.refl, np.float32)
        k["face_sky"] = np.ascontiguousarray(g.face_sky, np.uint8)
        k["face_parent"] = np.ascontiguousarray(g.face_parent, np.int32)
        k["row_faces_ptr"] = np.ascontiguousarray(g.row_faces_ptr, np.int32)
        k["row_faces"] = np.ascontiguousarray(g.row_faces, np.int32)
        k["receivers"] = np.ascontiguousarray(g.receivers, np.int32)
        d = _lib.VhPatchDesc(
            g.npatch, g.nface, g.ncluster, len(g.receivers),
            *[k[n].ctypes.data for n in (
                "origin", "normal", "plane_dist", "area", "child1", "child2", "next_parent",
                "face", "cluster", "winding", "reflectivity", "face_sky", "face_parent",
                "row_faces_ptr", "row_faces", "receivers")]
        )
        self._bumped = bumped is not None
        if bumped is not None:
            k["bumped"] = np.ascontiguousarray(bumped, np.uint8).reshape(g.npatch)
            k["bump_normals"] = np.ascontiguousarray(bump_normals, np.float32).reshape(g.npatch, 9)
            d.bumped = k["bumped"].ctypes.data
            d.bump_normals = k["bump_normals"].ctypes.data
        HipRayTracer._raise(_lib.lib().vh_patches_load(self._h, ctypes.byref(d)))
        self._loaded = True

    def build_transfers(self) -> int:
        """VisMatrix.BuildAsync on the GPU; returns the transfer count."""
        if not self._loaded:
            self.load_patches()
        t, c, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        HipRayTracer._raise(
            _lib.lib().vh_transfers_build(self._h, ctypes.byref(t), ctypes.byref(c), ctypes.byref(b))
        )
        self.transfer_count, self.candidate_count, self.blocked_count = t.value, c.value, b.value
        return t.value

    def load_transfers(self, row_ptr, src, weight) -> None:
        """Transfers built elsewhere (a TransferSet or the TRF1 cache)."""
        if not self._loaded:
            self.load_patches()
        rp = np.ascontiguousarray(row_ptr, np.int64)
        s = np.ascontiguousarray(src, np.int32)
        w = np.ascontiguousarray(weight, np.float32)
        HipRayTracer._raise(_lib.lib().vh_transfers_load(self._h, _p(rp), _p(s), _p(w)))
        self.transfer_count = int(rp[-1])

    def read_transfers(self):
        """The transfer lists as CSR (row_ptr (P+1), src, weight)."""
        P = self.patches.npatch
        rp = np.zeros(P + 1, np.int64)
        HipRayTracer._raise(_lib.lib().vh_transfers_read(self._h, _p(rp), None, None, 0))
        T = int(rp[-1])
        src = np.zeros(max(T, 1), np.int32)
        w = np.zeros(max(T, 1), np.float32)
        HipRayTracer._raise(_lib.lib().vh_transfers_read(self._h, _p(rp), _p(src), _p(w), max(T, 1)))
        return rp, src[:T], w[:T]

    def export_trf1(self, max_chunk_bytes: int = 64 << 20):
        """The built transfers in the reference's transfer-cache layout
        (TransferSetCodec.Index and Chunks, Rad/Bounce/TransferSetCodec.cs):
        (index blob, raw chunks).  The records are written on the device."""
        from . import trf1

        P = self.patches.npatch
        counts = np.zeros(max(P, 1), np.int32)
        HipRayTracer._raise(_lib.lib().vh_transfers_export(self._h, _p(counts), None, 0))
        counts = counts[:P]
        T = int(counts.astype(np.int64).sum())
        records = np.zeros(max(T, 1) * 2, np.uint32)
        HipRayTracer._raise(_lib.lib().vh_transfers_export(self._h, None, _p(records), max(T, 1)))
        return trf1.index(counts), trf1.chunks(records[:2 * T].tobytes(), max_chunk_bytes)

    def patch_lights(self, lighter, hdr: bool = False, read: bool = True):
        """PatchLighting.BuildPatchLights: the last gather of ``lighter`` (a
        HipFaceLighter over this patch system's faces, with sample areas)
        handed to the patches.  The result stays on the device for a
        ``bounce(None)``; returns it as (P, 3) when ``read``."""
        if not self._loaded:
            self.load_patches()
        f = lighter.faces
        if f is None or f.sample_area is None:
            raise ValueError("the lighter's faces carry no sample areas")
        area = np.ascontiguousarray(f.sample_area, np.float32)
        out = np.zeros((self.patches.npatch, 3), np.float32) if read else None
        HipRayTracer._raise(_lib.lib().vh_patch_lights(self._h, 1 if hdr else 0, _p(area), _p(out)))
        return out

    def bounce(self, emit0, start=None, weights: str = "f32", max_iter: int = 100, tol: float = 0.0,
               use_graph: bool = True, from_prediction: bool = False) -> BounceResult:
        """Radiosity.BounceAsync.  ``tol`` > 0 iterates to that relative residual;
        ``tol`` <= 0 keeps the reference's stop rule.  ``start`` (an estimate of
        the bounced light) switches to the from-start form; ``from_prediction``
        starts from the warm-start network's last prediction instead.
        ``emit0`` None starts from the device-resident result of
        ``patch_lights``."""
        P = self.patches.npatch
        e0 = None if emit0 is None else np.ascontiguousarray(emit0, np.float32).reshape(P, 3)
        st = None if start is None else np.ascontiguousarray(start, np.float32).reshape(P, 3)
        opts = _lib.VhBounceOpts(
            _lib.VH_WEIGHTS_F16 if weights == "f16" else _lib.VH_WEIGHTS_F32,
            _lib.VH_BOUNCE_FROM_PREDICTION if from_prediction
            else _lib.VH_BOUNCE_FROM_START if st is not None else _lib.VH_BOUNCE_PROGRESSIVE,
            int(max_iter), 1 if use_graph else 0, float(tol),
        )
        total = np.zeros((P, 3), np.float32)
        res = np.zeros(max_iter, np.float64)
        en = np.zeros((max_iter, 3), np.float64)
        it = ctypes.c_int32()
        HipRayTracer._raise(
            _lib.lib().vh_bounce(self._h, ctypes.byref(opts), _p(e0), _p(st), _p(total), _p(res), _p(en),
                                 ctypes.byref(it))
        )
        n = it.value
        bump = None
        if getattr(self, "_bumped", False):
            bump = np.zeros((P, 3, 3), np.float32)
            HipRayTracer._raise(_lib.lib().vh_bounce_bump_totals(self._h, bump.ctypes.data_as(ctypes.c_void_p)))
        return BounceResult(total, res[:n], en[:n], n, bump)

    async def bounce_async(self, emit0, **kw) -> BounceResult:
        import asyncio

        return await asyncio.to_thread(self.bounce, emit0, **kw)
