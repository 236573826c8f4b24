"""The lighting pass end to end on one GPU, in vrad's stage order.

The reference runs a pass as FaceLight (the per-luxel direct gather,
Rad/RadWorld.Facelights.cs), BuildPatchLights (the luxels' light handed to
the radiosity patches, Rad/Light/PatchLighting.cs), the transfer build
(Rad/Bounce/VisMatrix.cs), the bounce (Rad/Bounce/Radiosity.cs, behind
IVradBounceStage, Rad/VradStages.cs:112-122) and the lump write
(Rad/Final/FinalLightFace.cs with StockLightColor.Encode).  Here every stage
runs in libvradhip.so against one scene upload; the patch lights stay on the
device and feed the bounce without a host round trip, and the final stage
(the luxel radial filter, FinalLightFace.Run) reads the gather's samples and
the bounced light where they lie.  Without final-stage inputs the lump holds
each sample's direct light, encoded in sample order.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from .bounce import BounceResult, HipBounceStage
from .light import FaceSet, HipFaceLighter
from .rooms import patch_bumps
from .tracer import HipTracerFactory


@dataclass
class PassResult:
    """One lighting pass: per-stage wall seconds, the patches' direct light,
    the bounce, the lump bytes."""

    stage_s: Dict[str, float] = field(default_factory=dict)
    direct: Optional[np.ndarray] = None
    bounce: Optional[BounceResult] = None
    lump: Optional[np.ndarray] = None
    transfers: int = 0
    supersample_passes: list = field(default_factory=list)  # samples re-lit per supersampling pass

    @property
    def wall_s(self) -> float:
        return float(sum(self.stage_s.values()))


def light_pass(g, faces: FaceSet, lights: np.ndarray, light_pvs=None, *, hdr: bool = False,
               sun_extent: float = 0.0, sun_samples: int = 30, sky_dirs=None, ao_count: int = 0,
               ao_radius: float = 40.0, weights: str = "f32", tol: float = 1e-4, max_iter: int = 100,
               device: int = 0, read_direct: bool = False, schedule: str = "queue",
               warm_start: Optional[dict] = None, final_inputs=None, extra_passes: int = 0) -> PassResult:
    """Runs one pass (LDR, or HDR with ``hdr``) over room grid ``g`` whose
    faces, in the patch system's order, are ``faces``.  ``schedule`` is the
    gather's ('queue', or 'per_face' for one launch per face); ``warm_start``
    (the network's ``params``, ``gain`` and ``scale``) starts the bounce from
    the network's prediction instead of from zero.  ``final_inputs``
    ((FinalFaces, FinalPatches), e.g. rooms.room_final_inputs) makes the lump
    the final stage's: luxels filtered from the samples plus the bounced light
    of the patches round them, minlight, medians.  ``extra_passes`` > 0 runs
    that many supersampling passes after the gather, before the patch
    hand-off (FaceLightJob.Finish, Rad/Light/FaceLightJob.cs:935-960; the
    reference's default is 4, ``-noextra`` 0); they need ``final_inputs`` for
    the faces' luxel frames.  Raises if the HIP tracer declines (there is no
    CPU path)."""
    out = PassResult()
    clock = time.perf_counter

    t = clock()
    offer = HipTracerFactory(device=device).try_create(g.casters)
    if offer.tracer is None:
        raise RuntimeError(f"HIP tracer declined: {offer.decline_reason}")
    tr = offer.tracer
    out.stage_s["scene"] = clock() - t
    try:
        t = clock()
        fl = HipFaceLighter(tr)
        fl.load(faces, lights, light_pvs=light_pvs, sun_extent=sun_extent, sun_samples=sun_samples,
                sky_dirs=sky_dirs, ao_count=ao_count, ao_radius=ao_radius)
        fl.run(schedule)
        if extra_passes > 0:
            if final_inputs is None:
                raise ValueError("supersampling needs final_inputs for the faces' luxel frames")
            out.supersample_passes = fl.supersample(final_inputs[0].faces, which=1 if hdr else 0,
                                                    extra_passes=extra_passes)
        out.stage_s["facelight"] = clock() - t

        t = clock()
        st = HipBounceStage(tr, g)
        # bumped receivers gather per bump normal (Patch.NeedsBumpmap)
        if np.any(np.asarray(faces.face_normal_count) == 4):
            st.load_patches(*patch_bumps(g, faces))
        else:
            st.load_patches()
        out.direct = st.patch_lights(fl, hdr=hdr, read=read_direct)
        out.stage_s["patch_lights"] = clock() - t

        t = clock()
        out.transfers = st.build_transfers()
        out.stage_s["transfers"] = clock() - t

        t = clock()
        if warm_start is not None:
            from .mlp import WarmStartMLP

            net = WarmStartMLP(tr, params=warm_start["params"], gain=warm_start["gain"],
                               scale=warm_start["scale"])
            net.predict(None, read=False)
            out.bounce = st.bounce(None, weights=weights, tol=tol, max_iter=max_iter, from_prediction=True)
        else:
            out.bounce = st.bounce(None, weights=weights, tol=tol, max_iter=max_iter)
        out.stage_s["bounce"] = clock() - t

        t = clock()
        if final_inputs is not None:
            from .final import HipFinalLighter

            fin = HipFinalLighter(fl)
            fin.load(*final_inputs)
            out.lump = fin.run(hdr=hdr, bounce="device")
        else:
            out.lump = fl.lightmap(hdr=hdr)
        out.stage_s["lump"] = clock() - t
    finally:
        tr.dispose()
    return out
