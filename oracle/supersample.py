"""Supersampling restated -- TEST INFRASTRUCTURE ONLY (the oracle).

FaceLightJob's supersampling passes for one light style
(SourceSharp.MapTools/Rad/Light/FaceLightJob.cs):

* ``StyleSupersample.ComputeLuxelIntensity`` (:1091-1104): per normal, the
  light's ``Intensity()`` (R + G + B, LightingValue.cs:49) taken to a linear
  perception space, DetMath.PowToSingle(I / 256.0, 1.0 / 2.2): correctly
  rounded to double, then to float (vradcpu_math.c);
* ``ComputeGradients`` / ``Step`` (:1131-1207): the largest absolute step to
  the eight neighbouring luxels, every normal, in the reference's order; a
  neighbour without a sample is left out (Correct) or reads as infinity (the
  stock reading of uninitialised stack, ``garbage=True``, :1000-1018);
* ``SelectForPass`` (:1109-1130): unprocessed samples whose gradient is not
  below ``SupersampleGradient`` 0.0625 (LightConstants.cs:135) are selected
  and marked processed;
* the pass loop (:698-733): passes 1..``ExtraPasses`` (4,
  DirectLightingSettings.cs:45) while the last selection was not empty;
* ``EmitSupersamplePoint`` (:786-851): 4 x 4 direct points at luxel offsets
  ``csshift + c * cscale`` (cscale 1/4, csshift -0.375) and 2 x 2 ambient
  points (cscale 1/2, csshift -0.25) around ``WorldToLuxel(sample)``, each
  ``LuxelToWorld``; the direct points see every light but sky ambient, the
  ambient points only sky ambient (EmitResample, :465-518), every light's
  record added (ResolveResample, :526-555, no all-zero skip);
* ``ResolveSupersamplePoint`` / ``ResolveSupersampleItem`` (:740-775,
  :856-893): the points summed in row order, then
  ``v = direct * (1/16) + ambient * (1/4)`` (``AddWeighted``), and the
  sample's intensity recomputed.

The sample points take their sample's cluster (every point lies within half
a luxel of the sample, on its face) and every point counts (the samples here
carry no partial-luxel winding, so ``PointsInWinding`` never drops one).
"""
from __future__ import annotations

import numpy as np

from .gather import gather

f32 = np.float32
GRADIENT = f32(0.0625)
EXTRA_PASSES = 4

# StyleSupersample.ComputeGradients: the neighbours in the order the reference
# steps them, as (ds, dt)
NEIGHBOURS = ((-1, -1), (0, -1), (1, -1), (-1, 1), (0, 1), (1, 1), (-1, 0), (1, 0))


def intensity(light: np.ndarray, normal_count: np.ndarray) -> np.ndarray:
    """ComputeLuxelIntensity for every sample: (4, S) float32, slots past a
    sample's normal count zero."""
    L = light.astype(np.float32)
    i = (L[..., 0] + L[..., 1]) + L[..., 2]
    from .gather import pow_to_single

    x = i.astype(np.float64) / 256.0
    out = pow_to_single(x.reshape(-1), 1.0 / 2.2).reshape(x.shape)
    out[np.arange(4)[:, None] >= np.asarray(normal_count)[None, :]] = 0
    return out


def gradients(inten, lux, lux_ptr, face, s, t, width, height, normal_count, garbage=False):
    """Step's maximum per sample (float32), every sample (the caller skips the
    processed ones)."""
    S = len(face)
    g = np.zeros(S, np.float32)
    w, h = width[face], height[face]
    for n in range(4):
        live = normal_count > n
        c = inten[n]
        for ds, dt in NEIGHBOURS:
            ss, tt = s + ds, t + dt
            inb = live & (ss >= 0) & (ss < w) & (tt >= 0) & (tt < h)
            j = np.where(inb, lux_ptr[face] + np.clip(ss, 0, None) + np.clip(tt, 0, None) * w, 0)
            nb = np.where(inb, lux[j], -1)
            has = nb >= 0
            v = np.abs(c - inten[n][np.maximum(nb, 0)]).astype(np.float32)
            step = np.where(g > v, g, v)  # the max macro, a > b ? a : b
            g = np.where(inb & has, step, g)
            if garbage:
                g = np.where(inb & ~has, np.float32(np.inf), g)
    return g


def _world_to_luxel(fr, p):
    pos = p.astype(np.float32) - fr["origin"]
    ws, wt = fr["w2l_s"], fr["w2l_t"]
    s = ((pos[:, 0] * ws[:, 0]) + (pos[:, 1] * ws[:, 1])) + (pos[:, 2] * ws[:, 2])
    t = ((pos[:, 0] * wt[:, 0]) + (pos[:, 1] * wt[:, 1])) + (pos[:, 2] * wt[:, 2])
    return s - fr["min_s"].astype(np.float32), t - fr["min_t"].astype(np.float32)


def _luxel_to_world(fr, s, t):
    s = s + fr["min_s"].astype(np.float32)
    t = t + fr["min_t"].astype(np.float32)
    pos = fr["origin"] + fr["l2w_s"] * s[:, None]
    return pos + fr["l2w_t"] * t[:, None]


def offsets(width: int) -> np.ndarray:
    """csshift + c * cscale for c < width (4 direct, 2 ambient)."""
    cscale = f32(1.0) / f32(width)
    csshift = f32(-(float(f32(f32(width) - f32(1)) * cscale)) / 2.0)
    return np.array([csshift + f32(c) * cscale for c in range(width)], np.float32)


def subsample_points(frames, face, pos):
    """(direct (n, 16, 3), ambient (n, 4, 3)) for samples at ``pos`` on faces
    ``face``: direct point 4 * row + lane at (a[row], a[lane]), ambient lane
    l at (b[l >> 1], b[l & 1])."""
    fr = frames[face]
    os_, ot = _world_to_luxel(fr, pos)
    a, b = offsets(4), offsets(2)
    direct = np.stack([_luxel_to_world(fr, os_ + a[r], ot + a[l]) for r in range(4) for l in range(4)], 1)
    amb = np.stack([_luxel_to_world(fr, os_ + b[l >> 1], ot + b[l & 1]) for l in range(4)], 1)
    return direct.astype(np.float32), amb.astype(np.float32)


def _point_faces(fset, sel_face, points):
    """A FaceSet of the selected samples' points grouped by their face (in
    sample order), each group carrying its face's normals."""
    from hl2sdk_tools_sharp_amd.light import FaceSet

    k = points.shape[1]
    faces, first = np.unique(sel_face, return_index=True)
    order = np.argsort(first)
    faces = faces[order]
    cnt = np.array([np.count_nonzero(sel_face == f) for f in faces], np.int64) * k
    ptr = np.zeros(len(faces) + 1, np.int32)
    np.cumsum(cnt, out=ptr[1:])
    return FaceSet(ptr, fset.face_normal_count[faces].copy(), fset.face_normals[faces].copy(),
                   points.reshape(-1, 3).copy(), np.zeros(len(points) * k, np.int32))


def supersample(casters, fset, lights, light_pvs, sun_extent, sun_samples, sky_dirs, frames, light,
                extra_passes: int = EXTRA_PASSES, sample_st=None, garbage: bool = False, which: int = 0,
                threads=None, face_style=None, skybox=None):
    """The passes over one gather result ``light`` ((4, S, 4), the LDR or HDR
    result named by ``which``, which picks the resample's matching output).
    ``frames`` are the faces' luxel frames (final.FACE_DTYPE; width x height
    luxels); ``sample_st`` (S, 2) each sample's luxel, or None for sample k
    of a face at (k % width, k // width).  ``face_style`` (F,) makes this
    one style slot's StyleSupersample (FaceLightJob.cs:688-700): per face
    the slot's light style, 255 where the face has no such slot (its samples
    are never selected); the resample then takes only that style's lights
    (EmitResample, FaceLightJob.cs:486).  ``skybox`` (skybox.Skybox) gives
    the resample's sky tests their 3D-skybox stage, each point looking up its
    own leaf.  Returns (refined light, the number of samples each pass
    resampled)."""
    from hl2sdk_tools_sharp_amd.light import pack_desc

    light = np.array(light, np.float32, copy=True)
    ptr = np.asarray(fset.face_sample_ptr, np.int64)
    F, S = fset.nfaces, fset.nsamples
    nc_f = np.asarray(fset.face_normal_count, np.int64)
    face = np.repeat(np.arange(F), np.diff(ptr))
    width = frames["width"].astype(np.int64)
    height = frames["height"].astype(np.int64)
    if sample_st is None:
        k = np.arange(S) - ptr[face]
        s, t = k % width[face], k // width[face]
    else:
        s, t = np.asarray(sample_st, np.int64)[:, 0], np.asarray(sample_st, np.int64)[:, 1]
    lux_ptr = np.zeros(F + 1, np.int64)
    np.cumsum(width * height, out=lux_ptr[1:])
    lux = np.full(max(int(lux_ptr[-1]), 1), -1, np.int64)
    lux[lux_ptr[face] + s + t * width[face]] = np.arange(S)
    nc = nc_f[face]
    inten = intensity(light, nc)
    processed = np.zeros(S, bool)
    counts = []
    for _pass in range(1, extra_passes + 1):
        g = gradients(inten, lux, lux_ptr, face, s, t, width, height, nc, garbage)
        pick = ~processed & ~(g < GRADIENT)
        if face_style is not None:
            pick &= np.asarray(face_style)[face] != 255
        sel = np.nonzero(pick)[0]
        if len(sel) == 0:
            break
        processed[sel] = True
        counts.append(len(sel))
        direct, amb = subsample_points(frames, face[sel], fset.sample_pos[sel])
        res = []
        for pts, filt in ((direct, 2), (amb, 1)):
            sub = _point_faces(fset, face[sel], pts)
            sub.sample_cluster[:] = np.repeat(fset.sample_cluster[sel], pts.shape[1])
            desc = pack_desc(sub, lights, light_pvs, sun_extent, sun_samples, sky_dirs, skybox=skybox)
            fsf = None
            if face_style is not None:
                sub_faces = np.unique(face[sel], return_index=True)
                sub_faces = sub_faces[0][np.argsort(sub_faces[1])]
                fsf = np.asarray(face_style, np.int32)[sub_faces]
            ldr, hdr, _ = gather(casters, desc, sub.nsamples, threads, light_filter=filt, no_skip=True,
                                 face_style_filter=fsf)
            out = hdr if which else ldr
            k = pts.shape[1]
            acc = np.zeros((4, len(sel), 4), np.float32)
            for p in range(k):
                acc = acc + out[:, p::k][:, : len(sel)]
            res.append(acc)
        v = np.zeros((4, len(sel), 4), np.float32) + f32(1.0 / 16.0) * res[0]
        v = v + f32(0.25) * res[1]
        live = np.arange(4)[:, None] < nc[sel][None, :]
        cur = light[:, sel]
        light[:, sel] = np.where(live[..., None], v, cur)
        inten[:, sel] = intensity(light[:, sel], nc[sel])
    return light, counts


def supersample_styles(casters, fset, lights, light_pvs, sun_extent, sun_samples, sky_dirs, frames, slots,
                       face_styles, extra_passes: int = EXTRA_PASSES, sample_st=None, garbage: bool = False,
                       which: int = 0, threads=None, skybox=None):
    """Every style slot's passes: ``slots`` (4 slots, 4 normals, S, 4) the
    gather's light per slot, ``face_styles`` (F, 4) its style numbers.  Each
    slot is a StyleSupersample of its own (its own selection and passes).
    Returns (refined slots, samples re-lit per pass summed over the slots)."""
    out = np.array(slots, np.float32, copy=True)
    total = []
    for k in range(4):
        fs = np.asarray(face_styles)[:, k]
        if not np.any(fs != 255):
            continue
        out[k], counts = supersample(casters, fset, lights, light_pvs, sun_extent, sun_samples, sky_dirs, frames,
                                     out[k], extra_passes, sample_st, garbage, which, threads, face_style=fs,
                                     skybox=skybox)
        for i, n in enumerate(counts):
            if i < len(total):
                total[i] += n
            else:
                total.append(n)
    return out, total
