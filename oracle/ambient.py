"""CPU restatement of vrad's leaf-ambient ray fan -- TEST INFRASTRUCTURE ONLY.

AmbientSampler.ComputeRayCube (Rad/Ambient/AmbientSampler.cs) for one light
style: from a sample point, one ray along each of the 162 vertex normals
(Rad/Ambient/VertexNormals.cs) out to AmbientCube.RayLength; each ray's
light by RayAmbientLighting.Accumulate (Rad/Ambient/RayAmbientLighting.cs):
the hit face's average colour and the luxel under the hit, blended by the
cone's radius at the hit (RemapValClamped 20..40), both tinted by the face's
reflectivity, or the sky-ambient colour for a sky face; then
AmbientCube.Project (Rad/Ambient/AmbientCube.cs) onto the six box directions.

The hit surface and fraction come from the closest-hit seam (the oracle's
BVH closest hit, pinned to the reference's KD closest hit) rather than from
AmbientRayTracer's BSP walk, whose splits the synthetic maps do not have.
Then the flagged world lights (AmbientSampler.AddSurfaceLights with
AmbientCube.AddEmitSurfaceLights, classified by LeafAmbientSurfaceLights) are
added where the segment to the light is open.  Arithmetic is numpy float32, one IEEE operation at a time in
the reference's order; sums over the 162 rays run in ray order.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32

# VertexNormals.All: the 162 directions of the fan, in the reference's order
# (the anorms table; the summation order of AmbientCube.Project follows it).
VERTEX_NORMALS = np.array([
    (-0.525731, 0.000000, 0.850651), (-0.442863, 0.238856, 0.864188), (-0.295242, 0.000000, 0.955423),
    (-0.309017, 0.500000, 0.809017), (-0.162460, 0.262866, 0.951056), (0.000000, 0.000000, 1.000000),
    (0.000000, 0.850651, 0.525731), (-0.147621, 0.716567, 0.681718), (0.147621, 0.716567, 0.681718),
    (0.000000, 0.525731, 0.850651), (0.309017, 0.500000, 0.809017), (0.525731, 0.000000, 0.850651),
    (0.295242, 0.000000, 0.955423), (0.442863, 0.238856, 0.864188), (0.162460, 0.262866, 0.951056),
    (-0.681718, 0.147621, 0.716567), (-0.809017, 0.309017, 0.500000), (-0.587785, 0.425325, 0.688191),
    (-0.850651, 0.525731, 0.000000), (-0.864188, 0.442863, 0.238856), (-0.716567, 0.681718, 0.147621),
    (-0.688191, 0.587785, 0.425325), (-0.500000, 0.809017, 0.309017), (-0.238856, 0.864188, 0.442863),
    (-0.425325, 0.688191, 0.587785), (-0.716567, 0.681718, -0.147621), (-0.500000, 0.809017, -0.309017),
    (-0.525731, 0.850651, 0.000000), (0.000000, 0.850651, -0.525731), (-0.238856, 0.864188, -0.442863),
    (0.000000, 0.955423, -0.295242), (-0.262866, 0.951056, -0.162460), (0.000000, 1.000000, 0.000000),
    (0.000000, 0.955423, 0.295242), (-0.262866, 0.951056, 0.162460), (0.238856, 0.864188, 0.442863),
    (0.262866, 0.951056, 0.162460), (0.500000, 0.809017, 0.309017), (0.238856, 0.864188, -0.442863),
    (0.262866, 0.951056, -0.162460), (0.500000, 0.809017, -0.309017), (0.850651, 0.525731, 0.000000),
    (0.716567, 0.681718, 0.147621), (0.716567, 0.681718, -0.147621), (0.525731, 0.850651, 0.000000),
    (0.425325, 0.688191, 0.587785), (0.864188, 0.442863, 0.238856), (0.688191, 0.587785, 0.425325),
    (0.809017, 0.309017, 0.500000), (0.681718, 0.147621, 0.716567), (0.587785, 0.425325, 0.688191),
    (0.955423, 0.295242, 0.000000), (1.000000, 0.000000, 0.000000), (0.951056, 0.162460, 0.262866),
    (0.850651, -0.525731, 0.000000), (0.955423, -0.295242, 0.000000), (0.864188, -0.442863, 0.238856),
    (0.951056, -0.162460, 0.262866), (0.809017, -0.309017, 0.500000), (0.681718, -0.147621, 0.716567),
    (0.850651, 0.000000, 0.525731), (0.864188, 0.442863, -0.238856), (0.809017, 0.309017, -0.500000),
    (0.951056, 0.162460, -0.262866), (0.525731, 0.000000, -0.850651), (0.681718, 0.147621, -0.716567),
    (0.681718, -0.147621, -0.716567), (0.850651, 0.000000, -0.525731), (0.809017, -0.309017, -0.500000),
    (0.864188, -0.442863, -0.238856), (0.951056, -0.162460, -0.262866), (0.147621, 0.716567, -0.681718),
    (0.309017, 0.500000, -0.809017), (0.425325, 0.688191, -0.587785), (0.442863, 0.238856, -0.864188),
    (0.587785, 0.425325, -0.688191), (0.688191, 0.587785, -0.425325), (-0.147621, 0.716567, -0.681718),
    (-0.309017, 0.500000, -0.809017), (0.000000, 0.525731, -0.850651), (-0.525731, 0.000000, -0.850651),
    (-0.442863, 0.238856, -0.864188), (-0.295242, 0.000000, -0.955423), (-0.162460, 0.262866, -0.951056),
    (0.000000, 0.000000, -1.000000), (0.295242, 0.000000, -0.955423), (0.162460, 0.262866, -0.951056),
    (-0.442863, -0.238856, -0.864188), (-0.309017, -0.500000, -0.809017), (-0.162460, -0.262866, -0.951056),
    (0.000000, -0.850651, -0.525731), (-0.147621, -0.716567, -0.681718), (0.147621, -0.716567, -0.681718),
    (0.000000, -0.525731, -0.850651), (0.309017, -0.500000, -0.809017), (0.442863, -0.238856, -0.864188),
    (0.162460, -0.262866, -0.951056), (0.238856, -0.864188, -0.442863), (0.500000, -0.809017, -0.309017),
    (0.425325, -0.688191, -0.587785), (0.716567, -0.681718, -0.147621), (0.688191, -0.587785, -0.425325),
    (0.587785, -0.425325, -0.688191), (0.000000, -0.955423, -0.295242), (0.000000, -1.000000, 0.000000),
    (0.262866, -0.951056, -0.162460), (0.000000, -0.850651, 0.525731), (0.000000, -0.955423, 0.295242),
    (0.238856, -0.864188, 0.442863), (0.262866, -0.951056, 0.162460), (0.500000, -0.809017, 0.309017),
    (0.716567, -0.681718, 0.147621), (0.525731, -0.850651, 0.000000), (-0.238856, -0.864188, -0.442863),
    (-0.500000, -0.809017, -0.309017), (-0.262866, -0.951056, -0.162460), (-0.850651, -0.525731, 0.000000),
    (-0.716567, -0.681718, -0.147621), (-0.716567, -0.681718, 0.147621), (-0.525731, -0.850651, 0.000000),
    (-0.500000, -0.809017, 0.309017), (-0.238856, -0.864188, 0.442863), (-0.262866, -0.951056, 0.162460),
    (-0.864188, -0.442863, 0.238856), (-0.809017, -0.309017, 0.500000), (-0.688191, -0.587785, 0.425325),
    (-0.681718, -0.147621, 0.716567), (-0.442863, -0.238856, 0.864188), (-0.587785, -0.425325, 0.688191),
    (-0.309017, -0.500000, 0.809017), (-0.147621, -0.716567, 0.681718), (-0.425325, -0.688191, 0.587785),
    (-0.162460, -0.262866, 0.951056), (0.442863, -0.238856, 0.864188), (0.162460, -0.262866, 0.951056),
    (0.309017, -0.500000, 0.809017), (0.147621, -0.716567, 0.681718), (0.000000, -0.525731, 0.850651),
    (0.425325, -0.688191, 0.587785), (0.587785, -0.425325, 0.688191), (0.688191, -0.587785, 0.425325),
    (-0.955423, 0.295242, 0.000000), (-0.951056, 0.162460, 0.262866), (-1.000000, 0.000000, 0.000000),
    (-0.850651, 0.000000, 0.525731), (-0.955423, -0.295242, 0.000000), (-0.951056, -0.162460, 0.262866),
    (-0.864188, 0.442863, -0.238856), (-0.951056, 0.162460, -0.262866), (-0.809017, 0.309017, -0.500000),
    (-0.864188, -0.442863, -0.238856), (-0.951056, -0.162460, -0.262866), (-0.809017, -0.309017, -0.500000),
    (-0.681718, 0.147621, -0.716567), (-0.681718, -0.147621, -0.716567), (-0.850651, 0.000000, -0.525731),
    (-0.688191, 0.587785, -0.425325), (-0.587785, 0.425325, -0.688191), (-0.425325, 0.688191, -0.587785),
    (-0.425325, -0.688191, -0.587785), (-0.587785, -0.425325, -0.688191), (-0.688191, -0.587785, -0.425325),
], np.float32)

BOX = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], np.float32)
RAY_LENGTH = f32(32768.0) * f32(1.74)                     # AmbientCube.RayLength
CONE_ANGLE = f32(7.275) * (f32(math.pi) / f32(180.0))     # VertexNormals.ConeInnerAngleRadians
TAN_THETA = f32(math.tan(float(CONE_ANGLE)))              # DetMathF.Tan, correctly rounded

LIGHT_DTYPE = np.dtype([
    ("origin", "<f4", 3), ("normal", "<f4", 3), ("intensity", "<f4", 3), ("radius", "<f4"),
    ("type", "<i4"), ("style", "<i4"),
])
EMIT_SURFACE = 0

FACE_DTYPE = np.dtype([
    ("s", "<f4", 4), ("t", "<f4", 4), ("min_s", "<f4"), ("min_t", "<f4"), ("smax", "<i4"),
    ("tmax", "<i4"), ("light_ofs", "<i4"), ("sky", "<i4"), ("reflectivity", "<f4", 3),
])


def power2n() -> np.ndarray:
    """StockLightColor's Power2N: (float)(2^(i-128) / 255.0) in double
    (Rad/Ambient/StockLightColor.cs:262-271)."""
    return np.array([np.ldexp(1.0, i - 128) / 255.0 for i in range(256)], np.float64).astype(np.float32)


def tex_light_to_linear(words: np.ndarray) -> np.ndarray:
    """TexLightToLinear per ColorRGBExp32 word ((n, 4) uint8): mantissa *
    Power2N[exponent + 128]."""
    w = np.asarray(words, np.uint8).reshape(-1, 4)
    p = power2n()[w[:, 3].view(np.int8).astype(np.int64) + 128]
    return w[:, :3].astype(np.float32) * p[:, None]


def remap_clamped(val, a, b, c, d):
    """RemapValClamped (RayAmbientLighting.cs:318-330), float32."""
    val = np.asarray(val, np.float32)
    a, b, c, d = f32(a), f32(b), f32(c), f32(d)
    cv = (val - a) / (b - a)
    cv = np.minimum(np.maximum(cv, f32(0)), f32(1))
    return c + (d - c) * cv


def project(rad: np.ndarray) -> np.ndarray:
    """AmbientCube.Project: rad (n, 162, 3) -> (n, 6, 3); per box direction
    the cosine-weighted mean of the positive-cosine rays, summed in ray
    order, then times (1 / t)."""
    n = rad.shape[0]
    cube = np.zeros((n, 6, 3), np.float32)
    for j in range(5, -1, -1):
        t = f32(0)
        s = np.zeros((n, 3), np.float32)
        for i in range(162):
            a, b = VERTEX_NORMALS[i], BOX[j]
            c = f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2])
            if c > 0:
                t = f32(t + c)
                s = s + rad[:, i] * c
        cube[:, j] = s * (f32(1) / t)
    return cube


def face_averages(faces: np.ndarray, lump: np.ndarray) -> np.ndarray:
    """AmbientScene's AverageTinted for style 0: the face's average word
    (light_ofs - 4) decoded and times the reflectivity; zero for sky faces and
    faces without a lightmap."""
    out = np.zeros((len(faces), 3), np.float32)
    ok = (faces["sky"] == 0) & (faces["light_ofs"] >= 4)
    idx = np.nonzero(ok)[0]
    if len(idx):
        w = lump[(faces["light_ofs"][idx][:, None] - 4 + np.arange(4)).astype(np.int64)]
        out[idx] = tex_light_to_linear(w) * faces["reflectivity"][idx]
    return out


def ray_colors(faces: np.ndarray, sky_ambient, lump: np.ndarray, points: np.ndarray, surface: np.ndarray,
               fraction: np.ndarray) -> np.ndarray:
    """RayAmbientLighting.Accumulate for every (point, direction): surface
    (n, 162) the caller's triangle id (-1 for a miss), fraction (n, 162).
    Returns the rays' style-0 colours (n, 162, 3)."""
    n = points.shape[0]
    pts = np.asarray(points, np.float32)
    delta = VERTEX_NORMALS * RAY_LENGTH                           # (162, 3)
    length = np.sqrt(f32(f32(delta[:, 0] * delta[:, 0]) + f32(delta[:, 1] * delta[:, 1]))
                     + f32(delta[:, 2] * delta[:, 2]))
    color = np.zeros((n, 162, 3), np.float32)
    surface = np.asarray(surface, np.int64)
    frac = np.asarray(fraction, np.float32)
    face = np.where(surface >= 0, surface & 0x00FFFFFF, -1)
    hit = (face >= 0) & (face < len(faces))
    fidx = np.where(hit, face, 0)
    sky = hit & (faces["sky"][fidx] != 0)
    # the cone's radius where it meets the surface
    dist = (length[None, :] * TAN_THETA) * frac
    s_avg = remap_clamped(dist, 20.0, 40.0, 0.0, 1.0)
    s_avg = np.where(sky, f32(1), s_avg).astype(np.float32)
    s_smp = (f32(1) - s_avg).astype(np.float32)
    amb = None if sky_ambient is None else np.asarray(sky_ambient, np.float32)
    avg = face_averages(faces, lump)
    # AddFaceAverage
    use_avg = hit & (s_avg != 0)
    add = np.where(sky[..., None], (amb if amb is not None else np.zeros(3, np.float32)) * s_avg[..., None],
                   avg[fidx] * s_avg[..., None])
    if amb is None:
        use_avg &= ~sky
    color = np.where(use_avg[..., None], color + add, color)
    # AddFacePointSample: the luxel under the hit, truncated and clamped
    ofs = faces["light_ofs"][fidx]
    use_smp = hit & (s_smp != 0) & (ofs != -1) & ~sky
    pt = pts[:, None, :] + frac[..., None] * delta[None, :, :]
    S, T = faces["s"][fidx], faces["t"][fidx]
    sc = (((pt[..., 0] * S[..., 0] + pt[..., 1] * S[..., 1]) + pt[..., 2] * S[..., 2]) + S[..., 3])
    tc = (((pt[..., 0] * T[..., 0] + pt[..., 1] * T[..., 1]) + pt[..., 2] * T[..., 2]) + T[..., 3])
    ds = sc - faces["min_s"][fidx]
    dt = tc - faces["min_t"][fidx]
    smax, tmax = faces["smax"][fidx], faces["tmax"][fidx]
    with np.errstate(invalid="ignore"):
        di = np.clip(np.trunc(np.nan_to_num(ds)).astype(np.int64), 0, np.maximum(smax - 1, 0))
        ti = np.clip(np.trunc(np.nan_to_num(dt)).astype(np.int64), 0, np.maximum(tmax - 1, 0))
    index = ti * smax + di
    off = np.where(use_smp, ofs + index * 4, 0).astype(np.int64)
    words = lump[off[..., None] + np.arange(4)]
    c = tex_light_to_linear(words.reshape(-1, 4)).reshape(n, 162, 3) * faces["reflectivity"][fidx]
    color = np.where(use_smp[..., None], color + c * s_smp[..., None], color)
    return color.astype(np.float32)


def rays(points: np.ndarray) -> np.ndarray:
    """The fan's rays as 7-float records (origin, delta, reach 1), point-major."""
    n = points.shape[0]
    out = np.zeros((n, 162, 7), np.float32)
    out[:, :, :3] = np.asarray(points, np.float32)[:, None, :]
    out[:, :, 3:6] = (VERTEX_NORMALS * RAY_LENGTH)[None]
    out[:, :, 6] = 1.0
    return out.reshape(-1, 7)


def flagged_lights(lights) -> np.ndarray:
    """LeafAmbientSurfaceLights.Classify, Correct side (Rad/Ambient/LeafAmbientSurfaceLights.cs):
    surface lights of style 0 whose brightest channel times InvRSquared((0,
    0, 512)) is under 0.005, in light order."""
    if lights is None or len(lights) == 0:
        return np.zeros(0, LIGHT_DTYPE)
    L = np.asarray(lights, LIGHT_DTYPE)
    ratio = f32(1) / f32(512 * 512)
    m = np.maximum(np.maximum(L["intensity"][:, 0], L["intensity"][:, 1]), L["intensity"][:, 2])
    keep = (L["type"] == EMIT_SURFACE) & (L["style"] == 0) & (m * ratio < f32(0.005))
    return L[keep]


def segment_rays(points: np.ndarray, ends: np.ndarray) -> np.ndarray:
    """Ray.Segment (Tracing/IRayTracer.cs:71-88, Correct side) from every
    point to every end, point-major: origin, d / |d|, reach |d|."""
    p = np.asarray(points, np.float32)[:, None, :]
    e = np.asarray(ends, np.float32)[None, :, :]
    d = (e - p).astype(np.float32)
    length = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = f32(1) / length
    out = np.zeros(d.shape[:2] + (7,), np.float32)
    out[..., :3] = np.broadcast_to(p, d.shape)
    out[..., 3:6] = d * inv[..., None]
    out[..., 6] = length
    return out.reshape(-1, 7)


def add_surface_lights(cubes: np.ndarray, points: np.ndarray, flagged: np.ndarray, open_: np.ndarray) -> None:
    """AmbientCube.AddEmitSurfaceLights (Rad/Ambient/AmbientCube.cs) in
    place: open_ (n, lights) says which segments are unblocked."""
    for k, L in enumerate(flagged):
        delta = (L["origin"][None, :] - points).astype(np.float32)
        r2 = (delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1]) + delta[:, 2] * delta[:, 2]
        dscale = f32(1) / np.where(r2 < 1, f32(1), r2)
        if L["radius"] != 0:
            dscale = np.where(r2 > L["radius"] * L["radius"], f32(0), dscale)
        length = np.sqrt(r2)
        with np.errstate(divide="ignore", invalid="ignore"):
            dn = np.where((length != 0)[:, None], delta / length[:, None], f32(0)).astype(np.float32)
        dot = (dn[:, 0] * dn[:, 0] + dn[:, 1] * dn[:, 1]) + dn[:, 2] * dn[:, 2]
        dot2 = -((dn[:, 0] * L["normal"][0] + dn[:, 1] * L["normal"][1]) + dn[:, 2] * L["normal"][2])
        angle = np.where((dot < 0) | (dot2.astype(np.float64) <= 0.1 / 10), f32(0), dot * dot2)
        ratio = (dscale * angle) * f32(1)
        use = open_[:, k] & (ratio != 0)
        for j in range(6):
            b = BOX[j]
            t = (b[0] * dn[:, 0] + b[1] * dn[:, 1]) + b[2] * dn[:, 2]
            add = use & (t > 0)
            w = t * ratio
            cubes[:, j] = np.where(add[:, None], cubes[:, j] + L["intensity"][None, :] * w[:, None], cubes[:, j])


def ambient_cubes(casters, faces: np.ndarray, sky_ambient, lump: np.ndarray, points: np.ndarray,
                  threads: int = 1, bvh=None, lights=None) -> np.ndarray:
    """ComputeCube for every point: (n, 6, 3).  ``casters`` has ``vertices``
    (m, 9) and the caller's ``ids`` (the face index in their low 24 bits);
    ``bvh`` may be a CpuBvh already built over them; ``lights`` the pass's
    world lights (LIGHT_DTYPE)."""
    from . import CpuBvh, convert_triangles

    if bvh is None:
        bvh = CpuBvh(casters.vertices, convert_triangles(casters.vertices, casters.ids))
    points = np.asarray(points, np.float32).reshape(-1, 3)
    n = points.shape[0]
    idx, t = bvh.closest(rays(points), tmin=0.0, threads=threads)
    ids = np.asarray(casters.ids, np.int64)
    surf = np.where(idx >= 0, ids[np.maximum(idx, 0)], -1).reshape(n, 162)
    frac = np.where(idx >= 0, t, np.float32(1.0)).astype(np.float32).reshape(n, 162)
    rad = ray_colors(faces, sky_ambient, np.asarray(lump, np.uint8), points, surf, frac)
    cubes = project(rad)
    fl = flagged_lights(lights)
    if len(fl):
        seg = segment_rays(points, fl["origin"])
        bits = bvh.visibility(seg, tmin=0.0, threads=threads)
        m = seg.shape[0]
        blocked = ((bits[np.arange(m) >> 6] >> (np.arange(m) & 63).astype(np.uint64)) & np.uint64(1)).astype(bool)
        add_surface_lights(cubes, points, fl, ~blocked.reshape(n, len(fl)))
    return cubes
