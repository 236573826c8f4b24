"""CPU restatement of vrad's final-light stage -- TEST INFRASTRUCTURE ONLY.

FinalLightFace.Run (Rad/Final/FinalLightFace.cs:93-299) for faces that are
not displacements, one light style, with the luxel accumulation grid
LuxelRadial (Rad/Final/LuxelRadial.cs), on the Correct side of the
compliance switches the stage reads (SampleRadialEdgeOffByOne: a point one
past the last column or row is off the grid; PatchRadialNeighbourBumpFromSelf:
a neighbour's patches are bumped when the neighbour is).  Macro textures are
the identity (the synthetic maps have none).  Every float operation is a
numpy float32 scalar operation in the reference's order; the double
promotions the reference makes (EQUAL_EPSILON, 0.1, RADIALDIST, 1.0 / weight)
are made in Python floats (IEEE double).  Slow on purpose: a plain statement
of the arithmetic, for small inputs.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np

from .gather import encode_rgbe

f32 = np.float32
RADIAL_DIST2 = 2                   # RADIALDIST2, LuxelRadial.cs:42
RADIAL_DIST = 1.42                 # RADIALDIST (double), :45
WEIGHT_EPS = f32(0.00001)          # WEIGHT_EPS, :48
ONE_OVER_SQRT3 = f32(0.57735025882720947)  # OO_SQRT_3, :51
EQUAL_EPSILON_D = 0.001            # LightConstants.EqualEpsilonDouble
ERROR_RED = (f32(2550.0), f32(0.0), f32(0.0))


class Face:
    """FaceLightInfo's luxel frame (Rad/Light/FaceLightInfo.cs:264-290) and
    the face's final-light inputs."""

    def __init__(self, origin, w2l_s, w2l_t, l2w_s, l2w_t, min_s, min_t, width, height,
                 bumped=False, minlight=0.0, lit=True):
        self.origin = [f32(x) for x in origin]
        self.ws = [f32(x) for x in w2l_s]
        self.wt = [f32(x) for x in w2l_t]
        self.ls = [f32(x) for x in l2w_s]
        self.lt = [f32(x) for x in l2w_t]
        self.min_s, self.min_t = int(min_s), int(min_t)
        self.w, self.h = int(width), int(height)
        self.nb = 4 if bumped else 1
        self.minlight = f32(minlight)
        self.lit = bool(lit)

    def world_to_luxel(self, p):
        pos = [f32(f32(p[k]) - self.origin[k]) for k in range(3)]
        s = f32(f32(f32(pos[0] * self.ws[0]) + f32(pos[1] * self.ws[1])) + f32(pos[2] * self.ws[2]))
        t = f32(f32(f32(pos[0] * self.wt[0]) + f32(pos[1] * self.wt[1])) + f32(pos[2] * self.wt[2]))
        return f32(s - f32(self.min_s)), f32(t - f32(self.min_t))

    def luxel_to_world(self, s, t):
        s = f32(f32(s) + f32(self.min_s))
        t = f32(f32(t) + f32(self.min_t))
        pos = [f32(self.origin[k] + f32(self.ls[k] * s)) for k in range(3)]
        return [f32(pos[k] + f32(self.lt[k] * t)) for k in range(3)]


def _trunc(x) -> int:
    """C#'s (int) cast: toward zero."""
    return int(math.trunc(float(x)))


class Radial:
    """LuxelRadial over one face's grid (radial_t)."""

    def __init__(self, face: Face):
        self.face = face
        n = face.w * face.h
        self.weight = [f32(0.0)] * n
        self.light = [[[f32(0.0)] * 3 for _ in range(n)] for _ in range(4)]

    def add_direct(self, point, min_s, min_t, max_s, max_t, light, has_bump, nbr_bump):
        """AddDirect (LuxelRadial.cs:140-208)."""
        f = self.face
        cs, ct = f.world_to_luxel(point)
        min_s, min_t, max_s, max_t = f32(min_s), f32(min_t), f32(max_s), f32(max_t)
        s_min, t_min = _trunc(min_s), _trunc(min_t)
        s_max = _trunc(f32(max_s + f32(0.9999))) + 1
        t_max = _trunc(f32(max_t + f32(0.9999))) + 1
        s_min, t_min = max(s_min, 0), max(t_min, 0)
        s_max, t_max = min(s_max, f.w), min(t_max, f.h)
        for s in range(s_min, s_max):
            for t in range(t_min, t_max):
                a = float(f32(min_s - f32(s)))
                s0 = f32(a if a > -1.0 else -1.0)
                a = float(f32(min_t - f32(t)))
                t0 = f32(a if a > -1.0 else -1.0)
                a = float(f32(max_s - f32(s)))
                s1 = f32(a if a < 1.0 else 1.0)
                a = float(f32(max_t - f32(t)))
                t1 = f32(a if a < 1.0 else 1.0)
                area = f32(f32(s1 - s0) * f32(t1 - t0))
                if float(area) > EQUAL_EPSILON_D:
                    ds = f32(abs(f32(cs - f32(s))))
                    dt = f32(abs(f32(ct - f32(t))))
                    r = ds if ds > dt else dt
                    r = f32(float(area) / 0.1) if float(r) < 0.1 else f32(area / r)
                    i = s + t * f.w
                    if has_bump:
                        if nbr_bump:
                            for b in range(4):
                                self._add(b, i, light[b], r)
                        else:
                            self._add(0, i, light[0], r)
                            bw = f32(r * ONE_OVER_SQRT3)
                            for b in range(1, 4):
                                self._add(b, i, light[0], bw)
                    else:
                        self._add(0, i, light[0], r)
                    self.weight[i] = f32(self.weight[i] + r)

    def _add(self, b, i, c, r):
        """LightingValue.AddWeighted: Lighting += weight * src."""
        L = self.light[b][i]
        for k in range(3):
            L[k] = f32(L[k] + f32(r * f32(c[k])))

    def add_bounced(self, point, min_s, min_t, max_s, max_t, light, has_bump, nbr_bump):
        """AddBounced (LuxelRadial.cs:230-293)."""
        f = self.face
        cs, ct = f.world_to_luxel(point)
        dist_s = f32(f32(max_s) - f32(min_s))
        dist_t = f32(f32(max_t) - f32(min_t))
        dist_s = f32(1.0 if 1.0 > float(dist_s) else float(dist_s))
        dist_t = f32(1.0 if 1.0 > float(dist_t) else float(dist_t))
        s_min = _trunc(float(cs) - float(dist_s) * RADIAL_DIST)
        t_min = _trunc(float(ct) - float(dist_t) * RADIAL_DIST)
        s_max = _trunc(float(cs) + float(dist_s) * RADIAL_DIST + 1.0)
        t_max = _trunc(float(ct) + float(dist_t) * RADIAL_DIST + 1.0)
        s_min, t_min = max(s_min, 0), max(t_min, 0)
        s_max, t_max = min(s_max, f.w), min(t_max, f.h)
        for s in range(s_min, s_max):
            for t in range(t_min, t_max):
                ds = f32(f32(cs - f32(s)) / dist_s)
                dt = f32(f32(ct - f32(t)) / dist_t)
                r = f32(f32(RADIAL_DIST2) - f32(f32(ds * ds) + f32(dt * dt)))
                i = s + t * f.w
                if r > 0:
                    if has_bump:
                        if nbr_bump:
                            for b in range(4):
                                self._addc(b, i, light[b], r)
                        else:
                            self._addc(0, i, light[0], r)
                            bw = f32(r * ONE_OVER_SQRT3)
                            for b in range(1, 4):
                                self._addc(b, i, light[0], bw)
                    else:
                        self._addc(0, i, light[0], r)
                    self.weight[i] = f32(self.weight[i] + r)

    def _addc(self, b, i, c, r):
        """Lighting += r * light (Vec3 times float)."""
        L = self.light[b][i]
        for k in range(3):
            L[k] = f32(L[k] + f32(f32(c[k]) * r))

    def sample(self, point, nb, red_errors):
        """Sample (LuxelRadial.cs:314-359), the Correct edge test."""
        f = self.face
        cs, ct = f.world_to_luxel(point)
        u = _trunc(f32(cs + f32(0.5)))
        v = _trunc(f32(ct + f32(0.5)))
        i = u + v * f.w
        if u < 0 or u >= f.w or v < 0 or v >= f.h:
            return [list(ERROR_RED) for _ in range(nb)], False
        ok = True
        out = []
        for b in range(nb):
            wgt = self.weight[i] if i < f.w * f.h else f32(0.0)
            if wgt > WEIGHT_EPS:
                sc = f32(1.0 / float(wgt))
                out.append([f32(self.light[b][i][k] * sc) for k in range(3)])
            else:
                out.append(list(ERROR_RED) if red_errors else [f32(0.0)] * 3)
                if b == 0:
                    ok = False
        return out, ok


def patch_range(face: Face, winding):
    """PatchLightmapCoordRange (FinalLightFace.cs:519-536)."""
    mn_s, mn_t, mx_s, mx_t = f32(1e30), f32(1e30), f32(-1e30), f32(-1e30)
    for p in winding:
        s, t = face.world_to_luxel(p)
        mn_s = mn_s if mn_s < s else s
        mx_s = mx_s if mx_s > s else s
        mn_t = mn_t if mn_t < t else t
        mx_t = mx_t if mx_t > t else t
    return mn_s, mn_t, mx_s, mx_t


def face_patch_lists(nfaces: int, patch_face, patch_leaf) -> List[List[int]]:
    """Each face's leaf patches in the order FacePatches walks them: the list
    RebuildFaceLists makes by prepending every patch in index order, so the
    highest index first (Rad/Light/PatchSubdivider.cs:408-423), less the
    patches with children (FinalLightFace.cs:500-504)."""
    lists: List[List[int]] = [[] for _ in range(nfaces)]
    for p in range(len(patch_face) - 1, -1, -1):
        if patch_leaf[p]:
            lists[int(patch_face[p])].append(p)
    return lists


def final_light(faces: Sequence[Face], luxels: Sequence[np.ndarray], samples: Sequence[dict],
                neighbours: Sequence[Sequence[int]], patches: Optional[dict] = None,
                red_errors: bool = False, only=None, styles=None) -> List[Optional[dict]]:
    """Run's per-face loop (FinalLightFace.cs:93-299).

    samples[f]: {'pos' (n,3), 'bounds' (n,4) minS minT maxS maxT in face f's
    luxel space, 'light' (nb,n,3)}; luxels[f]: (numLuxels,3) world points;
    patches: None when there is no bounce, else {'origin' (P,3), 'winding'
    list of (k,3), 'face' (P,), 'leaf' (P,) bool, 'total' (P,4,3) TotalLight
    per normal}.  ``styles`` None: light style 0 alone; returns per face
    {'luxels' (nb, numLuxels, 3) after minlight, 'median' (3,)}.  ``styles``
    (F, 4) style numbers per slot (255 empty): samples[f]['light'] is then a
    list over the face's slots, each (nb,n,3); every style is lit in turn
    (a neighbour's samples of the slot with the same style NUMBER, bounced
    light in style 0 only) and each face gives {'luxels' (nstyles, nb,
    numLuxels, 3), 'median' (nstyles, 3)}.  None for a face that is not lit
    (or not in ``only``, a set of faces to run when given)."""
    out: List[Optional[dict]] = []
    plist = None
    if patches is not None:
        plist = face_patch_lists(len(faces), patches["face"], patches["leaf"])

    def nstyles(f):
        if styles is None:
            return 1
        return next((k for k in range(4) if styles[f][k] == 255), 4)

    def slot_light(f, k):
        return samples[f]["light"] if styles is None else samples[f]["light"][k]

    for fn, face in enumerate(faces):
        if not face.lit or nstyles(fn) == 0 or (only is not None and fn not in only):
            out.append(None)
            continue
        per = [_light_style(faces, luxels, samples, neighbours, patches, plist, red_errors, fn, k,
                            styles, slot_light) for k in range(nstyles(fn))]
        if styles is None:
            out.append(per[0])
        else:
            out.append({"luxels": np.stack([r["luxels"] for r in per]),
                        "median": np.stack([r["median"] for r in per])})
    return out


def _light_style(faces, luxels, samples, neighbours, patches, plist, red_errors, fn, k, styles, slot_light):
    """One face's style slot k."""
    face = faces[fn]
    nb = face.nb
    bumped = nb == 4
    rad = Radial(face)
    sm = samples[fn]
    own = slot_light(fn, k)
    wanted = 0 if styles is None else int(styles[fn][k])
    # BuildLuxelRadial (FinalLightFace.cs:301-389): own samples, then each
    # neighbour's samples of the same light style, their bounds carried
    # corner by corner into this grid
    for i in range(len(sm["pos"])):
        rad.add_direct(sm["pos"][i], *sm["bounds"][i], [own[b][i] for b in range(nb)], bumped, bumped)
    for n in neighbours[fn]:
        nface, ns = faces[n], samples[n]
        if styles is None:
            nk = 0
        else:
            nk = next((j for j in range(4) if int(styles[n][j]) == wanted), -1)
        if nk < 0 or not nface.lit or len(ns["pos"]) == 0:
            continue
        nl = slot_light(n, nk)
        nbump = nface.nb == 4
        for i in range(len(ns["pos"])):
            b0 = ns["bounds"][i]
            tmp = nface.luxel_to_world(b0[0], b0[1])
            mn_s, mn_t = face.world_to_luxel(tmp)
            tmp = nface.luxel_to_world(b0[2], b0[3])
            mx_s, mx_t = face.world_to_luxel(tmp)
            light = [nl[b][i] for b in range(nface.nb)]
            rad.add_direct(ns["pos"][i], mn_s, mn_t, mx_s, mx_t, light, bumped, nbump)
    prad = None
    if plist is not None and k == 0:
        # BuildPatchRadial (FinalLightFace.cs:404-416, 489-516), style 0 only
        # (:182): the face's leaf patches, then each neighbour's; a
        # neighbour's patches are bumped when that neighbour is
        prad = Radial(face)
        for src in [fn] + list(neighbours[fn]):
            nbump = bumped if src == fn else faces[src].nb == 4
            for p in plist[src]:
                mn_s, mn_t, mx_s, mx_t = patch_range(face, patches["winding"][p])
                light = [np.asarray(patches["total"][p][b], np.float32) for b in range(4)]
                prad.add_bounced(patches["origin"][p], mn_s, mn_t, mx_s, mx_t, light, bumped, nbump)
    lum = np.zeros((nb, len(luxels[fn]), 3), np.float32)
    red_, green_, blue_ = [], [], []
    no_samples = len(sm["pos"]) == 0
    for j in range(len(luxels[fn])):
        lb, ok = rad.sample(luxels[fn][j], nb, red_errors)
        if prad is not None:
            v, _ = prad.sample(luxels[fn][j], nb, red_errors)
            for b in range(nb):
                lb[b] = [f32(lb[b][c] + v[b][c]) for c in range(3)]
        if no_samples:
            lb = [[f32(255.0), f32(0.0), f32(0.0)] for _ in range(nb)]
            ok = False
        for b in range(nb):
            c = lb[b]
            c = [c[q] if c[q] > face.minlight else face.minlight for q in range(3)]
            if b == 0 and ok:
                red_.append(c[0])
                green_.append(c[1])
                blue_.append(c[2])
            lum[b, j] = c
    med = np.zeros(3, np.float32)
    if red_:
        med = np.array([median(red_), median(green_), median(blue_)], np.float32)
    return {"luxels": lum, "median": med}


def median(values) -> np.float32:
    """Median (FinalLightFace.cs:466-472): element count / 2 of the ascending
    order."""
    v = sorted(values)
    return v[len(v) >> 1]


def layout(faces: Sequence[Face], styles=None):
    """LightmapOffsets.Compute (Rad/Light/LightmapOffsets.cs:60-123): each lit
    face reserves an average-colour word per style, then 4 bytes per luxel
    per normal per style; its light offset is just past the average words.
    ``styles`` (F, 4) or None (style 0 alone).  Returns (offsets, -1 for a
    face not lit; total bytes)."""
    offs, size = [], 0
    for f, face in enumerate(faces):
        ns = 1 if styles is None else next((k for k in range(4) if styles[f][k] == 255), 4)
        if not face.lit or ns == 0:
            offs.append(-1)
            continue
        size += 4 * ns
        offs.append(size)
        size += 4 * face.nb * face.w * face.h * ns
    return offs, size


def lump_bytes(faces: Sequence[Face], result: Sequence[Optional[dict]]) -> bytes:
    """The stage's lump: for each lit face in order, the average colour words
    in reverse style order, then its luxels style by style, bump map by bump
    map (the face's light offset is just past the average words; style k's
    median sits 4 (k + 1) bytes before it: FinalLightFace.cs:270, 279)."""
    parts = []
    for face, r in zip(faces, result):
        if r is None:
            continue
        med = r["median"].reshape(-1, 3)
        parts.append(encode_rgbe(med[::-1].copy()).tobytes())
        parts.append(encode_rgbe(r["luxels"].reshape(-1, 3)).tobytes())
    return b"".join(parts)


# --- the C statement (vradcpu_final.c), for maps of realistic size ---

FACE_DTYPE = np.dtype([
    ("origin", "<f4", 3), ("w2l_s", "<f4", 3), ("w2l_t", "<f4", 3), ("l2w_s", "<f4", 3),
    ("l2w_t", "<f4", 3), ("min_s", "<i4"), ("min_t", "<i4"), ("width", "<i4"), ("height", "<i4"),
    ("minlight", "<f4"), ("lit", "<i4"),
])


def final_light_c(faces: np.ndarray, luxel_pos, sample_bounds, nbr_ptr, nbr, face_sample_ptr,
                  face_normal_count, sample_pos, light, patches: Optional[dict] = None,
                  patch_light=None, red_errors: bool = False, threads: int = 1, face_styles=None):
    """vc_final_light over whole maps: faces (FACE_DTYPE), the luxels, the
    samples' bounds, neighbour CSR, the gather's faces, sample points and
    light ((4, S, 4), or (slots, 4, S, 4) with ``face_styles`` (F, 4) style
    numbers per slot), and for a bounce the patches ({'origin', 'face',
    'child1', 'winding_ptr', 'winding'}) with their TotalLight per normal
    (npatch, 4, 3).  Returns (lump bytes as uint8, offsets)."""
    import ctypes

    from . import lib

    L = lib()
    P = ctypes.c_void_p

    class Desc(ctypes.Structure):
        _fields_ = [("nfaces", ctypes.c_int32), ("faces", P), ("luxel_pos", P), ("sample_bounds", P),
                    ("nbr_ptr", P), ("nbr", P), ("npatch", ctypes.c_int32), ("patch_origin", P),
                    ("patch_face", P), ("patch_child1", P), ("winding_ptr", P), ("winding", P),
                    ("face_sample_ptr", P), ("face_normal_count", P), ("sample_pos", P),
                    ("nsamples", ctypes.c_int32), ("light", P), ("patch_light", P),
                    ("red_errors", ctypes.c_int32), ("face_styles", P), ("light_slots", ctypes.c_int32)]

    keep = []

    def c(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data if a.size else None

    fa = np.ascontiguousarray(faces, FACE_DTYPE)
    keep.append(fa)
    d = Desc()
    d.nfaces = len(fa)
    d.faces = fa.ctypes.data
    d.luxel_pos = c(luxel_pos, np.float32)
    d.sample_bounds = c(sample_bounds, np.float32)
    d.nbr_ptr = c(nbr_ptr, np.int32)
    d.nbr = c(nbr, np.int32)
    d.face_sample_ptr = c(face_sample_ptr, np.int32)
    d.face_normal_count = c(face_normal_count, np.int32)
    d.sample_pos = c(sample_pos, np.float32)
    d.nsamples = int(np.asarray(sample_pos).shape[0])
    d.light = c(light, np.float32)
    d.red_errors = 1 if red_errors else 0
    d.light_slots = 1
    if face_styles is not None:
        d.face_styles = c(face_styles, np.uint8)
        d.light_slots = int(np.asarray(light).reshape(-1, 4, int(np.asarray(sample_pos).shape[0]), 4).shape[0])
    if patches is not None and patch_light is not None:
        d.npatch = len(patches["origin"])
        d.patch_origin = c(patches["origin"], np.float32)
        d.patch_face = c(patches["face"], np.int32)
        d.patch_child1 = c(patches["child1"], np.int32)
        d.winding_ptr = c(patches["winding_ptr"], np.int32)
        d.winding = c(patches["winding"], np.float32)
        d.patch_light = c(patch_light, np.float32)
    L.vc_final_layout.restype = ctypes.c_int64
    L.vc_final_layout.argtypes = [P, P]
    L.vc_final_light.restype = ctypes.c_int
    L.vc_final_light.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32]
    offs = np.zeros(max(len(fa), 1), np.int32)
    size = L.vc_final_layout(ctypes.byref(d), offs.ctypes.data)
    out = np.zeros(max(size, 1), np.uint8)
    rc = L.vc_final_light(ctypes.byref(d), out.ctypes.data, out.shape[0], int(threads))
    if rc != 0:
        raise RuntimeError("vc_final_light failed")
    return out[:size], offs[:len(fa)]
