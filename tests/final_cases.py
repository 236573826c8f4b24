"""Final-light test cases and their translation into the oracle's inputs.

A random case puts faces at arbitrary orientations, luxel scales and mins,
with samples scattered over and past their grids (bounds of random size),
luxels jittered and some pushed off the grid, random neighbour lists, bumped
faces among flat ones, one face lit without samples and one not lit, and
patches of three to five winding points, some of them parents."""
from __future__ import annotations

import numpy as np

from oracle import final as fo
from hl2sdk_tools_sharp_amd.final import FACE_DTYPE, FinalFaces, FinalPatches
from hl2sdk_tools_sharp_amd.light import FaceSet, bump_basis, make_lights, EMIT_POINT
from hl2sdk_tools_sharp_amd.scenes import box_triangles
from hl2sdk_tools_sharp_amd.tracer import ShadowCasterSet


def oracle_faces(ff: FinalFaces, fset: FaceSet):
    out = []
    for f, x in enumerate(ff.faces):
        out.append(fo.Face(x["origin"], x["w2l_s"], x["w2l_t"], x["l2w_s"], x["l2w_t"], x["min_s"], x["min_t"],
                           x["width"], x["height"], bumped=fset.face_normal_count[f] == 4,
                           minlight=x["minlight"], lit=bool(x["lit"])))
    return out


def oracle_inputs(ff: FinalFaces, fset: FaceSet, light: np.ndarray, fp: FinalPatches = None,
                  patch_total: np.ndarray = None, face_styles=None):
    """(faces, luxels, samples, neighbours, patches) for oracle.final.final_light;
    ``light`` is the gather's (4, S, 4) result the device reads, or with
    ``face_styles`` (F, 4) its (4 slots, 4, S, 4) light per style slot."""
    faces = oracle_faces(ff, fset)
    n = ff.faces["width"].astype(np.int64) * ff.faces["height"]
    lb = np.concatenate([[0], np.cumsum(n)])
    if ff.luxel_pos is not None:
        luxels = [ff.luxel_pos[lb[f]:lb[f + 1]] for f in range(len(faces))]
    else:  # BuildFaceLuxels: LuxelToWorld(s, t) of every grid point
        luxels = [np.array([face.luxel_to_world(s, t) for t in range(face.h) for s in range(face.w)],
                           np.float32).reshape(-1, 3) for face in faces]
    samples = []
    for f, face in enumerate(faces):
        s0, s1 = fset.face_sample_ptr[f], fset.face_sample_ptr[f + 1]
        if face_styles is None:
            lt = light[:face.nb, s0:s1, :3]
        else:
            lt = [light[k][:face.nb, s0:s1, :3] for k in range(4)]
        samples.append({"pos": fset.sample_pos[s0:s1], "bounds": ff.sample_bounds[s0:s1], "light": lt})
    nbrs = [list(ff.nbr[ff.nbr_ptr[f]:ff.nbr_ptr[f + 1]]) for f in range(len(faces))]
    patches = None
    if fp is not None:
        patches = {"origin": fp.origin, "face": fp.face, "leaf": fp.child1 < 0,
                   "winding": [fp.winding[fp.winding_ptr[p]:fp.winding_ptr[p + 1]] for p in range(fp.npatch)],
                   "total": patch_total}
    return faces, luxels, samples, nbrs, patches


def face_bytes(faces, result, offsets):
    """{face: (light offset, expected bytes from offset - 4)} for the faces the
    oracle ran."""
    out = {}
    for f, r in enumerate(result):
        if r is None:
            continue
        nst = 1 if r["median"].ndim == 1 else r["median"].shape[0]
        out[f] = (int(offsets[f]), fo.lump_bytes([faces[f]], [r]), nst)
    return out


def random_styles(nfaces: int, seed: int) -> np.ndarray:
    """(F, 4) style numbers: slot 0 style 0, then up to three of a few
    styles in a random order per face."""
    rng = np.random.default_rng(seed)
    out = np.full((nfaces, 4), 255, np.uint8)
    out[:, 0] = 0
    for f in range(nfaces):
        k = int(rng.integers(0, 4))
        out[f, 1:1 + k] = rng.choice([3, 5, 9, 33], size=k, replace=False)
    return out


def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    a = np.cross(n, [0.3, 0.5, 0.8])
    a /= np.linalg.norm(a)
    b = np.cross(n, a)
    return n, a, b


def random_case(seed: int, nfaces: int = 14, npatch: int = 40):
    """(casters, lights, FaceSet, FinalFaces, FinalPatches) for one seed."""
    rng = np.random.default_rng(seed)
    fa = np.zeros(nfaces, FACE_DTYPE)
    ptr = [0]
    pos, bounds, luxels, normals, ncount = [], [], [], [], []
    for f in range(nfaces):
        n, a, b = _frame(rng)
        sc = float(rng.choice([4.0, 8.0, 16.0]))
        w, h = int(rng.integers(1, 13)), int(rng.integers(1, 13))
        if f == 0:
            w, h = 21, 9  # several waves of cells
        org = rng.uniform(-300, 300, 3)
        ms, mt = int(rng.integers(-6, 7)), int(rng.integers(-6, 7))
        x = fa[f]
        x["origin"], x["w2l_s"], x["w2l_t"] = org, a / sc, b / sc
        x["l2w_s"], x["l2w_t"] = a * sc, b * sc
        x["min_s"], x["min_t"], x["width"], x["height"] = ms, mt, w, h
        x["minlight"] = 0.0 if f % 4 else 3.0
        x["lit"] = 0 if f == 5 else 1

        def world(s, t):
            return org + a * sc * (s + ms) + b * sc * (t + mt)

        k = 0 if f == 6 else int(rng.integers(w * h // 2, w * h + 6))
        st = np.stack([rng.uniform(-0.8, w - 0.2, k), rng.uniform(-0.8, h - 0.2, k)], 1)
        for s, t in st:
            pos.append(world(s, t) + n * rng.uniform(0, 1))
            lo = rng.uniform(0.05, 0.95, 2)
            hi = rng.uniform(0.05, 0.95, 2)
            bounds.append((s - lo[0], t - lo[1], s + hi[0], t + hi[1]))
        ptr.append(ptr[-1] + k)
        for t in range(h):
            for s in range(w):
                js, jt = s + rng.uniform(-0.3, 0.3), t + rng.uniform(-0.3, 0.3)
                if rng.random() < 0.03:
                    js += w + 1  # off the grid
                luxels.append(world(js, jt))
        bumped = f % 3 == 1
        ncount.append(4 if bumped else 1)
        nm = np.zeros((4, 3))
        nm[0] = n
        if bumped:
            nm[1:] = bump_basis(n.astype(np.float32))
        normals.append(nm)
    nbr_ptr, nbr = [0], []
    for f in range(nfaces):
        k = int(rng.integers(0, 5))
        others = [g for g in range(nfaces) if g != f]
        nbr += sorted(rng.choice(others, size=k, replace=False).tolist())
        nbr_ptr.append(len(nbr))
    fset = FaceSet(np.array(ptr, np.int32), np.array(ncount, np.int32), np.array(normals, np.float32),
                   np.array(pos, np.float32), np.zeros(len(pos), np.int32))
    ff = FinalFaces(fa, np.array(luxels, np.float32), np.array(bounds, np.float32),
                    np.array(nbr_ptr, np.int32), np.array(nbr, np.int32))
    # patches: a few winding points round a point near the face
    porg, pface, child, wptr, wpts = [], [], [], [0], []
    for p in range(npatch):
        f = int(rng.integers(0, nfaces))
        x = fa[f]
        s, t = rng.uniform(-1, x["width"] + 1), rng.uniform(-1, x["height"] + 1)
        c = x["origin"] + x["l2w_s"] * (s + x["min_s"]) + x["l2w_t"] * (t + x["min_t"])
        k = int(rng.integers(3, 6))
        ext = rng.uniform(0.2, 3.0) * np.linalg.norm(x["l2w_s"])
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        pts = [c + ext * (np.cos(q) * x["l2w_s"] + np.sin(q) * x["l2w_t"]) / np.linalg.norm(x["l2w_s"])
               for q in ang]
        porg.append(np.mean(pts, 0))
        pface.append(f)
        child.append(p + 1 if (p % 7 == 3 and p + 1 < npatch) else -1)
        wpts += pts
        wptr.append(len(wpts))
    fp = FinalPatches(np.array(porg, np.float32), np.array(pface, np.int32), np.array(child, np.int32),
                      np.array(wptr, np.int32), np.array(wpts, np.float32))
    # a box for the scene and two point lights
    v9, ids = box_triangles(np.array([[900.0, 900, 900]]), np.array([[950.0, 950, 950]]))
    casters = ShadowCasterSet(v9, ids)
    L = make_lights(2)
    for i, o in enumerate(([0.0, 0.0, 400.0], [-200.0, 300.0, -350.0])):
        L[i]["type"] = EMIT_POINT
        L[i]["origin"] = o
        L[i]["intensity"] = (90000.0, 60000.0, 30000.0)
        L[i]["intensity_hdr"] = (180000.0, 120000.0, 60000.0)
        L[i]["quadratic_attn"] = 1.0
    return casters, L, fset, ff, fp
