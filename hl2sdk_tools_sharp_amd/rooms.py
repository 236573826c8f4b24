"""Seeded room grids: the patch system vrad's bounce runs on.

The reference gets faces, patches, clusters and a PVS from vbsp and vvis;
here they are generated, with the reference's subdivision rules and the
reference's patch fields, so that the transfer build and the bounce see the
same structure they see on a compiled map.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass

import numpy as np

from .tracer import TRACE_ID_OPAQUE, TRACE_ID_SKY, ShadowCasterSet


@dataclass
class RoomGrid:
    """A grid of box rooms joined by doorways, with everything the bounce
    needs: faces, the patch tree of each face, clusters (one per room), a PVS
    row per cluster and shadow casters.

    Patch fields follow the reference's ``Patch`` (Rad/Light/Patch.cs):
    origin, plane normal, plane distance, area, ``child1``/``child2`` (-1 for
    a leaf), the next root patch of the same face (the FaceParents chain),
    face, cluster and a four-corner winding wound as vrad's are
    ((c1-c0) x (c2-c0) points away from the face normal, the sense
    FormFactors.PolyToDiff expects).  Children have higher indices than their
    parents, as in ``PatchSet``."""

    casters: ShadowCasterSet
    room_lo: np.ndarray  # (C, 3) interior bounds
    room_hi: np.ndarray
    face_lo: np.ndarray  # (F, 3)
    face_hi: np.ndarray
    face_normal: np.ndarray  # (F, 3)
    face_cluster: np.ndarray  # (F,)
    face_sky: np.ndarray  # (F,) uint8
    face_refl: np.ndarray  # (F, 3)
    face_parent: np.ndarray  # (F,) first root patch of the face
    origin: np.ndarray  # (P, 3) float32
    normal: np.ndarray
    plane_dist: np.ndarray
    area: np.ndarray
    child1: np.ndarray
    child2: np.ndarray
    parent: np.ndarray
    next_parent: np.ndarray
    face: np.ndarray
    cluster: np.ndarray
    winding: np.ndarray  # (P, 4, 3)
    row_faces_ptr: np.ndarray  # (C + 1,)
    row_faces: np.ndarray
    receivers: np.ndarray  # leaf patches, cluster by cluster
    lights: np.ndarray  # (L, 3) point lights
    light_intensity: np.ndarray  # (L,)
    to_sun: np.ndarray  # (3,) unit, towards the sun

    @property
    def npatch(self) -> int:
        return self.origin.shape[0]

    @property
    def nface(self) -> int:
        return self.face_lo.shape[0]

    @property
    def ncluster(self) -> int:
        return self.room_lo.shape[0]

    @property
    def leaf(self) -> np.ndarray:
        return self.child1 < 0

    @property
    def sky(self) -> np.ndarray:
        return self.face_sky[self.face]

    @property
    def refl(self) -> np.ndarray:
        return self.face_refl[self.face]

    def levels(self) -> np.ndarray:
        """Height of each patch above its deepest leaf (leaves are 0)."""
        h = np.zeros(self.npatch, np.int32)
        for i in range(self.npatch - 1, -1, -1):
            c1 = self.child1[i]
            if c1 >= 0:
                h[i] = 1 + max(h[c1], h[self.child2[i]])
        return h

    def pull_up(self, values: np.ndarray) -> np.ndarray:
        """Parents from their children, area weighted, children first
        (Radiosity.CollectLight's parent rule, Rad/Bounce/Radiosity.cs:268-316)."""
        v = np.array(values, dtype=np.float32, copy=True)
        a = self.area
        for i in range(self.npatch - 1, -1, -1):
            c1 = self.child1[i]
            if c1 >= 0:
                c2 = self.child2[i]
                s1 = np.float32(a[c1] / (a[c1] + a[c2]))
                s2 = np.float32(a[c2] / (a[c1] + a[c2]))
                v[i] = (v[c1] * s1) + (v[c2] * s2)
        return v


def _rect_winding(lo, hi, ax, sign):
    u, v = (ax + 1) % 3, (ax + 2) % 3
    c = np.zeros((4, 3), dtype=np.float64)
    for q, (su, sv) in enumerate([(0, 0), (1, 0), (1, 1), (0, 1)]):
        c[q, ax] = lo[ax]
        c[q, u] = hi[u] if su else lo[u]
        c[q, v] = hi[v] if sv else lo[v]
    # this order winds along +axis; vrad's windings wind against the normal
    if sign > 0:
        c = c[[0, 3, 2, 1]]
    return c


def room_grid(nx: int = 3, ny: int = 3, size=(512.0, 512.0, 256.0), wall: float = 8.0,
              door: float = 128.0, patch: float = 128.0, minchop: float = 64.0,
              nlights: int = 8, sky_ceiling: bool = False, seed: int = 1,
              tile: float = 0.0) -> RoomGrid:
    """A grid of ``nx`` x ``ny`` rooms of interior ``size``, walls ``wall`` thick,
    each shared wall pierced by a ``door`` x ``door`` doorway at floor level in
    its middle.  Faces are the rooms' inner surfaces (a wall with a doorway is
    split into three rectangles around it) and the doorway jambs, which belong
    to the room on their lower side.  Patches are subdivided by the
    reference's rules (PatchSubdivider.SubdividePatch and CreateChildPatch,
    Rad/Light/PatchSubdivider.cs) with ``patch`` as the chop and ``minchop``
    as its floor, both in world units.  The PVS row of a room is the room and
    its doorway neighbours.

    ``tile`` > 0 cuts every face into square tiles of that size first, as
    vbsp leaves a map whose faces are small: each tile is a face with a
    single patch (no subdivision, so no parent patches), and every source a
    receiver sees is then a separate candidate."""
    rng = np.random.default_rng(seed)
    W, D, H = (float(x) for x in size)
    nrooms = nx * ny
    room_lo = np.zeros((nrooms, 3))
    room_hi = np.zeros((nrooms, 3))
    for j in range(ny):
        for i in range(nx):
            r = j * nx + i
            room_lo[r] = (i * (W + wall), j * (D + wall), 0.0)
            room_hi[r] = room_lo[r] + (W, D, H)

    faces = []  # (lo, hi, axis, sign, cluster, sky)

    def add_face(lo, hi, ax, sign, cluster, sky=False):
        lo, hi = np.array(lo, float), np.array(hi, float)
        ext = np.delete(hi - lo, ax)
        if np.all(ext > 0):  # a doorway as tall as the room leaves no lintel face
            faces.append((lo, hi, ax, sign, cluster, sky))

    half = door / 2.0
    for j in range(ny):
        for i in range(nx):
            r = j * nx + i
            lo, hi = room_lo[r], room_hi[r]
            add_face((lo[0], lo[1], lo[2]), (hi[0], hi[1], lo[2]), 2, +1, r)  # floor
            add_face((lo[0], lo[1], hi[2]), (hi[0], hi[1], hi[2]), 2, -1, r, sky_ceiling)  # ceiling
            for ax, door_lo, door_hi in ((0, i > 0, i < nx - 1), (1, j > 0, j < ny - 1)):
                u = 1 - ax  # the wall's horizontal in-plane axis
                for side, has_door in ((0, door_lo), (1, door_hi)):
                    c = hi[ax] if side else lo[ax]
                    sign = -1 if side else +1
                    wl = lo.copy()
                    wh = hi.copy()
                    wl[ax] = wh[ax] = c
                    if not has_door:
                        add_face(wl, wh, ax, sign, r)
                        continue
                    mid = 0.5 * (lo[u] + hi[u])
                    a, b = wl.copy(), wh.copy()
                    b[u] = mid - half
                    add_face(a, b, ax, sign, r)
                    a, b = wl.copy(), wh.copy()
                    a[u] = mid + half
                    add_face(a, b, ax, sign, r)
                    a, b = wl.copy(), wh.copy()
                    a[u], b[u], a[2] = mid - half, mid + half, door
                    add_face(a, b, ax, sign, r)
            for ax, has in ((0, i < nx - 1), (1, j < ny - 1)):
                if not has:
                    continue
                u = 1 - ax
                mid = 0.5 * (lo[u] + hi[u])
                x0, x1 = hi[ax], hi[ax] + wall
                for s_u, sign in ((mid - half, +1), (mid + half, -1)):  # side jambs
                    a, b = np.zeros(3), np.zeros(3)
                    a[ax], b[ax] = x0, x1
                    a[u] = b[u] = s_u
                    a[2], b[2] = 0.0, door
                    add_face(a, b, u, sign, r)
                a, b = np.zeros(3), np.zeros(3)
                a[ax], b[ax] = x0, x1
                a[u], b[u] = mid - half, mid + half
                a[2] = b[2] = door
                add_face(a, b, 2, -1, r)  # lintel
                a, b = a.copy(), b.copy()
                a[2] = b[2] = 0.0
                add_face(a, b, 2, +1, r)  # threshold

    if tile > 0:
        cut = []
        for lo, hi, ax, sign, cl, sky in faces:
            u, v = (ax + 1) % 3, (ax + 2) % 3
            nu = max(1, int(np.ceil((hi[u] - lo[u]) / tile - 1e-9)))
            nv = max(1, int(np.ceil((hi[v] - lo[v]) / tile - 1e-9)))
            for a in range(nu):
                for b in range(nv):
                    tl, th = lo.copy(), hi.copy()
                    tl[u] = lo[u] + a * tile
                    th[u] = min(hi[u], lo[u] + (a + 1) * tile)
                    tl[v] = lo[v] + b * tile
                    th[v] = min(hi[v], lo[v] + (b + 1) * tile)
                    cut.append((tl, th, ax, sign, cl, sky))
        faces = cut

    F = len(faces)
    face_lo = np.array([f[0] for f in faces], np.float32)
    face_hi = np.array([f[1] for f in faces], np.float32)
    face_normal = np.zeros((F, 3), np.float32)
    for k, f in enumerate(faces):
        face_normal[k, f[2]] = f[3]
    face_cluster = np.array([f[4] for f in faces], np.int32)
    face_sky = np.array([1 if f[5] else 0 for f in faces], np.uint8)
    face_refl = rng.uniform(0.3, 0.7, size=(F, 3)).astype(np.float32)

    v9 = np.zeros((F, 2, 9), np.float32)
    ids = np.zeros((F, 2), np.int32)
    for k, f in enumerate(faces):
        w = _rect_winding(f[0], f[1], f[2], f[3])
        v9[k, 0] = np.concatenate([w[0], w[1], w[2]])
        v9[k, 1] = np.concatenate([w[0], w[2], w[3]])
        ids[k] = (TRACE_ID_SKY if f[5] else TRACE_ID_OPAQUE) | k
    casters = ShadowCasterSet(v9.reshape(-1, 9), ids.reshape(-1))

    # Patches: one root per face, then each root's subtree in the order
    # SubdividePatch appends it (both children, then child1's subtree, then child2's).
    P_lo, P_hi, P_area, P_face, P_chop = [], [], [], [], []
    P_child1, P_child2, P_parent = [], [], []

    def new_patch(lo, hi, face_k, chop, parent):
        ax = faces[face_k][2]
        u, v = (ax + 1) % 3, (ax + 2) % 3
        P_lo.append(lo.copy())
        P_hi.append(hi.copy())
        P_area.append(np.float32((hi[u] - lo[u]) * (hi[v] - lo[v])))
        P_face.append(face_k)
        P_chop.append(chop)
        P_child1.append(-1)
        P_child2.append(-1)
        P_parent.append(parent)
        return len(P_lo) - 1

    for k, f in enumerate(faces):
        new_patch(f[0], f[1], k, patch, -1)

    def subdivide(p):
        f = faces[P_face[p]]
        if f[5]:  # sky patches are never split
            return
        total = P_hi[p] - P_lo[p]
        chop = P_chop[p]
        widest_axis, widest, split = -1, -1.0, False
        for i in range(3):
            t = total[i]
            if t > widest:
                widest_axis, widest = i, t
            if t >= chop and t >= minchop:
                split = True
        if not split and widest_axis != -1:
            wv = total[widest_axis]
            if (wv > total[(widest_axis + 1) % 3] * 2 and wv > total[(widest_axis + 2) % 3] * 2
                    and chop > minchop):
                split = True
                chop = max(minchop, chop / 2)
                P_chop[p] = chop
        if not split:
            return
        ax = widest_axis
        dist = (P_lo[p][ax] + P_hi[p][ax]) * 0.5
        lo1, hi1 = P_lo[p].copy(), P_hi[p].copy()
        lo1[ax] = dist  # child1: the front of the split plane
        lo2, hi2 = P_lo[p].copy(), P_hi[p].copy()
        hi2[ax] = dist
        kids = []
        for lo, hi in ((lo1, hi1), (lo2, hi2)):
            cchop = chop
            t = hi - lo
            if cchop > minchop and all(t[i] < cchop for i in range(3)):
                fl, fh = f[0], f[1]
                for i in range(3):
                    if (fh[i] == hi[i] or fl[i] == lo[i]) and t[i] > minchop:
                        cchop = max(minchop, cchop / 2)
                        break
            kids.append(new_patch(lo, hi, P_face[p], cchop, p))
        P_child1[p], P_child2[p] = kids
        subdivide(kids[0])
        subdivide(kids[1])

    if tile <= 0:
        old = sys.getrecursionlimit()
        sys.setrecursionlimit(max(old, 10000))
        try:
            for k in range(F):
                subdivide(k)
        finally:
            sys.setrecursionlimit(old)

    P = len(P_lo)
    lo_a = np.array(P_lo)
    hi_a = np.array(P_hi)
    origin = (0.5 * (lo_a + hi_a)).astype(np.float32)
    face_arr = np.array(P_face, np.int32)
    normal = face_normal[face_arr].copy()
    fplane = np.array([np.float32(face_normal[k] @ faces[k][0]) for k in range(F)], np.float32)
    plane_dist = fplane[face_arr]
    winding = np.zeros((P, 4, 3), np.float32)
    for p in range(P):
        f = faces[P_face[p]]
        winding[p] = _rect_winding(P_lo[p], P_hi[p], f[2], f[3])
    child1 = np.array(P_child1, np.int32)
    child2 = np.array(P_child2, np.int32)
    face_parent = np.full(F, -1, np.int32)
    next_parent = np.full(P, -1, np.int32)
    for p in range(F):  # root patches chained per face, prepended in index order
        next_parent[p] = face_parent[P_face[p]]
        face_parent[P_face[p]] = p
    cluster = face_cluster[face_arr].copy()

    row_ptr, row = [0], []
    for c in range(nrooms):
        i, j = c % nx, c // nx
        neigh = ((c - 1, i > 0), (c + 1, i < nx - 1), (c - nx, j > 0), (c + nx, j < ny - 1))
        for cc in sorted({c} | {nb for nb, ok in neigh if ok}):
            row.extend(np.nonzero(face_cluster == cc)[0].tolist())
        row_ptr.append(len(row))
    leaf = child1 < 0
    receivers = np.concatenate(
        [np.nonzero(leaf & (cluster == c))[0] for c in range(nrooms)]
    ).astype(np.int32)

    lroom = rng.integers(0, nrooms, size=nlights)
    lpos = room_lo[lroom] + 16.0 + rng.uniform(0, 1, size=(nlights, 3)) * (
        room_hi[lroom] - room_lo[lroom] - 32.0)
    intensity = rng.uniform(50.0, 300.0, size=nlights).astype(np.float32)
    phi = rng.uniform(0, 2 * np.pi)
    z = rng.uniform(0, 1)
    rxy = np.sqrt(max(0.0, 1 - z * z))
    to_sun = np.array([rxy * np.cos(phi), rxy * np.sin(phi), z], np.float32)

    return RoomGrid(
        casters=casters, room_lo=room_lo.astype(np.float32), room_hi=room_hi.astype(np.float32),
        face_lo=face_lo, face_hi=face_hi, face_normal=face_normal, face_cluster=face_cluster,
        face_sky=face_sky, face_refl=face_refl, face_parent=face_parent, origin=origin,
        normal=normal, plane_dist=plane_dist, area=np.array(P_area, np.float32), child1=child1,
        child2=child2, parent=np.array(P_parent, np.int32), next_parent=next_parent, face=face_arr,
        cluster=cluster, winding=winding, row_faces_ptr=np.array(row_ptr, np.int32),
        row_faces=np.array(row, np.int32), receivers=receivers,
        lights=lpos.astype(np.float32), light_intensity=intensity, to_sun=to_sun,
    )


def bounce_grid(nx: int = 12, ny: int = 12, seed: int = 3) -> RoomGrid:
    """The bounce config: a 12 x 12 room grid of 64-unit patches, faces cut to
    the patch size (so every patch is its own source), 1344 x 1344 x 384
    rooms joined by 256 x 256 doorways, so the map has 200,000 +- 2 % patches
    (198,528 at 12 x 12) and its receivers see about 1,500 sources each; two
    point lights per room."""
    return room_grid(nx, ny, size=BOUNCE_ROOM, door=BOUNCE_DOOR, patch=64.0, minchop=64.0, tile=64.0,
                     nlights=2 * nx * ny, seed=seed)


def sample_points(g: RoomGrid) -> np.ndarray:
    """Each patch's sample point: its origin one unit along its normal, as
    FaceLightJob offsets samples (Rad/Light/FaceLightJob.cs:244-251)."""
    return (g.origin + g.normal).astype(np.float32)


def direct_light(g: RoomGrid, blocked: np.ndarray) -> np.ndarray:
    """Point-light irradiance at every patch from its sample point: for each
    light not blocked, intensity * max(0, n . l) / max(d, 1)^2 (quadratic
    falloff, LightFalloff with c = l = 0, q = 1), parents pulled up.
    ``blocked`` is one bool per (patch, light), patch-major."""
    pts = sample_points(g).astype(np.float64)
    L = g.lights.astype(np.float64)
    blocked = np.asarray(blocked, bool).reshape(g.npatch, len(L))
    e = np.zeros(g.npatch, np.float64)
    # only the unblocked (light, patch) pairs, light-major, so each patch
    # sums its lights in ascending order
    lk, pk = np.nonzero(~blocked.T)
    d = L[lk] - pts[pk]
    dist = np.sqrt((d * d).sum(1))
    dot = np.maximum(0.0, (d * g.normal[pk]).sum(1) / np.maximum(dist, 1e-9))
    np.add.at(e, pk, g.light_intensity[lk] * dot / np.maximum(dist, 1.0) ** 2)
    e = np.where(g.leaf & ~g.sky.astype(bool), e, 0.0)
    e3 = np.repeat(e[:, None], 3, axis=1).astype(np.float32)
    return g.pull_up(e3)


BOUNCE_ROOM = (1344.0, 1344.0, 384.0)
BOUNCE_DOOR = 256.0


def room_faces(g: RoomGrid, luxel: float = 16.0, bumped_every: int = 0):
    """The grid's faces as lightmaps: one sample at the centre of every
    ``luxel`` x ``luxel`` cell of each face (its area the cell's), the face's
    cluster for each sample; every ``bumped_every``-th face (0 = none) carries
    the three bump normals as well.  Sky faces carry no lightmap (vrad lights
    no sky surface) and so have no samples."""
    from .light import FaceSet, bump_basis

    F = g.nface
    fi = np.arange(F)
    lo = g.face_lo.astype(np.float64)
    hi = g.face_hi.astype(np.float64)
    n = g.face_normal
    ax = np.argmax(np.abs(n), axis=1)
    u, v = (ax + 1) % 3, (ax + 2) % 3
    ext_u = hi[fi, u] - lo[fi, u]
    ext_v = hi[fi, v] - lo[fi, v]
    nu = np.maximum(1, np.round(ext_u / luxel).astype(np.int64))
    nv = np.maximum(1, np.round(ext_v / luxel).astype(np.int64))
    cnt = np.where(g.face_sky.astype(bool), 0, nu * nv)
    ptr = np.zeros(F + 1, np.int64)
    np.cumsum(cnt, out=ptr[1:])
    S = int(ptr[-1])
    fs = np.repeat(fi, cnt)  # each sample's face, in face order
    k = np.arange(S) - ptr[fs]
    i, j = k // nv[fs], k % nv[fs]  # cell (i, j), j fastest
    pos = np.zeros((S, 3))
    rows = np.arange(S)
    pos[rows, ax[fs]] = lo[fs, ax[fs]]
    pos[rows, u[fs]] = lo[fs, u[fs]] + (i + 0.5) * ext_u[fs] / nu[fs]
    pos[rows, v[fs]] = lo[fs, v[fs]] + (j + 0.5) * ext_v[fs] / nv[fs]
    area = (ext_u / nu * ext_v / nv)[fs]
    nc = np.ones(F, np.int32)
    normals = np.zeros((F, 4, 3), np.float32)
    normals[:, 0] = n
    if bumped_every:
        for f in range(0, F, bumped_every):
            if not g.face_sky[f]:
                nc[f] = 4
                normals[f, 1:] = bump_basis(n[f])
    return FaceSet(ptr.astype(np.int32), nc, normals, pos.astype(np.float32),
                   g.face_cluster[fs].astype(np.int32), area.astype(np.float32))


def patch_bumps(g: RoomGrid, fset):
    """Per patch, whether it gathers per bump normal (Patch.NeedsBumpmap: its
    face carries bump normals) and its three bump normals.  The reference
    builds each patch's basis from the phong normal at its origin
    (Radiosity.PrepareNormals, Rad/Bounce/Radiosity.cs:322-356); the grid's
    faces are flat and unsmoothed, so every patch takes its face's basis."""
    bumped = (np.asarray(fset.face_normal_count)[g.face] == 4).astype(np.uint8)
    normals = np.where(bumped[:, None, None] == 1, np.asarray(fset.face_normals, np.float32)[g.face, 1:4], 0)
    return bumped, np.ascontiguousarray(normals, np.float32)


def room_lights(g: RoomGrid, sun: bool = True, sky_ambient: bool = False, spot_every: int = 0,
                surface_every: int = 0):
    """The grid's point lights (quadratic falloff 1/d^2, white, intensity from
    the generator, HDR = 2x LDR), each visible from its room's PVS row; plus a
    sun along -to_sun and optionally a sky-ambient light.  Returns (lights,
    light_pvs rows)."""
    from .light import EMIT_POINT, EMIT_SKYAMBIENT, EMIT_SKYLIGHT, EMIT_SPOTLIGHT, EMIT_SURFACE, make_lights

    nl = len(g.lights) + (1 if sun else 0) + (1 if sky_ambient else 0)
    L = make_lights(nl)
    C = g.ncluster
    pvs_bytes = (C + 7) // 8
    pvs = np.zeros((nl, pvs_bytes), np.uint8)
    for k in range(len(g.lights)):
        p = g.lights[k]
        inside = np.all((p >= g.room_lo) & (p <= g.room_hi), axis=1)
        room = int(np.argmax(inside))
        L[k]["type"] = EMIT_POINT
        L[k]["origin"] = p
        L[k]["intensity"] = g.light_intensity[k]
        L[k]["intensity_hdr"] = 2.0 * g.light_intensity[k]
        L[k]["quadratic_attn"] = 1.0
        if spot_every and k % spot_every == 1:
            L[k]["type"] = EMIT_SPOTLIGHT
            L[k]["normal"] = (0.0, 0.0, -1.0)
            L[k]["stopdot"] = np.float32(np.cos(np.radians(30.0)))
            L[k]["stopdot2"] = np.float32(np.cos(np.radians(45.0)))
            L[k]["exponent"] = 2.0
            L[k]["start_fade"], L[k]["end_fade"] = 100.0, 900.0
        elif surface_every and k % surface_every == 2:
            L[k]["type"] = EMIT_SURFACE
            L[k]["normal"] = (0.0, 0.0, -1.0)
            L[k]["quadratic_attn"] = 0.0
        faces = g.row_faces[g.row_faces_ptr[room]:g.row_faces_ptr[room + 1]]
        for c in np.unique(g.face_cluster[faces]):
            pvs[k, c >> 3] |= np.uint8(1 << (c & 7))
    k = len(g.lights)
    if sun:
        L[k]["type"] = EMIT_SKYLIGHT
        L[k]["normal"] = -g.to_sun
        L[k]["intensity"] = (255.0, 240.0, 200.0)
        L[k]["intensity_hdr"] = (510.0, 480.0, 400.0)
        pvs[k] = 0xFF
        k += 1
    if sky_ambient:
        L[k]["type"] = EMIT_SKYAMBIENT
        L[k]["intensity"] = (40.0, 60.0, 90.0)
        L[k]["intensity_hdr"] = (80.0, 120.0, 180.0)
        pvs[k] = 0xFF
    return L, pvs


def face_neighbours(g: RoomGrid):
    """FaceNeighbours for the grid's faces (Rad/Light/FaceNeighbours.cs): the
    faces that share a corner with a face and smooth with it.  Every face is
    an axis-aligned rectangle, so two faces smooth exactly when their normals
    agree (the default threshold passes 0 degrees and stops 90); each list is
    in ascending face order.  Returns (ptr (F + 1), nbr) as int32."""
    F = g.nface
    lo = g.face_lo.astype(np.float64)
    hi = g.face_hi.astype(np.float64)
    n = g.face_normal
    ax = np.argmax(np.abs(n), axis=1)
    nid = ax * 2 + (n[np.arange(F), ax] > 0)
    corners = []
    for su in (0, 1):
        for sv in (0, 1):
            c = lo.copy()
            u, v = (ax + 1) % 3, (ax + 2) % 3
            rows = np.arange(F)
            c[rows, u] = np.where(su, hi[rows, u], lo[rows, u])
            c[rows, v] = np.where(sv, hi[rows, v], lo[rows, v])
            corners.append(c)
    C = np.concatenate(corners)                        # (4F, 3)
    fid = np.tile(np.arange(F), 4)
    key = np.concatenate([np.round(C * 64).astype(np.int64), np.tile(nid, 4)[:, None]], axis=1)
    order = np.lexsort(key.T[::-1])
    key, fid = key[order], fid[order]
    same = np.all(key[1:] == key[:-1], axis=1)
    pairs = []
    for d in range(1, 5):
        if len(fid) <= d:
            break
        run = np.ones(len(fid) - d, bool)
        for k in range(d):
            run &= same[k:len(same) - d + 1 + k]
        a, b = fid[:-d][run], fid[d:][run]
        pairs.append(np.stack([a, b], 1))
        pairs.append(np.stack([b, a], 1))
    P = np.concatenate(pairs) if pairs else np.zeros((0, 2), np.int64)
    P = P[P[:, 0] != P[:, 1]]
    P = np.unique(P, axis=0)
    ptr = np.zeros(F + 1, np.int64)
    np.add.at(ptr, P[:, 0] + 1, 1)
    np.cumsum(ptr, out=ptr)
    return ptr.astype(np.int32), P[:, 1].astype(np.int32)


def room_final_inputs(g: RoomGrid, faces, luxel: float = 16.0):
    """The final stage's inputs for ``room_faces(g, luxel)``: each face's
    luxel frame, with s along the face's second in-plane axis and t along its
    first so that luxel s + t * width is sample s + t * width (samples run
    along the second axis fastest); the luxels are BuildFaceLuxels' grid
    points (luxel_pos None: made on the device from the frames); each
    sample's bounds its cell, (s -+ 0.5, t -+ 0.5); a face without samples
    (sky) is not lit.  The patches are the grid's, four-corner windings.
    Returns (FinalFaces, FinalPatches)."""
    from .final import FACE_DTYPE, FinalFaces, FinalPatches

    F = g.nface
    fi = np.arange(F)
    lo = g.face_lo.astype(np.float64)
    hi = g.face_hi.astype(np.float64)
    n = g.face_normal
    ax = np.argmax(np.abs(n), axis=1)
    u, v = (ax + 1) % 3, (ax + 2) % 3
    ext_u = hi[fi, u] - lo[fi, u]
    ext_v = hi[fi, v] - lo[fi, v]
    nu = np.maximum(1, np.round(ext_u / luxel).astype(np.int64))
    nv = np.maximum(1, np.round(ext_v / luxel).astype(np.int64))
    cu, cv = ext_u / nu, ext_v / nv
    cnt = np.diff(faces.face_sample_ptr)
    lit = cnt > 0
    fa = np.zeros(F, FACE_DTYPE)
    eu = np.zeros((F, 3))
    ev = np.zeros((F, 3))
    eu[fi, u] = 1.0
    ev[fi, v] = 1.0
    org = lo.copy()
    org[fi, u] += 0.5 * cu
    org[fi, v] += 0.5 * cv
    fa["origin"] = org
    fa["w2l_s"] = ev / cv[:, None]
    fa["w2l_t"] = eu / cu[:, None]
    fa["l2w_s"] = ev * cv[:, None]
    fa["l2w_t"] = eu * cu[:, None]
    fa["width"] = np.where(lit, nv, 0)
    fa["height"] = np.where(lit, nu, 0)
    fa["lit"] = lit
    S = int(faces.face_sample_ptr[-1])
    fs = np.repeat(fi, cnt)
    k = np.arange(S) - faces.face_sample_ptr[fs]
    t_, s_ = k // nv[fs], k % nv[fs]
    bounds = np.stack([s_ - 0.5, t_ - 0.5, s_ + 0.5, t_ + 0.5], 1).astype(np.float32)
    ptr, nbr = face_neighbours(g)
    ff = FinalFaces(fa, None, bounds, ptr, nbr)
    P = g.npatch
    fp = FinalPatches(g.origin.astype(np.float32), g.face.astype(np.int32), g.child1.astype(np.int32),
                      (4 * np.arange(P + 1)).astype(np.int32), g.winding.reshape(-1, 3).astype(np.float32))
    return ff, fp


def room_ambient_faces(g: RoomGrid, ff, offsets) -> np.ndarray:
    """The leaf-ambient stage's face records for the final stage's faces
    ``ff`` (FinalFaces) and lump ``offsets`` (vh_final_layout): lightmap
    vectors from each face's luxel frame (s = dot(p, w2l_s) - dot(origin,
    w2l_s)), mins, grid, light offset, the sky flag and reflectivity."""
    from .ambient import FACE_DTYPE

    fa = ff.faces
    out = np.zeros(len(fa), FACE_DTYPE)
    o = fa["origin"].astype(np.float64)
    for key, src in (("s", "w2l_s"), ("t", "w2l_t")):
        v = fa[src].astype(np.float64)
        out[key][:, :3] = fa[src]
        out[key][:, 3] = -(o * v).sum(1)
    out["min_s"] = fa["min_s"]
    out["min_t"] = fa["min_t"]
    out["smax"] = fa["width"]
    out["tmax"] = fa["height"]
    out["light_ofs"] = np.asarray(offsets, np.int32)
    out["sky"] = g.face_sky.astype(np.int32)
    out["reflectivity"] = g.face_refl
    return out


def room_ambient_points(g: RoomGrid, per_room: int = 32, seed: int = 5, margin: float = 8.0) -> np.ndarray:
    """Leaf-ambient sample points: ``per_room`` seeded points inside each
    room's interior, ``margin`` units off its walls (room-major)."""
    rng = np.random.default_rng(seed)
    lo = g.room_lo + margin
    hi = g.room_hi - margin
    u = rng.random((len(lo), per_room, 3))
    return (lo[:, None, :] + u * (hi - lo)[:, None, :]).reshape(-1, 3).astype(np.float32)


@dataclass
class SkyboxMap:
    """A room grid lit through sky ceilings, with two 3D skyboxes beside it:
    the grid, the casters (the grid's and the skyboxes'), the map's sky-camera
    entities and compiled BSP, and the skybox table the gather loads."""

    grid: RoomGrid
    casters: ShadowCasterSet
    entities: list
    tree: object  # skybox.CompiledBspTree
    leaf_area: np.ndarray
    area_count: int
    skybox: object  # skybox.Skybox


SKYBOX_AWAY = -8192.0  # the skyboxes' x, far outside the grid


def skybox_grid(nx: int = 2, ny: int = 2, seed: int = 11, nlights: int = 4, closed: bool = False,
                open_skybox: bool = False) -> SkyboxMap:
    """An ``nx`` x ``ny`` room grid (``ny`` >= 2) whose ceilings are sky, and
    two skybox rooms 512 x 512 x 256 at x = SKYBOX_AWAY (opaque floors, sky
    walls and ceilings), each with a sky camera that sees the whole grid
    scaled down inside it: camera 0 at scale 16, camera 1 at scale 32 (a
    third camera, scale 0, is ignored).  An opaque roof over the scaled image
    of room (0, 0) in skybox 0 and over room (nx-1, ny-1) in skybox 1 shades
    those rooms' sky rays in the skybox, and a tower in skybox 0 stands in
    the way of slanted ones.

    The BSP: x = -4096 splits the skyboxes (leaf 1, area 1, the cameras'
    area) from the grid; y between the first two rows of rooms splits row 1
    and beyond (leaf 2, area 2) from row 0, which the plane 0.6 x + 0.8 y =
    300 cuts into its far side (leaf 3, area 0) and a corner of room (0, 0)
    (leaf 4, area 1: no recursion there, as in the skybox itself).

    ``closed`` gives the grid opaque ceilings (no sky ray of the grid reaches
    sky); ``open_skybox`` leaves out the roofs and the tower (every skybox ray
    goes to sky)."""
    from .scenes import box_triangles
    from .skybox import CompiledBspTree, Skybox

    if ny < 2:
        raise ValueError("skybox_grid needs two rows of rooms")
    g = room_grid(nx, ny, sky_ceiling=not closed, nlights=nlights, seed=seed)
    W, D, H = (float(x) for x in (g.room_hi[0] - g.room_lo[0]))
    wall = float(g.room_lo[1][0] - g.room_hi[0][0]) if nx > 1 else 8.0
    ext = np.array([g.room_hi[:, 0].max(), g.room_hi[:, 1].max(), H], np.float64)
    v9 = [g.casters.vertices]
    ids = [g.casters.ids]
    base = int(np.max(g.casters.ids & 0x00FFFFFF)) + 1
    entities, cams = [], []
    for k, (scale, cy, roof_room) in enumerate(((16.0, -2048.0, 0), (32.0, 2048.0, nx * ny - 1))):
        c = np.array([SKYBOX_AWAY, cy, 0.0])
        lo, hi = c - (256.0, 256.0, 0.0), c + (256.0, 256.0, 256.0)
        for ax, sign, at, sky in ((2, +1, lo, False), (2, -1, hi, True), (0, +1, lo, True), (0, -1, hi, True),
                                  (1, +1, lo, True), (1, -1, hi, True)):
            flo, fhi = lo.copy(), hi.copy()
            flo[ax] = fhi[ax] = at[ax]
            w = _rect_winding(flo, fhi, ax, sign)
            v9.append(np.array([np.concatenate([w[0], w[1], w[2]]), np.concatenate([w[0], w[2], w[3]])],
                               np.float32))
            ids.append(np.full(2, (TRACE_ID_SKY if sky else TRACE_ID_OPAQUE) | base, np.int32))
            base += 1
        # the camera sees the grid's box [0, ext] scaled by 1 / scale, centred
        # on the skybox room, 32 units over its floor
        origin = c + np.array([0.0, 0.0, 32.0]) - 0.5 * np.array([ext[0], ext[1], 0.0]) / scale
        origin = origin.astype(np.float32)
        cams.append((origin, scale))
        rlo = (g.room_lo[roof_room] / scale + origin).astype(np.float64)
        rhi = (g.room_hi[roof_room] / scale + origin).astype(np.float64)
        # the roof overhangs the room's image by less than half a scaled wall
        roof_lo = np.array([rlo[0] - 0.1, rlo[1] - 0.1, rhi[2] + 24.0])
        roof_hi = np.array([rhi[0] + 0.1, rhi[1] + 0.1, rhi[2] + 28.0])
        boxes_lo, boxes_hi = ([], []) if open_skybox else ([roof_lo], [roof_hi])
        if k == 0 and not open_skybox:  # a tower beside the scaled grid
            boxes_lo.append(np.array([origin[0] + ext[0] / scale + 8.0, origin[1], 0.0]))
            boxes_hi.append(np.array([origin[0] + ext[0] / scale + 40.0, origin[1] + ext[1] / scale, 120.0]))
        if boxes_lo:
            bv, bid = box_triangles(np.array(boxes_lo), np.array(boxes_hi), first_id=base)
            base += 6 * len(boxes_lo)
            v9.append(bv)
            ids.append(bid)
        entities.append({"classname": "SKY_CAMERA" if k == 0 else "sky_camera",
                         "origin": "%r %r %r" % tuple(float(x) for x in origin), "scale": "%g" % scale})
        if k == 0:
            entities.append({"classname": "sky_camera", "origin": "%r %r %r" % tuple(float(x) for x in origin),
                             "scale": "0"})
    casters = ShadowCasterSet(np.concatenate(v9).astype(np.float32), np.concatenate(ids).astype(np.int32))
    row = float(g.room_hi[0][1]) + 0.5 * wall
    tree = CompiledBspTree(
        node_plane=[0, 1, 2],
        node_children=[[1, -2], [-3, 2], [-4, -5]],
        plane_normal=[[1, 0, 0], [0, 1, 0], [0.6, 0.8, 0.0]],
        plane_dist=[-4096.0, row, 300.0],
        plane_type=[0, 1, 3])
    leaf_area = np.array([0, 1, 2, 0, 1], np.int32)
    sb = Skybox.from_map(entities, tree, leaf_area, 3)
    return SkyboxMap(g, casters, entities, tree, leaf_area, 3, sb)
