"""Hand-built gather inputs shared by the oracle's closed-form tests and the
GPU parity tests: each case fixes a quantity of DirectLightGatherer exactly.
"""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np

from hl2sdk_tools_sharp_amd import light
from hl2sdk_tools_sharp_amd.light import FaceSet
from hl2sdk_tools_sharp_amd.tracer import ShadowCasterSet

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "detmath_pow.json")


def f32(bits: int) -> np.float32:
    return np.frombuffer(struct.pack("<I", bits), np.float32)[0]


def f64(bits: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def pow_rows():
    """The correctly rounded pow rows: the reference's table plus the hard
    arguments, as (kind, a, b, expected bits)."""
    with open(GOLDEN) as fh:
        d = json.load(fh)
    return d["reference"] + d["hard"]


def spot_pow_rows():
    """F2 rows whose base can be a spot cone's multiplier: 0 < x < 1, an
    exponent other than 0 and 1, a finite power."""
    out = []
    for r in pow_rows():
        if r["kind"] != "F2":
            continue
        x, y = f32(r["a"]), f32(r["b"])
        if 0 < x < 1 and np.isfinite(y) and y not in (0.0, 1.0):
            out.append((x, y, f32(r["expected"])))
    return out


def far_scene():
    """One triangle far away: nothing near the samples occludes."""
    v9 = np.array([[1e5, 1e5, 1e5, 1e5 + 1, 1e5, 1e5, 1e5, 1e5 + 1, 1e5]], np.float32)
    return ShadowCasterSet(v9, np.array([0x02000000], np.int32))


def spot_pow_case(xs, ys, height: float = 64.0):
    """One face per (x, y): a sample at (1000 k, 0, 0) facing up, a spot
    `height` (a power of two) straight above it with stopdot 1, stopdot2 0,
    exponent y, and the (unnormalised) direction (0, 0, -x).  The ray to the
    sample is exactly (0, 0, 1), so the cone dot -- and the fringe's
    multiplier -- is exactly x, and the light is
    pow(x, y) * (x * 2^-12) with q = 1.  Each light reaches only its own
    sample's cluster.  Returns (faces, lights, pvs, expected ldr R)."""
    xs = np.asarray(xs, np.float32)
    ys = np.asarray(ys, np.float32)
    n = len(xs)
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = 1000.0 * np.arange(n)
    pos[:, 2] = -1.0  # the point is one unit off the face
    normals = np.zeros((n, 4, 3), np.float32)
    normals[:, 0, 2] = 1.0
    faces = FaceSet(np.arange(n + 1, dtype=np.int32), np.ones(n, np.int32), normals, pos,
                    np.arange(n, dtype=np.int32))
    L = light.make_lights(n)
    L["type"] = light.EMIT_SPOTLIGHT
    L["origin"][:, 0] = pos[:, 0]
    L["origin"][:, 2] = height
    L["normal"][:, 2] = -xs
    L["intensity"] = 1.0
    L["intensity_hdr"] = 1.0
    L["quadratic_attn"] = 1.0
    L["stopdot"] = 1.0
    L["stopdot2"] = 0.0
    L["exponent"] = ys
    nb = (n + 7) // 8
    pvs = np.zeros((n, nb), np.uint8)
    for k in range(n):
        pvs[k, k >> 3] = 1 << (k & 7)
    return faces, L, pvs


def spot_pow_expected(xs, powers, height: float = 64.0):
    """The light spot_pow_case gives: mult * (x / height^2), each product
    rounded once in float32."""
    xs = np.asarray(xs, np.float32)
    inv = np.float32(1.0) / np.float32(np.float32(height) * np.float32(height))
    return (np.asarray(powers, np.float32) * (inv * xs)).astype(np.float32)


def one_face(points, normal=(0, 0, 1), bumped=False):
    n = len(points)
    normals = np.zeros((1, 4, 3), np.float32)
    normals[0, 0] = normal
    if bumped:
        normals[0, 1:] = light.bump_basis(np.array(normal, np.float32))
    return FaceSet(np.array([0, n], np.int32), np.array([4 if bumped else 1], np.int32), normals,
                   np.asarray(points, np.float32), np.zeros(n, np.int32))


# The GatherTests facts.  Each takes run(faces, lights, **kw) -> an object with
# ldr (4 slots, 4 normals, S, 4), face_styles and rays (traced), and returns
# whether the reference's closed form holds.

def _light(kind, **kw):
    L = light.make_lights(1)
    L["type"] = kind
    L["intensity"] = (1, 1, 1)
    L["intensity_hdr"] = (1, 1, 1)
    if kind in (light.EMIT_POINT, light.EMIT_SPOTLIGHT):
        L["constant_attn"] = 1.0  # DirectLight's default attenuation (GatherTests.Point)
    for k, v in kw.items():
        L[k] = v
    return L


def _plate(z, size=1000.0):
    """A square of two triangles at height z over the origin."""
    a = size
    v9 = np.array([[-a, -a, z, a, -a, z, a, a, z], [-a, -a, z, a, a, z, -a, a, z]], np.float32)
    return ShadowCasterSet(v9, np.array([0x02000005, 0x02000005], np.int32))


def _closed_box(r=50.0):
    from hl2sdk_tools_sharp_amd import scenes

    v9, ids = scenes.box_triangles(np.array([[-r, -r, -r]]), np.array([[r, r, r]]))
    return ShadowCasterSet(v9, ids)


def fact_point_overhead_unit_dot(run):
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_POINT, origin=(0, 0, 100)))
    return o.ldr[0, 0, 0, 0] == 1.0 and o.rays == 1


def fact_light_behind_asks_for_no_ray(run):  # GatherTests.cs:131-137
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_POINT, origin=(0, 0, -100)))
    return o.rays == 0 and o.ldr[0, 0, 0, 0] == 0.0


def fact_forty_five_degrees_is_the_cosine(run):  # :139-144
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_POINT, origin=(100, 0, 100)))
    return abs(float(o.ldr[0, 0, 0, 0]) - math.sqrt(0.5)) < 5e-7


def fact_distance_clamped_at_one(run):  # :118-124
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_POINT, origin=(0, 0, 0.5), constant_attn=0.0,
                                              quadratic_attn=1.0))
    return o.ldr[0, 0, 0, 0] == 1.0


def fact_spot_outside_outer_cone(run):  # :146-157
    L = _light(light.EMIT_SPOTLIGHT, origin=(100, 0, 100), normal=(0, 0, -1),
               stopdot=math.cos(math.radians(10)), stopdot2=math.cos(math.radians(20)))
    o = run(one_face([[0, 0, -1]]), L)
    return o.rays == 0 and o.ldr[0, 0, 0, 0] == 0.0


def fact_spot_inside_inner_cone_full_strength(run):  # :159-169
    L = _light(light.EMIT_SPOTLIGHT, origin=(0, 0, 100), normal=(0, 0, -1),
               stopdot=math.cos(math.radians(30)), stopdot2=math.cos(math.radians(45)))
    o = run(one_face([[0, 0, -1]]), L)
    return o.ldr[0, 0, 0, 0] == 1.0


def fact_spot_fringe_interpolates(run):  # :171-184 -- falloff = mult * dot2
    n = np.array([0.3, 0, -1], np.float32)
    n = (n / np.float32(np.sqrt(np.float32(np.dot(n, n))))).astype(np.float32)
    L = _light(light.EMIT_SPOTLIGHT, origin=(0, 0, 100), normal=n, stopdot=0.99, stopdot2=0.9)
    o = run(one_face([[0, 0, -1]]), L)
    dot2 = np.float32(-np.float32(n[2]))
    mult = np.float32((dot2 - np.float32(0.9)) / (np.float32(0.99) - np.float32(0.9)))
    return abs(float(o.ldr[0, 0, 0, 0]) - float(mult * dot2)) < 1e-5


def fact_spot_exponent_is_the_correctly_rounded_power(run):
    """Exponent 2 gives (float)((double)m * m), exponent 0.5 the correctly
    rounded square root, and every golden row's power (the reference's mpmath
    table) comes out of the fringe exactly."""
    rng = np.random.default_rng(11)
    xs = rng.uniform(0.01, 1.0, 40).astype(np.float32)
    want2 = (xs.astype(np.float64) * xs.astype(np.float64)).astype(np.float32)
    want5 = np.sqrt(xs)  # float32 sqrt is correctly rounded
    rows = spot_pow_rows()
    X = np.concatenate([xs, xs, [r[0] for r in rows]]).astype(np.float32)
    Y = np.concatenate([np.full(40, 2.0), np.full(40, 0.5), [r[1] for r in rows]]).astype(np.float32)
    P = np.concatenate([want2, want5, [r[2] for r in rows]]).astype(np.float32)
    faces, L, pvs = spot_pow_case(X, Y)
    o = run(faces, L, pvs=pvs)
    return np.array_equal(o.ldr[0, 0, :, 0], spot_pow_expected(X, P))


def fact_surface_light_falls_off_and_ends_off_the_emitter(run):  # :197-210
    L = _light(light.EMIT_SURFACE, origin=(0, 0, 10), normal=(0, 0, -1))
    f = one_face([[0, 0, -1]])
    open_ = run(f, L)
    # the ray ends DistEpsilon (1/32) short of the emitter at z = 10: a plate
    # at 9.98 lies past its end, one at 9.95 across it
    past = run(f, L, casters=_plate(9.98))
    across = run(f, L, casters=_plate(9.95))
    return (open_.ldr[0, 0, 0, 0] == np.float32(1.0) / np.float32(100.0) and past.ldr[0, 0, 0, 0] == open_.ldr[0, 0, 0, 0]
            and across.ldr[0, 0, 0, 0] == 0.0)


def fact_beyond_hard_falloff_end_nothing(run):  # :212-221
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_POINT, origin=(0, 0, 100), start_fade=10.0,
                                              end_fade=50.0))
    return o.rays == 0 and o.ldr[0, 0, 0, 0] == 0.0


def fact_quintic_fade_midpoint(run):  # :223-232
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_POINT, origin=(0, 0, 30), start_fade=10.0,
                                              end_fade=50.0))
    return abs(float(o.ldr[0, 0, 0, 0]) - 0.5) < 5e-7


def fact_sun_casts_back_along_its_direction(run):  # :234-243 -- to MaxTraceLength
    L = _light(light.EMIT_SKYLIGHT, normal=(0, 0, -1))
    f = one_face([[0, 0, -1]])
    o = run(f, L)
    near = run(f, L, casters=_plate(50000.0, 1e6))  # inside sqrt(3) 32768 = 56755
    far = run(f, L, casters=_plate(60000.0, 1e6))
    return (o.rays == 1 and o.ldr[0, 0, 0, 0] == 1.0 and o.ldr[0, 0, 0, 3] == 10000.0
            and near.ldr[0, 0, 0, 0] == 0.0 and far.ldr[0, 0, 0, 0] == 1.0)


def fact_soft_sun_casts_thirty_rays(run):  # :253-258
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_SKYLIGHT, normal=(0, 0, -1)),
               sun_samples=light.SUN_AREA_LIGHT_SAMPLES, sun_extent=0.05)
    return o.rays == light.SUN_AREA_LIGHT_SAMPLES


def fact_open_sky_ambient_unit_dot(run):  # :267-278
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_SKYAMBIENT), sky_dirs=light.sky_directions())
    return abs(float(o.ldr[0, 0, 0, 0]) - 1.0) < 1e-6 and 1 <= o.rays <= light.VERTEX_NORMAL_COUNT - 1


def fact_closed_sky_ambient_nothing(run):  # :280-286
    o = run(one_face([[0, 0, -1]]), _light(light.EMIT_SKYAMBIENT), sky_dirs=light.sky_directions(),
               casters=_closed_box())
    return o.ldr[0, 0, 0, 0] == 0.0


def fact_padding_lanes_not_traced(run):  # :288-295
    o = run(one_face([[0, 0, -1], [10, 0, -1]]), _light(light.EMIT_POINT, origin=(0, 0, 100)))
    return o.rays == 2 and o.ldr[0, 0, 0, 0] > 0 and o.ldr[0, 0, 1, 0] > 0


def fact_bump_dot_zeroed_with_the_flat_dot(run):  # :297-312
    f = one_face([[0, 0, -1]])
    f.face_normal_count[:] = 4
    f.face_normals[0, 1:] = (0, 0, -1)
    o = run(f, _light(light.EMIT_POINT, origin=(0, 0, -100)))
    return not np.any(o.ldr[0, :, 0, :3])


# --- the 3D skybox (TestLineDoesHitSky's recursion, DirectLightGatherer.cs:510-714;
# SkyRecursionTests, GatherTests.cs:361-555), on rooms.skybox_grid ---

def _skybox_case(closed=False, open_skybox=False):
    from hl2sdk_tools_sharp_amd import rooms

    m = rooms.skybox_grid(closed=closed, open_skybox=open_skybox)
    return m, rooms.room_faces(m.grid, luxel=32.0)


def _sun_only(m, normal):
    from hl2sdk_tools_sharp_amd import rooms

    L, pvs = rooms.room_lights(m.grid, sun=True)
    k = len(m.grid.lights)
    L, pvs = L[k:k + 1].copy(), pvs[k:k + 1].copy()
    L[0]["normal"] = normal
    return L, pvs


def fact_skybox_roof_shades_its_rooms(run):
    """A vertical sun over sky ceilings: every floor sample's first ray sees
    sky, so a sample whose leaf recurses is shaded exactly where its point,
    scaled into a camera's skybox, lies under that skybox's roof -- room (0, 0)
    in skybox 0, the last room in skybox 1 -- and keeps its sunlight
    elsewhere, as the samples of the corner leaf (the cameras' own area)
    do under the roof."""
    m, f = _skybox_case()
    L, pvs = _sun_only(m, (0.0, 0.0, -1.0))
    base = run(f, L, pvs=pvs, casters=m.casters).ldr[0, 0, :, 3]
    got = run(f, L, pvs=pvs, casters=m.casters, skybox=m.skybox).ldr[0, 0, :, 3]
    g = m.grid
    p = f.sample_pos + f.face_normals[np.repeat(np.arange(f.nfaces), np.diff(f.face_sample_ptr)), 0]
    rec = np.array([m.skybox.leaf_recurses[m.tree.leaf_from_point(q)] for q in p], bool)
    last = len(g.room_lo) - 1
    under = (np.all((p[:, :2] >= g.room_lo[0, :2]) & (p[:, :2] <= g.room_hi[0, :2]), axis=1)
             | np.all((p[:, :2] >= g.room_lo[last, :2]) & (p[:, :2] <= g.room_hi[last, :2]), axis=1))
    want = np.where(rec & under, np.float32(0.0), base)
    lit_under = (base > 0) & under
    return (np.array_equal(got, want) and np.any(lit_under & rec) and np.any(lit_under & ~rec)
            and np.any(base[~under] > 0))


def fact_skybox_only_adds_occlusion(run):
    """The skybox rays can only add to a lane's occlusion: with a slanted,
    jittered sun, sky ambient and point lights, no output of any sample
    grows, and some fall."""
    from hl2sdk_tools_sharp_amd import rooms

    m, f = _skybox_case()
    L, pvs = rooms.room_lights(m.grid, sun=True, sky_ambient=True)
    sun = np.array([0.15, 0.1, 1.0], np.float32)
    L[len(m.grid.lights)]["normal"] = -sun / np.float32(np.linalg.norm(sun))
    kw = dict(pvs=pvs, casters=m.casters, sun_samples=light.SUN_AREA_LIGHT_SAMPLES, sun_extent=0.05,
              sky_dirs=light.sky_directions())
    a = run(f, L, **kw)
    b = run(f, L, skybox=m.skybox, **kw)
    return bool(np.all(b.ldr <= a.ldr) and np.all(b.hdr <= a.hdr) and np.any(b.ldr < a.ldr))


def fact_skybox_closed_world_casts_no_skybox_rays(run):
    """Opaque ceilings: every group's first rays are all blocked, so no
    skybox ray is cast (AFullyOccludedDeferredTestTracesNoSkyboxRays) -- the
    rays equal those of a skybox no leaf recurses into, and so does the light."""
    from hl2sdk_tools_sharp_amd import rooms
    from hl2sdk_tools_sharp_amd.skybox import Skybox

    m, f = _skybox_case(closed=True)
    L, pvs = rooms.room_lights(m.grid, sun=True, sky_ambient=True)
    kw = dict(pvs=pvs, casters=m.casters, sky_dirs=light.sky_directions())
    none = Skybox(m.skybox.camera_origin, m.skybox.camera_world_to_sky, m.tree,
                  np.zeros_like(m.skybox.leaf_recurses))
    a = run(f, L, skybox=none, **kw)
    b = run(f, L, skybox=m.skybox, **kw)
    return a.rays == b.rays and np.array_equal(a.ldr, b.ldr) and a.rays > 0


def fact_skybox_open_changes_nothing(run):
    """Skyboxes with nothing in them but sky: every skybox ray is answered
    sky, and the light is the light without skyboxes, bit for bit."""
    from hl2sdk_tools_sharp_amd import rooms

    m, f = _skybox_case(open_skybox=True)
    L, pvs = rooms.room_lights(m.grid, sun=True, sky_ambient=True)
    kw = dict(pvs=pvs, casters=m.casters, sky_dirs=light.sky_directions())
    a = run(f, L, **kw)
    b = run(f, L, skybox=m.skybox, **kw)
    return np.array_equal(a.ldr.view(np.uint32), b.ldr.view(np.uint32)) and b.rays > a.rays


FACTS = [
    (fact_point_overhead_unit_dot, None),
    (fact_light_behind_asks_for_no_ray, 5),
    (fact_forty_five_degrees_is_the_cosine, None),
    (fact_distance_clamped_at_one, 10),
    (fact_spot_outside_outer_cone, None),
    (fact_spot_inside_inner_cone_full_strength, None),
    (fact_spot_fringe_interpolates, 2),
    (fact_spot_fringe_interpolates, 3),
    (fact_spot_exponent_is_the_correctly_rounded_power, 7),
    (fact_surface_light_falls_off_and_ends_off_the_emitter, 1),
    (fact_beyond_hard_falloff_end_nothing, None),
    (fact_quintic_fade_midpoint, 8),
    (fact_sun_casts_back_along_its_direction, 11),
    (fact_sun_casts_back_along_its_direction, 12),
    (fact_soft_sun_casts_thirty_rays, None),
    (fact_open_sky_ambient_unit_dot, None),
    (fact_closed_sky_ambient_nothing, 9),
    (fact_padding_lanes_not_traced, 6),
    (fact_bump_dot_zeroed_with_the_flat_dot, 4),
    (fact_skybox_roof_shades_its_rooms, 14),
    (fact_skybox_roof_shades_its_rooms, 15),
    (fact_skybox_only_adds_occlusion, 17),
    (fact_skybox_closed_world_casts_no_skybox_rays, 16),
    (fact_skybox_open_changes_nothing, None),
]


