"""The oracle's gather, sampler and encode against the reference's known
answers and closed forms.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.gather as og
from hl2sdk_tools_sharp_amd import light
from hl2sdk_tools_sharp_amd.light import FaceSet
from hl2sdk_tools_sharp_amd.tracer import ShadowCasterSet


# --- HaltonSequence / DirectionalSampler (src/SourceSharp.Tests/MapTools/Rad/Light/LightMathTests.cs:293-345) ---

def test_halton_known_values():
    assert [og.halton(2, i) for i in (1, 2, 3)] == [0.25, 0.75, 0.125]
    assert og.halton(3, 1) == np.float32(2.0 / 3.0)


def test_directional_sampler_first_value_and_unit_length():
    d = og.directional_sampler(162)
    assert d[0, 2] == np.float32(-0.5)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-7)
    # the product's table is the same bits
    assert np.array_equal(d.view(np.uint32), light.directional_sampler(162).view(np.uint32))


# --- StockLightColor.Encode (Rad/Ambient/StockLightColor.cs:136-185) ---

def encode_independent(v):
    """ColorRGBExp32 re-derived from the format: the exponent e is chosen so the
    largest channel's mantissa lands in [128, 256) (biased exponent - 134), and
    each channel is truncated at 2^-e."""
    v = np.asarray(v, np.float32)
    out = []
    for r, g, b in v:
        m = max(r, g, b) if not (r > g) else (r if r > b else b)
        m = (r if r > b else b) if r > g else (g if g > b else b)
        e = 0 if m == 0 else int(np.frexp(np.float64(m))[1]) - 1 - 7
        if m != 0 and not np.isfinite(m):
            e = 255 - 134
        s = np.float32(2.0 ** (-e)) if -126 <= -e <= 127 else None
        def q(x):
            if s is None:
                return None
            y = np.float32(x * s)
            if np.isnan(y):
                return 0
            return int(np.clip(np.trunc(np.float64(y)), -2**31, 2**31 - 1)) & 0xFF
        out.append((q(r), q(g), q(b), e & 0xFF))
    return out


def test_encode_known_values():
    got = og.encode_rgbe(np.array([[0, 0, 0], [1, 1, 1], [255, 128, 0], [0.5, 2.0, 0.25]], np.float32))
    assert got[0].tolist() == [0, 0, 0, 0]
    assert got[1].tolist() == [128, 128, 128, np.uint8(np.int8(-7))]
    assert got[2].tolist() == [255, 128, 0, 0]
    assert got[3].tolist() == [32, 128, 16, np.uint8(np.int8(-6))]


def test_encode_matches_an_independent_derivation():
    rng = np.random.default_rng(0)
    v = (rng.uniform(0, 1, (4000, 3)) * 10.0 ** rng.uniform(-6, 6, (4000, 1))).astype(np.float32)
    got = og.encode_rgbe(v)
    want = encode_independent(v)
    assert [tuple(int(x) for x in r) for r in got] == want


def test_encode_roundtrip_within_a_mantissa_step():
    rng = np.random.default_rng(1)
    v = (rng.uniform(0.01, 1, (2000, 3)) * 10.0 ** rng.uniform(-3, 4, (2000, 1))).astype(np.float32)
    e = og.encode_rgbe(v)
    dec = e[:, :3].astype(np.float64) * 2.0 ** e[:, 3].astype(np.int8)[:, None]
    m = v.max(axis=1, keepdims=True)
    assert np.all(np.abs(dec - v) <= m / 128.0 + 1e-30)


def test_encode_ties_go_to_the_later_channel():
    # equal red and green: the maximum is taken from the later comparison
    got = og.encode_rgbe(np.array([[3.0, 3.0, 1.0]], np.float32))
    assert got[0, 3] == np.uint8(np.int8(-6))


# --- the gather's closed forms (src/SourceSharp.Tests/MapTools/Rad/Light/GatherTests.cs:101-285) ---

def far_scene():
    """One triangle far away: nothing near the samples occludes."""
    v9 = np.array([[1e5, 1e5, 1e5, 1e5 + 1, 1e5, 1e5, 1e5, 1e5 + 1, 1e5]], np.float32)
    return ShadowCasterSet(v9, np.array([0x02000000], np.int32))


def one_face(points, normal=(0, 0, 1), bumped=False):
    n = len(points)
    normals = np.zeros((1, 4, 3), np.float32)
    normals[0, 0] = normal
    if bumped:
        normals[0, 1:] = light.bump_basis(np.array(normal, np.float32))
    return FaceSet(np.array([0, n], np.int32), np.array([4 if bumped else 1], np.int32), normals,
                   np.asarray(points, np.float32), np.zeros(n, np.int32))


def run(faces, L, **kw):
    d = light.pack_desc(faces, L, kw.get("pvs"), kw.get("sun_extent", 0.0), kw.get("sun_samples", 1),
                        kw.get("sky_dirs"), kw.get("ao_count", 0), kw.get("ao_radius", 40.0))
    return og.gather(kw.get("casters", far_scene()), d, faces.nsamples)


def run_full(faces, L, **kw):
    """The gather with style slots and the ray count."""
    d = light.pack_desc(faces, L, kw.get("pvs"), kw.get("sun_extent", 0.0), kw.get("sun_samples", 1),
                        kw.get("sky_dirs"), kw.get("ao_count", 0), kw.get("ao_radius", 40.0),
                        kw.get("skybox"))
    return og.gather_styles(kw.get("casters", far_scene()), d, faces.nsamples)


def test_point_light_inverse_square_and_unit_dot():
    f = one_face([[0, 0, -1]])  # the sample point is one unit off: (0, 0, 0)
    L = light.make_lights(1)
    L["type"], L["origin"], L["intensity"], L["quadratic_attn"] = light.EMIT_POINT, (0, 0, 10), (1, 1, 1), 1
    ldr, _, _ = run(f, L)
    assert ldr[0, 0, 0] == pytest.approx(0.01, rel=1e-6)


def test_falloff_distance_is_clamped_at_one():
    f = one_face([[0, 0, -1]])
    L = light.make_lights(1)
    L["type"], L["origin"], L["intensity"], L["quadratic_attn"] = light.EMIT_POINT, (0, 0, 0.5), (1, 1, 1), 1
    ldr, _, _ = run(f, L)
    assert ldr[0, 0, 0] == pytest.approx(1.0, rel=1e-6)


def test_quintic_fade_is_half_at_the_midpoint():
    f = one_face([[0, 0, -1]])
    L = light.make_lights(1)
    L["type"], L["origin"], L["intensity"], L["constant_attn"] = light.EMIT_POINT, (0, 0, 50), (1, 1, 1), 1
    L["start_fade"], L["end_fade"] = 0.0, 100.0
    ldr, _, _ = run(f, L)
    assert ldr[0, 0, 0] == pytest.approx(0.5, rel=1e-6)


def test_sun_with_open_sky_is_full_and_amount_ten_thousand():
    f = one_face([[0, 0, -1]])
    L = light.make_lights(1)
    L["type"], L["normal"], L["intensity"] = light.EMIT_SKYLIGHT, (0, 0, -1), (1, 2, 3)
    ldr, _, _ = run(f, L, sun_samples=30, sun_extent=0.01)
    np.testing.assert_allclose(ldr[0, 0], [1, 2, 3, 10000], rtol=1e-6)


def test_sky_ambient_open_sky_gives_dot_one():
    f = one_face([[0, 0, -1]], bumped=True)
    L = light.make_lights(1)
    L["type"], L["intensity"] = light.EMIT_SKYAMBIENT, (1, 1, 1)
    ldr, _, _ = run(f, L, sky_dirs=light.sky_directions())
    assert ldr[0, 0, 0] == pytest.approx(1.0, rel=1e-5)


def test_spot_inside_cone_and_outside():
    f = one_face([[0, 0, -1], [500, 0, -1]])
    L = light.make_lights(1)
    L["type"], L["origin"], L["normal"], L["intensity"] = light.EMIT_SPOTLIGHT, (0, 0, 100), (0, 0, -1), (1, 1, 1)
    L["quadratic_attn"], L["stopdot"], L["stopdot2"] = 1.0, np.cos(np.radians(20)), np.cos(np.radians(30))
    ldr, _, _ = run(f, L)
    assert ldr[0, 0, 0] == pytest.approx(1.0 / 10000.0, rel=1e-5)  # straight below: dot2 = 1
    assert ldr[0, 1, 0] == 0.0  # 79 degrees off the axis


def test_a_blocked_light_contributes_nothing_and_sun_amount_is_group_coupled():
    """Two samples of one 4-sample group: one sees the sun, one is under a
    roof.  The shaded sample gets neither light nor sun amount."""
    roof = np.array([[50, -50, 10, 150, -50, 10, 150, 50, 10], [50, -50, 10, 150, 50, 10, 50, 50, 10]],
                    np.float32)
    casters = ShadowCasterSet(roof, np.array([0x02000001, 0x02000001], np.int32))
    f = one_face([[0, 0, -1], [100, 0, -1]])
    L = light.make_lights(1)
    L["type"], L["normal"], L["intensity"] = light.EMIT_SKYLIGHT, (0, 0, -1), (1, 1, 1)
    ldr, _, _ = run(f, L, casters=casters)
    assert ldr[0, 0, 0] == 1.0 and ldr[0, 0, 3] == 10000.0
    assert ldr[0, 1, 0] == 0.0 and ldr[0, 1, 3] == 0.0


def test_ao_open_and_closed():
    f = one_face([[0, 0, -1]])
    L = light.make_lights(0)
    _, _, ao = run(f, L, ao_count=32, ao_radius=40.0)
    assert ao[0] == 1.0
    lid = np.array([[-100, -100, 5, 100, -100, 5, 100, 100, 5], [-100, -100, 5, 100, 100, 5, -100, 100, 5]],
                   np.float32)
    _, _, ao = run(f, L, ao_count=32, ao_radius=40.0,
                   casters=ShadowCasterSet(lid, np.array([0x02000000, 0x02000000], np.int32)))
    # the lid at height 5 over the sample (0, 0, 0) blocks exactly the
    # directions that reach it inside the 40-unit radius: d_z > 5/40
    dz = light.ao_directions(np.array([[0, 0, 1]], np.float32), 32)[0, :, 2]
    assert 0 < np.count_nonzero(dz <= 0.125) < 32
    assert ao[0] == np.float32(np.count_nonzero(dz <= 0.125) / 32)


# --- every GatherTests fact (src/SourceSharp.Tests/MapTools/Rad/Light/GatherTests.cs:101-312),
# each with the planted mistake it must catch; the facts live in gather_cases
# so the GPU tests hold the device to the same closed forms ---

import math  # noqa: E402

import gather_cases as gc  # noqa: E402


def gc_run(faces, L, **kw):
    return run_full(faces, L, **kw)


FACTS = gc.FACTS


@pytest.mark.parametrize("fact", sorted({f for f, _ in FACTS}, key=lambda f: f.__name__),
                         ids=lambda f: f.__name__)
def test_gather_fact(fact):
    og.set_mutation(0)
    assert fact(gc_run)


@pytest.mark.parametrize("fact,mutation", [(f, m) for f, m in FACTS if m is not None],
                         ids=lambda v: v.__name__ if callable(v) else str(v))
def test_planted_mistake_breaks_its_fact(fact, mutation):
    """The oracle with one planted mistake fails the fact that pins it."""
    og.set_mutation(mutation)
    try:
        assert not fact(gc_run)
    finally:
        og.set_mutation(0)


# --- DetMathF.Pow / DetMath.PowToSingle: the reference's mpmath table and hard arguments ---

def test_pow_matches_the_reference_table_and_hard_rows():
    rows = gc.pow_rows()
    f2 = [r for r in rows if r["kind"] == "F2"]
    s2 = [r for r in rows if r["kind"] == "S2"]
    x = np.array([gc.f32(r["a"]) for r in f2], np.float32)
    y = np.array([gc.f32(r["b"]) for r in f2], np.float32)
    got = og.pow_f32(x, y).view(np.uint32)
    assert got.tolist() == [r["expected"] for r in f2]
    xd = np.array([gc.f64(r["a"]) for r in s2])
    yd = np.array([gc.f64(r["b"]) for r in s2])
    got = og.pow_to_single(xd, yd).view(np.uint32)
    assert got.tolist() == [r["expected"] for r in s2]


def test_hard_rows_take_the_exact_step_and_double_rounding_would_miss():
    """The hard rows exercise the double-double step, and at least one of them
    is one where (float)pow(double) is not the correctly rounded answer."""
    hard = json_rows("hard")
    assert hard and all(og.pow_needs_exact(0 if r["kind"] == "F2" else 1,
                                           float(gc.f32(r["a"])) if r["kind"] == "F2" else gc.f64(r["a"]),
                                           float(gc.f32(r["b"])) if r["kind"] == "F2" else gc.f64(r["b"]))
                        for r in hard)
    naive_wrong = 0
    for r in hard:
        if r["kind"] != "F2":
            continue
        v = np.float32(math.pow(float(gc.f32(r["a"])), float(gc.f32(r["b"]))))
        naive_wrong += int(v.view(np.uint32)) != r["expected"]
    assert naive_wrong >= 1


def json_rows(part):
    import json

    with open(gc.GOLDEN) as fh:
        return json.load(fh)[part]


def test_pow_against_mpmath_on_random_arguments():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 120
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, 300).astype(np.float32)
    y = rng.uniform(0.1, 9, 300).astype(np.float32)
    got = og.pow_f32(x, y)
    for xi, yi, g in zip(x, y, got):
        exact = mpmath.power(mpmath.mpf(float(xi)), mpmath.mpf(float(yi)))
        lo = np.nextafter(g, np.float32(0))
        hi = np.nextafter(g, np.float32(2))
        # g is the nearest float: the exact value lies within half an ulp of it
        assert abs(exact - float(g)) <= min(abs(exact - float(lo)), abs(exact - float(hi)))


# --- light styles: FaceLight.FindOrAllocateStyle (Rad/Light/FaceLight.cs:204-222) ---

def _styled(kind_origin_style):
    L = light.make_lights(len(kind_origin_style))
    for k, (origin, style) in enumerate(kind_origin_style):
        L[k]["type"] = light.EMIT_POINT
        L[k]["origin"] = origin
        L[k]["constant_attn"] = 1.0
        L[k]["intensity"] = (1 + k, 2 + k, 3 + k)
        L[k]["intensity_hdr"] = (1 + k, 2 + k, 3 + k)
        L[k]["style"] = style
    return L


def test_style_zero_is_always_slot_zero_and_other_styles_take_the_next_slots():
    f = one_face([[0, 0, -1]])
    L = _styled([((0, 0, 100), 2), ((5, 0, 100), 0)])
    o = gc_run(f, L)
    assert o.face_styles[0].tolist() == [0, 2, 255, 255]
    one = gc_run(f, L[1:])
    two = gc_run(f, L[:1])
    assert np.array_equal(o.ldr[0, 0, 0], one.ldr[0, 0, 0])  # the style-0 light in slot 0
    assert two.face_styles[0].tolist() == [0, 2, 255, 255] and not np.any(two.ldr[0])
    assert np.array_equal(o.ldr[1, 0, 0], two.ldr[1, 0, 0])  # style 2 in slot 1, alone or not


def test_style_slots_follow_the_first_touch_group_by_group():
    """Light 0 (style 7) reaches only the second group's samples, light 1
    (style 3) every sample: group 0 meets style 3 first, so slot 1 is style 3
    and slot 2 style 7, whatever the light order."""
    pts = [[x, 0, -1] for x in (0, 1, 2, 3)] + [[x, 0, -1] for x in (1000, 1001, 1002, 1003)]
    f = one_face(pts)
    L = _styled([((1000, 0, 30), 7), ((0, 0, 100), 3)])
    L[0]["start_fade"], L[0]["end_fade"] = 0.0, 60.0  # only the far group is within 60 units
    o = gc_run(f, L)
    assert o.face_styles[0].tolist() == [0, 3, 7, 255]
    assert not np.any(o.ldr[2, 0, :4, :3]) and np.all(o.ldr[2, 0, 4:, 0] > 0)
    og.set_mutation(13)
    try:
        assert gc_run(f, L).face_styles[0].tolist() != [0, 3, 7, 255]
    finally:
        og.set_mutation(0)


def test_a_fifth_style_is_dropped():
    f = one_face([[0, 0, -1]])
    L = _styled([((k, 0, 100), s) for k, s in enumerate([1, 2, 3, 4, 0])])
    o = gc_run(f, L)
    assert o.face_styles[0].tolist() == [0, 1, 2, 3]
    assert o.dropped == 1
    alone = gc_run(f, L[4:])
    assert np.array_equal(o.ldr[0, 0, 0], alone.ldr[0, 0, 0])
