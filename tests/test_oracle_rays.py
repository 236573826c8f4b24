"""The oracle's ray functions against the reference's golden vectors and
against closed forms.  CPU only."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import kd_fixtures


def correct_digest(oracle, scene, tris) -> str:
    """VradCpuIndependenceTests.TheCorrectKdSceneDistancesArePinnedOnEveryCpu
    (src/SourceSharp.Tests/MapTools/Rad/VradCpuIndependenceTests.cs:114-149).
    Closest hits come from the oracle's brute force, clipped at the end; a hit
    that stock's walk found BEYOND the end (a leaf-membership effect of the KD
    tree, HitId.cs:23-31) is re-evaluated on that same triangle."""
    idx, t = oracle.closest_brute(tris, scene.rays, clip=True)
    h = hashlib.sha256()
    for tr in tris:
        h.update(np.asarray(list(tr["n"]) + [tr["d"]], np.float32).tobytes())
        h.update(np.asarray(tr["e"], np.float32).tobytes())
        h.update(np.int32(tr["id"]).tobytes())
        h.update(bytes([int(tr["cs0"]), int(tr["cs1"])]))
    for i in range(len(scene.rays)):
        if idx[i] >= 0:
            s, f = int(idx[i]), np.float32(t[i])
        elif scene.stock_hit[i] >= 0:
            s = int(scene.stock_hit[i])
            k, tt = oracle.closest_brute(tris[s : s + 1], scene.rays[i : i + 1], clip=False)
            s, f = (s, np.float32(tt[0])) if k[0] == 0 else (-1, np.float32(1.0))
        else:
            s, f = -1, np.float32(1.0)
        h.update(np.int32(s).tobytes())
        h.update(np.float32(f).tobytes())
    return h.hexdigest().upper()


@pytest.mark.parametrize("name", ["kd-scene", "kd-boxes"])
def test_correct_mode_digest_matches_reference(oracle_lib, kd_scenes, name):
    scene = kd_scenes[name]
    tris = oracle_lib.convert_triangles(scene.v9)
    assert correct_digest(oracle_lib, scene, tris) == kd_fixtures.CORRECT_DIGESTS[name]


def test_planted_mistake_moves_the_digest(oracle_lib, kd_scenes):
    """The digest is sensitive: swapping the projection axes in the oracle's
    output (a plausible transposition) changes it."""
    scene = kd_scenes["kd-boxes"]
    tris = oracle_lib.convert_triangles(scene.v9)
    bad = tris.copy()
    bad["cs0"], bad["cs1"] = tris["cs1"], tris["cs0"]
    assert correct_digest(oracle_lib, scene, bad) != kd_fixtures.CORRECT_DIGESTS["kd-boxes"]


@pytest.mark.parametrize("name", ["kd-scene", "kd-boxes"])
def test_hit_triangles_match_stock(oracle_lib, kd_scenes, name):
    """EveryRayHitsTheSameTriangleAsStock (KdRayTracerStockParityTests.cs:236-265),
    for every ray stock hit short of the end; no other ray hits."""
    scene = kd_scenes[name]
    tris = oracle_lib.convert_triangles(scene.v9)
    idx, _ = oracle_lib.closest_brute(tris, scene.rays, clip=True)
    short = scene.stock_blocked
    assert short.sum() > len(scene.rays) // 4
    assert np.array_equal(idx[short], scene.stock_hit[short])
    assert np.all(idx[~short] == -1)


@pytest.mark.parametrize("name", ["kd-scene", "kd-boxes"])
def test_visibility_bits_match_stock(oracle_lib, kd_scenes, name):
    """VisibilityBitsAgreeWithStocksHitsShortOfTheEnd (KdRayTracerStockParityTests.cs:284-302): 0 differing."""
    scene = kd_scenes[name]
    tris = oracle_lib.convert_triangles(scene.v9)
    bits = oracle_lib.visibility_brute(tris, scene.rays)
    got = oracle_lib.unpack_bits(bits, len(scene.rays))
    assert np.array_equal(got, scene.stock_blocked)


@pytest.mark.parametrize("name", ["kd-scene", "kd-boxes"])
def test_triangle_format_close_to_stock_planes(oracle_lib, kd_scenes, name):
    """Stock normalises with an rsqrt estimate; the Correct side divides.  The
    projection axes must be identical and every coefficient within a few ulps."""
    scene = kd_scenes[name]
    tris = oracle_lib.convert_triangles(scene.v9)
    assert np.array_equal(tris["cs0"], scene.stock_cs[:, 0])
    assert np.array_equal(tris["cs1"], scene.stock_cs[:, 1])
    np.testing.assert_allclose(tris["n"], scene.stock_normals, rtol=0, atol=2e-6)
    np.testing.assert_allclose(tris["d"], scene.stock_d, rtol=2e-6, atol=1e-3)
    np.testing.assert_allclose(tris["e"], scene.stock_edges, rtol=1e-4, atol=1e-6)


def _random_rays(rng, n, lo, hi, reach_mode="one"):
    a = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    b = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    r = np.empty((n, 7), np.float32)
    r[:, :3] = a
    r[:, 3:6] = b - a
    r[:, 6] = 1.0
    return r


@pytest.mark.parametrize("name", ["kd-scene", "kd-boxes"])
def test_cpu_bvh_equals_brute_force(oracle_lib, kd_scenes, name):
    scene = kd_scenes[name]
    tris = oracle_lib.convert_triangles(scene.v9)
    bvh = oracle_lib.CpuBvh(scene.v9, tris)
    for tmin in (0.0, 1e-3, 0.25):
        assert np.array_equal(
            bvh.visibility(scene.rays, tmin=tmin), oracle_lib.visibility_brute(tris, scene.rays, tmin=tmin)
        )
        i1, t1 = bvh.closest(scene.rays, tmin=tmin)
        i2, t2 = oracle_lib.closest_brute(tris, scene.rays, tmin=tmin, clip=True)
        assert np.array_equal(i1, i2)
        assert np.array_equal(t1[i1 >= 0].view(np.uint32), t2[i2 >= 0].view(np.uint32))


def test_cpu_bvh_options_equal_brute_force(oracle_lib):
    from hl2sdk_tools_sharp_amd import scenes

    rng = np.random.default_rng(7)
    lo, hi = scenes.random_boxes(60, rng, volume=1000.0, edge_min=8, edge_max=200)
    v9, ids = scenes.box_triangles(lo, hi)
    ids = ids.copy()
    ids[:24] |= 0x01000000  # a couple of sky boxes
    tris = oracle_lib.convert_triangles(v9, ids)
    bvh = oracle_lib.CpuBvh(v9, tris)
    rays = _random_rays(rng, 20000, -50, 1050)
    for skip in (-1, int(ids[30])):
        for sky in (False, True):
            a = bvh.visibility(rays, skip_id=skip, sky_passes=sky)
            b = oracle_lib.visibility_brute(tris, rays, skip_id=skip, sky_passes=sky)
            assert np.array_equal(a, b)


def test_swapped_endpoints_same_bit_and_closed_room(oracle_lib):
    """Closed forms: a segment inside one closed box room is never blocked by
    its walls; a segment through a wall is; swapping a segment's endpoints
    does not change whether it is blocked (the exact test has no preferred side)."""
    from hl2sdk_tools_sharp_amd import scenes

    rng = np.random.default_rng(11)
    lo = np.array([[0, 0, 0]], np.float32)
    hi = np.array([[512, 512, 256]], np.float32)
    v9, ids = scenes.box_triangles(lo, hi)
    tris = oracle_lib.convert_triangles(v9, ids)
    inside = _random_rays(rng, 5000, 1.0, 255.0)
    assert not oracle_lib.unpack_bits(oracle_lib.visibility_brute(tris, inside), 5000).any()
    a = rng.uniform(1, 255, size=(2000, 3)).astype(np.float32)
    b = a.copy()
    b[:, 0] += 600.0  # always crosses the x = 512 wall
    through = np.concatenate([a, b - a, np.ones((2000, 1), np.float32)], axis=1)
    assert oracle_lib.unpack_bits(oracle_lib.visibility_brute(tris, through), 2000).all()
    back = np.concatenate([b, a - b, np.ones((2000, 1), np.float32)], axis=1)
    assert oracle_lib.unpack_bits(oracle_lib.visibility_brute(tris, back), 2000).all()


def test_nonfinite_and_degenerate_rays_miss(oracle_lib, kd_scenes):
    scene = kd_scenes["kd-boxes"]
    tris = oracle_lib.convert_triangles(scene.v9)
    r = scene.rays[:4].copy()
    r[0, 3:6] = np.nan
    r[1, 0] = np.inf
    r[2, 3:6] = 0.0  # zero-length segment
    r[3, 6] = np.nan  # NaN reach: nothing is short of it
    assert not oracle_lib.unpack_bits(oracle_lib.visibility_brute(tris, r), 4).any()
