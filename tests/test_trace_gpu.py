"""Parity of the HIP traversal kernels with the oracle.  Needs an MI355X.

The bar is bit-exact: every visibility bit, every hit id and every hit
fraction equal to the oracle's (the oracle is pinned to the reference's
golden vectors in test_oracle_rays.py)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from hl2sdk_tools_sharp_amd import (
    HIT_DTYPE,
    HipTracerFactory,
    NotSupportedError,
    ObjectDisposedError,
    RayTraceOptions,
    ShadowCasterSet,
    _lib,
    scenes,
    unpack_bits,
)

pytestmark = pytest.mark.gpu


def make_tracer(v9, ids=None, **kw):
    v9 = np.ascontiguousarray(v9, np.float32).reshape(-1, 9)
    ids = np.arange(len(v9), dtype=np.int32) if ids is None else ids
    offer = HipTracerFactory(**kw).try_create(ShadowCasterSet(v9, ids))
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer


def gpu_vis(tr, rays, opts=RayTraceOptions()):
    bits = np.zeros((len(rays) + 63) // 64, np.uint64)
    tr.trace_visibility(rays, bits, opts)
    return bits


def gpu_closest(tr, rays, opts=RayTraceOptions()):
    hits = np.zeros(len(rays), HIT_DTYPE)
    tr.trace_closest(rays, hits, opts)
    return hits


def expected_closest(tris, ids, rays, tmin=0.0, skip_id=-1):
    idx, t = oracle.closest_brute(tris, rays, tmin=tmin, skip_id=skip_id, clip=True)
    surf = np.where(idx >= 0, ids[np.maximum(idx, 0)], -1).astype(np.int32)
    reach = rays[:, 6]
    frac = np.where(reach == 1.0, t, (t / reach).astype(np.float32)).astype(np.float32)
    frac = np.where(idx >= 0, frac, np.float32(1.0)).astype(np.float32)
    return surf, frac


@pytest.mark.parametrize("name", ["kd-scene", "kd-boxes"])
def test_kd_fixtures_bit_exact(kd_scenes, name):
    s = kd_scenes[name]
    print(f"[kd fixture {name}] {len(s.v9)} tris, {len(s.rays)} rays", flush=True)
    ids = np.arange(len(s.v9), dtype=np.int32)
    tris = oracle.convert_triangles(s.v9, ids)
    with make_tracer(s.v9, ids) as tr:
        bits = gpu_vis(tr, s.rays)
        assert np.array_equal(unpack_bits(bits, len(s.rays)), s.stock_blocked)
        assert np.array_equal(bits, oracle.visibility_brute(tris, s.rays))
        hits = gpu_closest(tr, s.rays)
        surf, frac = expected_closest(tris, ids, s.rays)
        assert np.array_equal(hits["surface"], surf)
        assert np.array_equal(hits["fraction"].view(np.uint32), frac.view(np.uint32))
        # stock's own answers: same triangle for every ray stock hit short of the end
        short = s.stock_blocked
        assert np.array_equal(hits["surface"][short], s.stock_hit[short])


def _random_segments(rng, n, lo, hi, reach_scale=False):
    a = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    b = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    r = np.empty((n, 7), np.float32)
    r[:, :3] = a
    if reach_scale:  # Ray.Segment parameterisation: unit direction, reach = length
        from hl2sdk_tools_sharp_amd import ray_segment

        return ray_segment(a, b)
    r[:, 3:6] = b - a
    r[:, 6] = 1.0
    return r


@pytest.fixture(scope="module")
def box_case():
    rng = np.random.default_rng(12345)
    lo, hi = scenes.random_boxes(300, rng, volume=2000.0, edge_min=8, edge_max=300)
    v9, ids = scenes.box_triangles(lo, hi)
    ids = ids.copy()
    ids[:36] |= 0x01000000  # three sky boxes
    tris = oracle.convert_triangles(v9, ids)
    bvh = oracle.CpuBvh(v9, tris)
    rays = np.concatenate(
        [_random_segments(rng, 150_000, -100, 2100), _random_segments(rng, 50_000, -100, 2100, True)]
    )
    return v9, ids, tris, bvh, rays


@pytest.mark.parametrize("tmin", [0.0, 1e-3, 0.3])
def test_box_scene_visibility_bit_exact(box_case, tmin):
    v9, ids, tris, bvh, rays = box_case
    print(f"[box scene visibility] tmin={tmin} rays={len(rays)}", flush=True)
    with make_tracer(v9, ids) as tr:
        got = gpu_vis(tr, rays, RayTraceOptions(tmin))
    assert np.array_equal(got, bvh.visibility(rays, tmin=tmin))


def test_box_scene_closest_bit_exact(box_case):
    v9, ids, tris, bvh, rays = box_case
    print("[box scene closest]", flush=True)
    with make_tracer(v9, ids) as tr:
        hits = gpu_closest(tr, rays)
    idx, t = bvh.closest(rays)
    surf = np.where(idx >= 0, ids[np.maximum(idx, 0)], -1)
    reach = rays[:, 6]
    frac = np.where(idx >= 0, np.where(reach == 1.0, t, (t / reach).astype(np.float32)), 1.0).astype(np.float32)
    assert np.array_equal(hits["surface"], surf)
    assert np.array_equal(hits["fraction"].view(np.uint32), frac.view(np.uint32))


@pytest.mark.parametrize("sky", [False, True])
def test_skip_id_and_sky_passes(box_case, sky):
    v9, ids, tris, bvh, rays = box_case
    skip = int(ids[100])
    print(f"[options] skip={skip:#x} sky={sky}", flush=True)
    opts = RayTraceOptions(0.0, skip_id=skip, sky_does_not_block=sky)
    with make_tracer(v9, ids) as tr:
        assert tr.supports(opts)
        got = gpu_vis(tr, rays[:60000], opts)
    assert np.array_equal(got, bvh.visibility(rays[:60000], skip_id=skip, sky_passes=sky))


def test_shared_reach_records(box_case):
    """Six-float records with one reach for the batch (MapTools.Gpu/RayRecord.cs:78-144)."""
    v9, ids, tris, bvh, rays = box_case
    r6 = np.ascontiguousarray(rays[:40000, :6])
    with make_tracer(v9, ids) as tr:
        bits = np.zeros((len(r6) + 63) // 64, np.uint64)
        st = _lib.lib().vh_trace_visibility(
            tr.handle, r6.ctypes.data_as(ctypes.c_void_p), 6, 0.75, len(r6), 0.0, -1, 0,
            bits.ctypes.data_as(ctypes.c_void_p), len(bits),
        )
        assert st == 0, _lib.last_error()
    r7 = rays[:40000].copy()
    r7[:, 6] = 0.75
    assert np.array_equal(bits, bvh.visibility(r7))


@pytest.mark.parametrize("nlights", [64, 37])
def test_point_lights_equal_materialised_rays(nlights):
    w = scenes.visibility_workload(nboxes=400, npoints=3000, nlights=nlights, seed=5, volume=3000.0)
    print(f"[point lights] {w.rays} rays, {nlights} lights", flush=True)
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        pl = tr.trace_point_lights(w.points, w.lights)
        rays = scenes.point_light_rays(w.points, w.lights)
        vis = gpu_vis(tr, rays)
    assert np.array_equal(pl, vis)
    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    bvh = oracle.CpuBvh(w.casters.vertices, tris)
    assert np.array_equal(pl, bvh.point_lights(w.points, w.lights))


def _walk_form_points(w, rng):
    """Points that send a wave's rays into several direction octants (points
    all round each light), rays with an exactly zero component (a shared
    coordinate with the light) and with components under 2^-60, plus the
    workload's own points."""
    parts = [w.points]
    for L in w.lights[:6]:
        parts.append((L + rng.normal(0.0, 40.0, size=(128, 3))).astype(np.float32))
        same = np.repeat(L[None, :], 64, axis=0).astype(np.float32)
        same[:, 1:] += rng.uniform(-300.0, 300.0, size=(64, 2)).astype(np.float32)  # d.x == 0
        parts.append(same)
    pts = np.concatenate(parts).astype(np.float32)
    return pts


def _tiny_component_case(w, rng):
    """A light moved to x = 0 and points at x = +-1e-25 and +-1e-30 round it:
    direction components under 2^-60 that are not zero."""
    L = w.lights[0].copy()
    L[0] = 0.0
    pts = np.repeat(L[None, :], 256, axis=0).astype(np.float32)
    pts[:, 1:] += rng.uniform(-300.0, 300.0, size=(256, 2)).astype(np.float32)
    pts[:, 0] = np.tile(np.array([1e-25, -1e-25, 1e-30, -1e-30], np.float32), 64)
    return pts, np.concatenate([w.lights, L[None, :]]).astype(np.float32)


@pytest.mark.parametrize("form", ["octants", "tmin"])
def test_point_lights_walk_forms(form):
    """The point-light kernel walks the tree per direction octant when a
    wave's rays share one and by the general form otherwise (points all round
    a light); direction components that are zero or under 2^-60 take the
    saturated inverse.  The bits equal the brute-force oracle's on the
    materialised rays, with tmin 0 and with tmin 0.25."""
    rng = np.random.default_rng(17)
    w = scenes.visibility_workload(nboxes=150, npoints=1500, nlights=12, seed=13, volume=1500.0)
    v9 = w.casters.vertices.copy()
    pts = _walk_form_points(w, rng)
    tiny_pts, lights = _tiny_component_case(w, rng)
    pts = np.concatenate([pts, tiny_pts]).astype(np.float32)
    tmin = 0.25 if form == "tmin" else 0.0
    ids = np.arange(len(v9), dtype=np.int32)
    rays = scenes.point_light_rays(pts, lights)
    print(f"[walk forms] {form}: {len(v9)} tris, {len(rays)} rays, tmin {tmin}", flush=True)
    tris = oracle.convert_triangles(v9, ids)
    want = oracle.visibility_brute(tris, rays, tmin=tmin)
    with make_tracer(v9, ids) as tr:
        got = tr.trace_point_lights(pts, lights, RayTraceOptions(tmin))
    assert np.array_equal(got, want)
    frac = unpack_bits(got, len(rays)).mean()
    assert 0.05 < frac < 0.95


def test_point_lights_single_leaf_scene():
    """A scene of one triangle is one leaf under the root, whose other child
    is empty; the octant walks drop the empty-child check, which this covers."""
    v9 = np.array([[0, 0, 0, 10, 0, 0, 0, 10, 0]], np.float32)
    tris = oracle.convert_triangles(v9)
    rng = np.random.default_rng(3)
    pts = rng.uniform([-5, -5, -10], [15, 15, 10], size=(500, 3)).astype(np.float32)
    lights = np.array([[2, 2, 20], [2, 2, -20], [30, 1, 1], [1, 1, 0.5]], np.float32)
    with make_tracer(v9) as tr:
        got = tr.trace_point_lights(pts, lights)
    want = oracle.visibility_brute(tris, scenes.point_light_rays(pts, lights))
    assert np.array_equal(got, want)
    assert unpack_bits(got, len(pts) * len(lights)).any()


def test_edge_cases():
    v9 = np.array([[0, 0, 0, 10, 0, 0, 0, 10, 0]], np.float32)  # one triangle, z = 0
    tris = oracle.convert_triangles(v9)
    with make_tracer(v9) as tr:
        # empty batch
        tr.trace_visibility(np.zeros((0, 7), np.float32), np.zeros(0, np.uint64))
        rays = np.array(
            [
                [1, 1, 5, 0, 0, -10, 1],  # hit at t = 0.5
                [1, 1, 5, 0, 0, -5, 1],  # ends exactly on the plane: not short of the end
                [np.nan, 1, 5, 0, 0, -10, 1],
                [1, 1, 5, np.inf, 0, -10, 1],
                [1, 1, 5, 0, 0, 0, 1],  # zero direction
                [1, 1, 5, 0, 0, -10, np.nan],
                [20, 20, 5, 0, 0, -10, 1],  # outside the triangle
                [1, 1, 5, 1e-30, 0, -10, 1],
                [1, 1, 0, 1, 0, 0, 1],  # in the plane: parallel reject
                [1, 1, -5, 0, 0, 10, 1],  # from below
            ],
            np.float32,
        )
        got = unpack_bits(gpu_vis(tr, rays), len(rays))
        want = unpack_bits(oracle.visibility_brute(tris, rays), len(rays))
        assert np.array_equal(got, want)
        assert got[0] and not got[1] and got[9]
        hits = gpu_closest(tr, rays)
        surf, frac = expected_closest(tris, np.zeros(1, np.int32), rays)
        assert np.array_equal(hits["surface"], surf)
        assert np.array_equal(hits["fraction"].view(np.uint32), frac.view(np.uint32))
        assert hits["surface"][1] == 0 and hits["fraction"][1] == 1.0  # t == reach counts for closest
        with pytest.raises(ValueError):
            tr.trace_visibility(rays, np.zeros(0, np.uint64))
        with pytest.raises(ValueError):
            tr.trace_closest(rays, np.zeros(3, HIT_DTYPE))
        with pytest.raises(ValueError):
            tr.trace_closest(rays, np.zeros(len(rays), HIT_DTYPE), RayTraceOptions(sky_does_not_block=True))
        with pytest.raises(NotSupportedError):
            tr.trace_visibility(rays, np.zeros(1, np.uint64), RayTraceOptions(float("nan")))
    with pytest.raises(ObjectDisposedError):
        tr.trace_visibility(rays, np.zeros(1, np.uint64))
    tr.dispose()  # a second dispose is harmless


def test_determinism_and_slab_independence(box_case):
    v9, ids, tris, bvh, rays = box_case
    with make_tracer(v9, ids) as tr:
        a = gpu_vis(tr, rays)
        b = gpu_vis(tr, rays)
        half = (len(rays) // 2) // 64 * 64
        c = np.concatenate([gpu_vis(tr, rays[:half]), gpu_vis(tr, rays[half:])])
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_factory_fault_hooks_decline():
    v9 = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    for stage in (1, 2, 3):
        offer = HipTracerFactory(fail_stage=stage).try_create(ShadowCasterSet(v9, np.zeros(1, np.int32)))
        assert offer.tracer is None and f"stage {stage}" in offer.decline_reason
    tr = make_tracer(v9)
    assert tr.tracer_identity.startswith("hip-bvh2-exact-1:gfx950")
    tr.dispose()


def test_visibility_config_sample_bit_exact():
    """The 24k-triangle box scene of the visibility config (seed 2) with a
    20,000-point sample of its points against all 64 lights (1.28M rays)."""
    w = scenes.visibility_workload(npoints=20_000)
    print(f"[visibility config sample] {w.casters.count} tris, {w.rays} rays", flush=True)
    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    bvh = oracle.CpuBvh(w.casters.vertices, tris)
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        got = tr.trace_point_lights(w.points, w.lights)
    want = bvh.point_lights(w.points, w.lights)
    assert np.array_equal(got, want)
    frac = unpack_bits(got, w.rays).mean()
    assert 0.05 < frac < 0.95  # a scene that both blocks and passes
