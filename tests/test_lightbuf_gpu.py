"""The light-buffer form of the point-light trace against the BVH walk and
the oracle.  Needs an MI355X.

Both forms answer with the exact triangle test (KdRayTracer.IntersectTriangle,
Tracing/KdRayTracer.cs:1194-1267) over a superset of the triangles a segment
can hit, so their bits must be equal to each other and to the brute-force
oracle's, bit for bit.  The cases aim at the light buffer's own geometry:
lights inside, on the faces, edges and corners of triangle boxes, points that
share one, two or three coordinates with a light (directions on cell and face
boundaries, zero directions), coarse and odd cell counts, minimum distances
above zero, light counts that are not whole words and span several 64-light
groups, and a scene replaced between calls."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from hl2sdk_tools_sharp_amd import HipTracerFactory, RayTraceOptions, ShadowCasterSet, _lib, scenes, unpack_bits

pytestmark = pytest.mark.gpu

BVH, LB = _lib.VH_POINT_LIGHTS_BVH, _lib.VH_POINT_LIGHTS_LIGHT_BUFFER


def make_tracer(v9, ids=None):
    v9 = np.ascontiguousarray(v9, np.float32).reshape(-1, 9)
    ids = np.arange(len(v9), dtype=np.int32) if ids is None else ids
    offer = HipTracerFactory().try_create(ShadowCasterSet(v9, ids))
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer


def both(tr, pts, lights, tmin=0.0, cells=0):
    tr.set_point_light_method(BVH)
    a = tr.trace_point_lights(pts, lights, RayTraceOptions(tmin))
    tr.set_point_light_method(LB, cells)
    b = tr.trace_point_lights(pts, lights, RayTraceOptions(tmin))
    return a, b


def adversarial(w, rng, nlights=40):
    """Lights inside triangle boxes and on their faces, edges and corners, at
    vertices; points that share 1-3 coordinates with a light."""
    v9 = w.casters.vertices
    lo = np.minimum(np.minimum(v9[:, 0:3], v9[:, 3:6]), v9[:, 6:9])
    hi = np.maximum(np.maximum(v9[:, 0:3], v9[:, 3:6]), v9[:, 6:9])
    k = rng.integers(0, len(v9), size=nlights)
    f = rng.uniform(0.0, 1.0, size=(nlights, 3))
    L = lo[k] + f * (hi[k] - lo[k])
    L[1::4, 0] = lo[k[1::4], 0]  # on a face plane
    L[2::4, :2] = np.stack([lo[k[2::4], 0], hi[k[2::4], 1]], axis=1)  # on an edge
    L[3::4] = v9[k[3::4], 0:3]  # at a vertex
    lights = np.concatenate([w.lights, L.astype(np.float32)]).astype(np.float32)
    extra = []
    for j in range(0, len(lights), 3):
        q = w.points[rng.integers(0, len(w.points), size=12)].copy()
        q[0:3, 0] = lights[j, 0]
        q[3:6, :2] = lights[j, :2]
        q[6:8] = lights[j]
        q[8:12, 1] = lights[j, 1] + (q[8:12, 2] - lights[j, 2])  # on a 45-degree face boundary
        extra.append(q)
    pts = np.concatenate([w.points] + extra).astype(np.float32)
    return pts, lights


@pytest.mark.parametrize("cells,tmin", [(0, 0.0), (5, 0.0), (64, 0.25), (120, 1e-3)])
def test_light_buffer_equals_walk_and_oracle(cells, tmin):
    rng = np.random.default_rng(23 + cells)
    w = scenes.visibility_workload(nboxes=300, npoints=2500, nlights=24, seed=29, volume=2500.0)
    pts, lights = adversarial(w, rng)
    rays = scenes.point_light_rays(pts, lights)
    print(f"[light buffer] cells {cells} tmin {tmin}: {w.casters.count} tris, {len(rays)} rays", flush=True)
    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    want = oracle.visibility_brute(tris, rays, tmin=tmin)
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        a, b = both(tr, pts, lights, tmin, cells)
    assert np.array_equal(a, want)
    assert np.array_equal(b, want)
    frac = unpack_bits(b, len(rays)).mean()
    assert 0.05 < frac < 0.95


@pytest.mark.parametrize("nlights", [37, 64, 130])
def test_light_buffer_light_groups(nlights):
    """Light counts that are not whole words (bits shared between points) and
    that span two and three 64-light groups."""
    w = scenes.visibility_workload(nboxes=200, npoints=1999, nlights=nlights, seed=31, volume=2000.0)
    print(f"[light buffer] {nlights} lights, {w.rays} rays", flush=True)
    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    bvh = oracle.CpuBvh(w.casters.vertices, tris)
    want = bvh.point_lights(w.points, w.lights)
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        a, b = both(tr, w.points, w.lights)
    assert np.array_equal(a, want) and np.array_equal(b, want)


def test_light_buffer_single_triangle_and_reload():
    """A one-triangle scene (one leaf under the root), then a scene replaced on
    the same context: the triangle boxes are rebuilt for the new scene."""
    v9 = np.array([[0, 0, 0, 10, 0, 0, 0, 10, 0]], np.float32)
    rng = np.random.default_rng(3)
    pts = rng.uniform([-5, -5, -10], [15, 15, 10], size=(500, 3)).astype(np.float32)
    lights = np.array([[2, 2, 20], [2, 2, -20], [30, 1, 1], [1, 1, 0.5], [2, 2, 0]], np.float32)
    w = scenes.visibility_workload(nboxes=100, npoints=800, nlights=9, seed=37, volume=1000.0)
    with make_tracer(v9) as tr:
        tr.set_point_light_method(LB)
        got1 = tr.trace_point_lights(pts, lights)
        tr.load_scene(w.casters)
        got2 = tr.trace_point_lights(w.points, w.lights)
    want1 = oracle.visibility_brute(oracle.convert_triangles(v9), scenes.point_light_rays(pts, lights))
    assert np.array_equal(got1, want1) and unpack_bits(got1, len(pts) * len(lights)).any()
    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    assert np.array_equal(got2, oracle.CpuBvh(w.casters.vertices, tris).point_lights(w.points, w.lights))


def test_light_buffer_nonfinite_and_degenerate():
    """Non-finite points and lights give unblocked rays, as in the walk; a
    point on its light has a zero direction and is never blocked."""
    w = scenes.visibility_workload(nboxes=80, npoints=600, nlights=6, seed=41, volume=800.0)
    pts = w.points.copy()
    pts[:5, 0] = np.nan
    pts[5:9, 1] = np.inf
    pts[9] = w.lights[2]
    lights = np.concatenate([w.lights, np.array([[np.inf, 0, 0], [0, np.nan, 0]], np.float32)])
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        a, b = both(tr, pts, lights)
    assert np.array_equal(a, b)
    bits = unpack_bits(b, len(pts) * len(lights)).reshape(len(pts), len(lights))
    assert not bits[:9].any() and not bits[:, 6:].any() and not bits[9, 2]
    want = oracle.visibility_brute(oracle.convert_triangles(w.casters.vertices, w.casters.ids),
                                   scenes.point_light_rays(pts, lights))
    assert np.array_equal(b, want)


def test_light_buffer_counters():
    """The counters of a light-buffer trace: every ray reads at least the
    entries it tests, and each wave's steps cover its lanes' entries."""
    w = scenes.visibility_workload(nboxes=300, npoints=4096, nlights=64, seed=43, volume=3000.0)
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        tr.set_point_light_method(LB)
        h = tr.handle
        assert _lib.lib().vh_count_enable(h, 1) == 0
        tr.trace_point_lights(w.points, w.lights)
        import ctypes
        v, t = ctypes.c_int64(), ctypes.c_int64()
        s0, s1 = ctypes.c_int64(), ctypes.c_int64()
        assert _lib.lib().vh_count_read(h, ctypes.byref(v), ctypes.byref(t)) == 0
        assert _lib.lib().vh_count_read_slots(h, ctypes.byref(s0), ctypes.byref(s1)) == 0
        _lib.lib().vh_count_enable(h, 0)
    assert 0 < t.value <= v.value <= s0.value
    assert s1.value == 0


def test_light_buffer_visibility_config_full():
    """The whole visibility config (2,000 boxes, 2M points x 64 lights =
    128M rays): the light buffer's 16 MB of hit words equal the BVH walk's
    (both exact; the walk is checked against the oracle on a sample in
    test_trace_gpu), and the auto method picks the light buffer here."""
    w = scenes.visibility_workload()
    print(f"[light buffer] visibility config: {w.rays} rays", flush=True)
    with make_tracer(w.casters.vertices, w.casters.ids) as tr:
        a, b = both(tr, w.points, w.lights)
        tr.set_point_light_method(_lib.VH_POINT_LIGHTS_AUTO)
        h = tr.handle
        _lib.lib().vh_timing_enable(h, 1)
        c = tr.trace_point_lights(w.points, w.lights)
        import ctypes
        ms, n = ctypes.c_double(), ctypes.c_int64()
        assert _lib.lib().vh_timing_read(h, _lib.VH_KERNEL_LIGHT_BUFFER, ctypes.byref(ms), ctypes.byref(n)) == 0
        _lib.lib().vh_timing_enable(h, 0)
    assert np.array_equal(a, b) and np.array_equal(b, c)
    assert n.value == 1  # one 64-light group, built once
    frac = unpack_bits(b, w.rays).mean()
    assert 0.5 < frac < 0.95


@pytest.mark.parametrize("nx,tile,cells", [(3, 0.0, 0), (5, 128.0, 0), (4, 256.0, 9)])
def test_light_buffer_room_grid(nx, tile, cells):
    """Closed rooms joined by doorways, the lights inside the rooms and the
    points one unit off the walls (the bounce map's direct-light call, at a
    size the oracle finishes in seconds): most rays end on a wall, every
    triangle lies in an axis plane (many in a light's own), and tiled faces
    give many small triangles near each light."""
    from hl2sdk_tools_sharp_amd import rooms
    if tile:
        g = rooms.room_grid(nx, nx, size=rooms.BOUNCE_ROOM, door=rooms.BOUNCE_DOOR, patch=tile,
                            minchop=tile, tile=tile, nlights=2 * nx * nx, seed=7)
    else:
        g = rooms.room_grid(nx, nx, nlights=12, seed=5)
    pts = rooms.sample_points(g)
    print(f"[light buffer] room grid {nx}x{nx} tile {tile}: {g.casters.count} tris, "
          f"{len(pts)} points x {len(g.lights)} lights", flush=True)
    tris = oracle.convert_triangles(g.casters.vertices, g.casters.ids)
    want = oracle.CpuBvh(g.casters.vertices, tris).point_lights(pts, g.lights)
    with make_tracer(g.casters.vertices, g.casters.ids) as tr:
        a, b = both(tr, pts, g.lights, 0.0, cells)
    assert np.array_equal(a, want)
    assert np.array_equal(b, want)
    frac = unpack_bits(b, len(pts) * len(g.lights)).mean()
    assert 0.3 < frac < 0.999
