"""Host-side logic of the seam mirror that needs no device."""
from __future__ import annotations

import numpy as np

from hl2sdk_tools_sharp_amd import RayTraceOptions, ray_segment, unpack_bits
from hl2sdk_tools_sharp_amd import scenes


def test_options_mirror_reference_presets():
    """RayTraceOptions.StockExact / SelfIntersectionSafe / TestLine (Tracing/RayTraceOptions.cs:127-144)."""
    assert RayTraceOptions.stock_exact().min_distance == 0.0
    assert RayTraceOptions.self_intersection_safe().min_distance == np.float32(1e-3)
    tl = RayTraceOptions.test_line()
    assert tl.skip_id is None and tl.isolated_rays and tl.is_plain
    tl2 = RayTraceOptions.test_line(5, True)
    assert tl2.skip_id == 5 and tl2.sky_does_not_block and not tl2.is_plain


def test_ray_segment_matches_reference_arithmetic():
    """Ray.Segment, exact reciprocal: dir = d * (1/len), reach = len (Tracing/IRayTracer.cs:71-88)."""
    a = np.array([[1.0, 2.0, 3.0]], np.float32)
    b = np.array([[4.0, 6.0, 3.0]], np.float32)
    r = ray_segment(a, b)
    assert r[0, 6] == 5.0
    assert r[0, 3] == np.float32(3.0) * (np.float32(1.0) / np.float32(5.0))
    z = ray_segment(a, a)  # zero length: NaN direction, hits nothing
    assert np.isnan(z[0, 3:6]).all()


def test_unpack_bits_lsb_first():
    w = np.array([0b1011, 1 << 63], dtype=np.uint64)
    b = unpack_bits(w, 128)
    assert b[0] and b[1] and not b[2] and b[3] and b[127] and b.sum() == 4


def test_box_scene_shape_and_ids():
    rng = np.random.default_rng(0)
    lo, hi = scenes.random_boxes(10, rng)
    v9, ids = scenes.box_triangles(lo, hi)
    assert v9.shape == (120, 9) and ids.shape == (120,)
    assert np.all(ids[0::2] == ids[1::2])  # one id per quad
    assert np.all((ids & 0x02000000) != 0)
    # every triangle has area and lies on its box
    v = v9.reshape(-1, 3, 3)
    cross = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    assert np.all(np.linalg.norm(cross, axis=1) > 0)


def test_points_sit_one_unit_off_a_face():
    rng = np.random.default_rng(1)
    lo, hi = scenes.random_boxes(5, rng)
    pts = scenes.points_on_boxes(lo, hi, 1000, rng)
    # each point is exactly 1 unit outside its box along one axis: distance to the box is ~1
    d = np.maximum(np.maximum(lo[None] - pts[:, None], 0), np.maximum(pts[:, None] - hi[None], 0))
    dist = np.linalg.norm(d, axis=2).min(axis=1)
    np.testing.assert_allclose(dist, 1.0, atol=1e-3)


def test_point_light_rays_layout():
    p = np.array([[0, 0, 0], [1, 1, 1]], np.float32)
    l = np.array([[10, 0, 0], [0, 10, 0], [0, 0, 10]], np.float32)
    r = scenes.point_light_rays(p, l)
    assert r.shape == (6, 7)
    np.testing.assert_array_equal(r[4, :3], [1, 1, 1])
    np.testing.assert_array_equal(r[4, 3:6], [-1, 9, -1])
    assert np.all(r[:, 6] == 1.0)
