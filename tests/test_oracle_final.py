"""The final-light oracle (oracle/final.py) pinned against the closed forms
of the reference's own tests for LuxelRadial and FinalLightFace
(SourceSharp.Tests/MapTools/Rad/Final/LuxelRadialTests.cs,
FinalLightFaceTests.cs) and what the arithmetic fixes: the weights of a
sample or patch at known distances, the double reciprocal in Sample, the
minlight clamp, the median rule, the lump layout, exact scaling by powers of
two and the normalisation of a uniform field."""
import numpy as np
import pytest

from oracle import final as fo

f32 = np.float32


def floor_face(**kw):
    """The reference tests' floor: a 17 x 17 grid of 16-unit luxels at the
    world origin (LightBox's face 0)."""
    args = dict(origin=(0, 0, 0), w2l_s=(1 / 16, 0, 0), w2l_t=(0, 1 / 16, 0), l2w_s=(16, 0, 0),
                l2w_t=(0, 16, 0), min_s=0, min_t=0, width=17, height=17)
    args.update(kw)
    return fo.Face(**args)


def radial(**kw):
    return fo.Radial(floor_face(**kw))


def L(r, s, t):
    return r.face.luxel_to_world(s, t)


def one(v):
    return [np.full(3, v, np.float32)]


def test_luxel_frame_round_trips_on_the_floor():
    f = floor_face(min_s=3, min_t=-2)
    for s, t in [(0, 0), (4.5, 4.5), (16, 3.25)]:
        assert f.world_to_luxel(f.luxel_to_world(s, t)) == (f32(s), f32(t))


def test_a_whole_sample_half_a_luxel_away_weighs_two():
    r = radial()
    r.add_direct(L(r, 4.5, 4.5), 4, 4, 5, 5, one(10), False, False)
    assert r.weight[4 + 4 * 17] == 2.0
    assert r.weight[5 + 5 * 17] == 2.0
    assert r.light[0][4 + 4 * 17][0] == 20.0


def test_a_sample_on_a_luxel_centre_is_area_over_a_tenth_in_double():
    r = radial()
    r.add_direct(L(r, 4, 4), 3.5, 3.5, 4.5, 4.5, one(1), False, False)
    assert r.weight[4 + 4 * 17] == f32(1.0 / 0.1)
    # the neighbouring luxel: overlap 0.5 x 1 at distance 1
    assert r.weight[5 + 4 * 17] == 0.5


def test_an_overlap_under_equal_epsilon_gets_nothing():
    r = radial()
    r.add_direct(L(r, 4, 4.5), 4, 4, 4.0005, 5, one(10), False, False)
    assert r.weight[4 + 4 * 17] == 0.0
    assert r.weight[5 + 4 * 17] == 0.0


def test_a_sample_hanging_off_the_grid_only_touches_luxels_on_it():
    r = radial()
    r.add_direct(L(r, 16.5, 16.5), 16, 16, 17, 17, one(10), False, False)
    assert r.weight[16 + 16 * 17] == 2.0
    assert sum(1 for w in r.weight if w != 0) == 1


def test_an_unbumped_sample_on_a_bumped_face_feeds_each_direction_at_one_over_root_three():
    r = radial()
    r.add_direct(L(r, 4.5, 4.5), 4, 4, 5, 5, one(3), True, False)
    i = 4 + 4 * 17
    assert r.light[0][i][0] == f32(3) * f32(2)
    assert r.light[2][i][0] == f32(3) * (f32(2) * fo.ONE_OVER_SQRT3)


def test_a_bumped_sample_on_a_bumped_face_feeds_each_direction_its_own_light():
    r = radial()
    light = [np.full(3, b + 1, np.float32) for b in range(4)]
    r.add_direct(L(r, 4.5, 4.5), 4, 4, 5, 5, light, True, True)
    i = 4 + 4 * 17
    assert [r.light[b][i][0] for b in range(4)] == [2.0, 4.0, 6.0, 8.0]


def test_a_bounced_patch_weighs_two_minus_its_squared_distance():
    r = radial()
    r.add_bounced(L(r, 4.5, 4.5), 4, 4, 5, 5, [np.ones(3, np.float32)], False, False)
    assert r.weight[4 + 4 * 17] == 1.5
    # (4.5 - 3) in patch units: 2 - (2.25 + 0.25) < 0, so nothing
    assert r.weight[3 + 4 * 17] == 0.0
    # a two-luxel patch halves the distances: 2 - (0.5625 + 0.0625)
    r2 = radial()
    r2.add_bounced(L(r2, 4.5, 4.5), 4, 4, 6, 6, [np.ones(3, np.float32)], False, False)
    assert r2.weight[3 + 4 * 17] == 2 - (0.5625 + 0.0625)


def test_a_bounced_patch_smaller_than_a_luxel_is_filtered_as_one():
    small, whole = radial(), radial()
    small.add_bounced(L(small, 4.5, 4.5), 4.4, 4.4, 4.6, 4.6, [np.ones(3, np.float32)], False, False)
    whole.add_bounced(L(whole, 4.5, 4.5), 4, 4, 5, 5, [np.ones(3, np.float32)], False, False)
    assert small.weight == whole.weight


def test_sample_divides_by_the_weight_as_a_double_reciprocal():
    r = radial()
    r.add_direct(L(r, 4.5, 4.5), 4, 4, 5, 5, one(7), False, False)
    r.add_direct(L(r, 3.5, 3.5), 3, 3, 4, 4, one(3), False, False)
    out, ok = r.sample(L(r, 4, 4), 1, False)
    i = 4 + 4 * 17
    assert ok
    w = r.weight[i]
    assert out[0][0] == r.light[0][i][0] * f32(1.0 / float(w))
    # 7 and 3 at weight 2 each: the mean
    assert out[0][0] == 5.0


def test_a_luxel_nothing_reached_is_black_and_not_a_base_sample():
    r = radial()
    out, ok = r.sample(L(r, 8, 8), 1, False)
    assert not ok and list(out[0]) == [0, 0, 0]


def test_red_errors_paint_a_luxel_nothing_reached_red():
    r = radial()
    out, ok = r.sample(L(r, 8, 8), 1, True)
    assert not ok and tuple(out[0]) == fo.ERROR_RED


def test_a_point_off_the_grid_is_red_whatever_red_errors_says():
    r = radial()
    out, ok = r.sample(L(r, 40, 8), 1, False)
    assert not ok and tuple(out[0]) == fo.ERROR_RED
    # one past the last column is off the grid too
    out, ok = r.sample(L(r, 17, 3), 1, False)
    assert not ok and tuple(out[0]) == fo.ERROR_RED
    # -0.4 rounds (truncating -0.4 + 0.5) onto column 0
    r.add_direct(L(r, 0, 3), -0.5, 2.5, 0.5, 3.5, one(4), False, False)
    out, ok = r.sample(L(r, -0.4, 3), 1, False)
    assert ok and out[0][0] == 4.0


def test_the_median_is_element_count_over_two_of_the_sorted_values():
    assert fo.median([f32(4), f32(1), f32(3), f32(2)]) == 3.0
    assert fo.median([f32(3), f32(1), f32(2)]) == 2.0


def _grid_samples(face, light):
    """One sample per luxel at the luxel, bounds the luxel's square, as
    vrad's sample builder makes them for a face that fills its grid."""
    pos, bounds = [], []
    for t in range(face.h):
        for s in range(face.w):
            pos.append(face.luxel_to_world(s, t))
            bounds.append((s - 0.5, t - 0.5, s + 0.5, t + 0.5))
    n = len(pos)
    lt = np.broadcast_to(np.asarray(light, np.float32), (face.nb, n, 3)).copy()
    return {"pos": np.array(pos, np.float32), "bounds": np.array(bounds, np.float32), "light": lt}


def _luxels(face):
    return np.array([face.luxel_to_world(s, t) for t in range(face.h) for s in range(face.w)],
                    np.float32)


def test_minlight_raises_a_dark_face_to_128_and_its_median_is_128():
    face = floor_face(width=5, height=4, minlight=128.0)
    res = fo.final_light([face], [_luxels(face)], [_grid_samples(face, [0, 0, 0])], [[]])
    assert np.all(res[0]["luxels"] == 128.0)
    lump = fo.lump_bytes([face], res)
    assert lump[:4] == bytes([128, 128, 128, 0])
    assert lump[4:8] == bytes([128, 128, 128, 0])
    assert len(lump) == 4 + 4 * 20


def test_a_uniform_field_comes_back_uniform():
    face = floor_face(width=6, height=5)
    c = np.array([3.0, 0.5, 7.25], np.float32)
    res = fo.final_light([face], [_luxels(face)], [_grid_samples(face, c)], [[]])
    np.testing.assert_allclose(res[0]["luxels"][0], np.broadcast_to(c, (30, 3)), rtol=1e-6)
    np.testing.assert_allclose(res[0]["median"], c, rtol=1e-6)


def test_scaling_the_light_by_a_power_of_two_scales_the_result_exactly():
    rng = np.random.default_rng(3)
    face = floor_face(width=7, height=6, bumped=True)
    sm = _grid_samples(face, [0, 0, 0])
    sm["light"] = rng.uniform(0, 50, sm["light"].shape).astype(np.float32)
    sm2 = dict(sm, light=sm["light"] * 4)
    a = fo.final_light([face], [_luxels(face)], [sm], [[]])[0]
    b = fo.final_light([face], [_luxels(face)], [sm2], [[]])[0]
    assert np.array_equal(a["luxels"] * 4, b["luxels"])
    assert np.array_equal(a["median"] * 4, b["median"])


def test_a_face_with_no_samples_is_red_and_has_no_median():
    face = floor_face(width=3, height=3)
    empty = {"pos": np.zeros((0, 3), np.float32), "bounds": np.zeros((0, 4), np.float32),
             "light": np.zeros((1, 0, 3), np.float32)}
    res = fo.final_light([face], [_luxels(face)], [empty], [[]])
    assert np.all(res[0]["luxels"][0] == np.array([255, 0, 0], np.float32))
    assert np.all(res[0]["median"] == 0)


def test_a_neighbours_sample_is_carried_into_this_grid_by_its_corners():
    # B sits right of A: B's luxel s = 0 is A's s = 16
    a = floor_face()
    b = floor_face(min_s=16)
    la, lb_ = _luxels(a), _luxels(b)
    sa = _grid_samples(a, [0, 0, 0])
    sb = {"pos": np.array([b.luxel_to_world(0.5, 3)], np.float32),
          "bounds": np.array([(0, 2.5, 1, 3.5)], np.float32),
          "light": np.full((1, 1, 3), 8, np.float32)}
    r = fo.Radial(a)
    for k in range(len(sa["pos"])):
        r.add_direct(sa["pos"][k], *sa["bounds"][k], [sa["light"][0][k]], False, False)
    w_before = r.weight[16 + 3 * 17]
    res = fo.final_light([a, b], [la, lb_], [sa, sb], [[1], []])
    # in A's grid B's sample is at s = 16.5 with bounds [16, 17]: weight 1 / 0.5
    # on luxel (16, 3), on top of A's own samples there: 10 from the one on
    # it, 0.5 from each of the three edge neighbours, 0.25 from each of the
    # two diagonal ones (the grid ends at column 16)
    assert w_before == 12.0
    expect = f32(f32(8) * f32(2)) * f32(1.0 / float(f32(w_before + f32(2))))
    assert res[0]["luxels"][0, 16 + 3 * 17, 0] == expect


def test_face_patch_lists_walk_the_highest_index_first_and_skip_parents():
    face = [0, 0, 1, 0, 0]
    leaf = [False, True, True, True, True]
    assert fo.face_patch_lists(2, face, leaf) == [[4, 3, 1], [2]]


def test_the_patch_radial_adds_a_patch_total_to_the_luxels_it_covers():
    face = floor_face(width=9, height=9)
    sm = _grid_samples(face, [1, 1, 1])
    corners = [face.luxel_to_world(s, t) for s, t in [(2, 2), (4, 2), (4, 4), (2, 4)]]
    patches = {"origin": np.array([face.luxel_to_world(3, 3)], np.float32), "winding": [np.array(corners)],
               "face": np.array([0]), "leaf": np.array([True]), "total": np.full((1, 4, 3), 6, np.float32)}
    res = fo.final_light([face], [_luxels(face)], [sm], [[]], patches)[0]
    # every luxel the patch reaches reads its weighted mean, 6, on top of 1
    j = 3 + 3 * 9
    assert res["luxels"][0, j, 0] == 7.0
    # (0, 0) is (3/2)^2 * 2 = 4.5 patch units away: unreached, so only the direct light
    assert res["luxels"][0, 0, 0] == 1.0


def test_layout_reserves_the_average_word_before_each_lit_face():
    faces = [floor_face(width=2, height=3), floor_face(width=2, height=2, lit=False),
             floor_face(width=1, height=1, bumped=True)]
    offs, size = fo.layout(faces)
    assert offs == [4, -1, 4 + 24 + 4]
    assert size == 4 + 24 + 4 + 16


@pytest.mark.parametrize("seed,red,bounce,grid", [(1, False, False, False), (2, True, True, False),
                                                  (3, False, True, True), (5, True, False, True)])
def test_the_c_statement_equals_the_python_one(seed, red, bounce, grid):
    """vradcpu_final.c, used for whole maps, against this module's statement
    on seeded random faces: every lump byte (grid: the luxels are the frames'
    grid points)."""
    from final_cases import oracle_inputs, random_case

    _, _, fset, ff, fp = random_case(seed)
    if grid:
        ff.luxel_pos = None
    rng = np.random.default_rng(seed + 10)
    light = rng.uniform(0, 8, (4, fset.nsamples, 4)).astype(np.float32)
    total = rng.uniform(0, 30, (fp.npatch, 4, 3)).astype(np.float32) if bounce else None
    faces, lux, sm, nb, pt = oracle_inputs(ff, fset, light, fp if bounce else None, total)
    want = fo.lump_bytes(faces, fo.final_light(faces, lux, sm, nb, pt, red))
    pd = {"origin": fp.origin, "face": fp.face, "child1": fp.child1, "winding_ptr": fp.winding_ptr,
          "winding": fp.winding}
    got, offs = fo.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                                 fset.face_sample_ptr, fset.face_normal_count, fset.sample_pos, light,
                                 pd if bounce else None, total, red, threads=3)
    assert list(offs) == fo.layout(faces)[0]
    assert got.tobytes() == want


# --- light styles: Run's loop over a face's styles (FinalLightFace.cs:114-125,155,182,270-279) ---

def _styled(face, lights):
    """Grid samples whose light per slot is uniform ``lights[k]``."""
    sm = _grid_samples(face, lights[0])
    n = len(sm["pos"])
    sm["light"] = [np.broadcast_to(np.asarray(c, np.float32), (face.nb, n, 3)).copy() for c in lights]
    return sm


def test_each_style_is_its_own_map_and_the_medians_sit_before_lightofs_in_reverse_order():
    from oracle.gather import encode_rgbe

    face = floor_face(width=4, height=3)
    c0, c1 = [2.0, 4.0, 8.0], [1.0, 0.5, 0.25]
    styles = np.array([[0, 5, 255, 255]], np.uint8)
    res = fo.final_light([face], [_luxels(face)], [_styled(face, [c0, c1])], [[]], styles=styles)
    assert res[0]["luxels"].shape == (2, 1, 12, 3)
    assert np.all(res[0]["luxels"][0] == np.float32(c0)) and np.all(res[0]["luxels"][1] == np.float32(c1))
    offs, size = fo.layout([face], styles)
    assert offs == [8] and size == 8 + 2 * 4 * 12
    lump = fo.lump_bytes([face], res)
    e0, e1 = encode_rgbe(np.array([c0], np.float32)).tobytes(), encode_rgbe(np.array([c1], np.float32)).tobytes()
    assert lump[0:4] == e1 and lump[4:8] == e0  # lightofs - 8: style 1, lightofs - 4: style 0
    assert lump[8:8 + 48] == e0 * 12 and lump[8 + 48:] == e1 * 12


def test_a_neighbours_samples_are_matched_by_style_number_not_slot():
    """A (styles 0, 5) beside B (styles 0, 7, 5): A's style-5 map takes B's
    slot 2, so a uniform style-5 light stays uniform across the seam; B's
    slot 1 (style 7, ten times brighter) never reaches it."""
    a, b = floor_face(width=4, height=4), floor_face(width=4, height=4, min_s=4)
    v = [1.0, 2.0, 3.0]
    sa = _styled(a, [[0, 0, 0], v])
    sb = _styled(b, [[0, 0, 0], [10.0, 20.0, 30.0], v])
    styles = np.array([[0, 5, 255, 255], [0, 7, 5, 255]], np.uint8)
    res = fo.final_light([a, b], [_luxels(a), _luxels(b)], [sa, sb], [[1], [0]], styles=styles)
    np.testing.assert_allclose(res[0]["luxels"][1, 0], np.broadcast_to(np.float32(v), (16, 3)), rtol=1e-6)
    np.testing.assert_allclose(res[1]["luxels"][2, 0], np.broadcast_to(np.float32(v), (16, 3)), rtol=1e-6)
    # matched by slot instead, A's seam column would take B's style-7 light
    wrong = fo.final_light([a, b], [_luxels(a), _luxels(b)], [sa, dict(sb, light=[sb["light"][0], sb["light"][1]])],
                           [[1], [0]], styles=np.array([[0, 5, 255, 255], [0, 5, 255, 255]], np.uint8))
    assert np.any(wrong[0]["luxels"][1, 0, :, 0] > 1.5)


def test_bounced_light_goes_into_style_zero_only():
    face = floor_face(width=9, height=9)
    sm = _styled(face, [[1, 1, 1], [2, 2, 2]])
    corners = [face.luxel_to_world(s, t) for s, t in [(2, 2), (4, 2), (4, 4), (2, 4)]]
    patches = {"origin": np.array([face.luxel_to_world(3, 3)], np.float32), "winding": [np.array(corners)],
               "face": np.array([0]), "leaf": np.array([True]), "total": np.full((1, 4, 3), 6, np.float32)}
    res = fo.final_light([face], [_luxels(face)], [sm], [[]], patches, styles=np.array([[0, 4, 255, 255]], np.uint8))
    j = 3 + 3 * 9
    assert res[0]["luxels"][0, 0, j, 0] == 7.0 and res[0]["luxels"][1, 0, j, 0] == 2.0


@pytest.mark.parametrize("seed,bounce", [(7, True), (8, False)])
def test_the_c_statement_equals_the_python_one_with_styles(seed, bounce):
    from final_cases import oracle_inputs, random_case, random_styles

    _, _, fset, ff, fp = random_case(seed)
    styles = random_styles(fset.nfaces, seed)
    rng = np.random.default_rng(seed + 20)
    light = rng.uniform(0, 8, (4, 4, fset.nsamples, 4)).astype(np.float32)
    total = rng.uniform(0, 30, (fp.npatch, 4, 3)).astype(np.float32) if bounce else None
    faces, lux, sm, nb, pt = oracle_inputs(ff, fset, light, fp if bounce else None, total, face_styles=styles)
    want = fo.lump_bytes(faces, fo.final_light(faces, lux, sm, nb, pt, styles=styles))
    pd = {"origin": fp.origin, "face": fp.face, "child1": fp.child1, "winding_ptr": fp.winding_ptr,
          "winding": fp.winding}
    got, offs = fo.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                                 fset.face_sample_ptr, fset.face_normal_count, fset.sample_pos, light,
                                 pd if bounce else None, total, threads=3, face_styles=styles)
    assert list(offs) == fo.layout(faces, styles)[0]
    assert got.tobytes() == want
