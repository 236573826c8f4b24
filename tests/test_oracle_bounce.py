"""The oracle's transfer build and bounce against closed forms, the
reference's own known answers, and a dense linear solve.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.bounce as ob
from hl2sdk_tools_sharp_amd import rooms


# --- FormFactors (src/SourceSharp.Tests/MapTools/Rad/Bounce/FormFactorsTests.cs:24-110) ---

def test_facing_patches_give_the_inverse_square():
    assert ob.diff_to_diff([0, 0, 10], [0, 0, -1], [0, 0, 0], [0, 0, 1]) == pytest.approx(0.01, abs=1e-6)


def test_one_facing_away_is_negative_both_away_positive():
    assert ob.diff_to_diff([0, 0, 10], [0, 0, 1], [0, 0, 0], [0, 0, 1]) < 0
    assert ob.diff_to_diff([0, 0, 10], [0, 0, 1], [0, 0, 0], [0, 0, -1]) == pytest.approx(0.01, abs=1e-6)


def test_square_overhead_matches_the_analytic_form_factor():
    """4 x 0.13853 = 0.55412 for a 2x2 square 1 above, returned times pi / area."""
    sq = [[1, -1, 1], [1, 1, 1], [-1, 1, 1], [-1, -1, 1]]
    r = ob.poly_to_diff(sq, 4.0, [0, 0, 0], [0, 0, 1])
    assert r * 4.0 / np.pi == pytest.approx(0.55412, abs=1e-4)
    rev = [[-1, -1, 1], [-1, 1, 1], [1, 1, 1], [1, -1, 1]]
    assert ob.poly_to_diff(rev, 4.0, [0, 0, 0], [0, 0, 1]) < 0


def test_poly_to_diff_converges_to_diff_to_diff_far_away():
    """A small square far away: the polygon form equals the differential one
    (both divided by the source area)."""
    h = 0.5
    sq = np.array([[h, -h, 50], [h, h, 50], [-h, h, 50], [-h, -h, 50]], np.float32)
    poly = ob.poly_to_diff(sq, 1.0, [0, 0, 0], [0, 0, 1])
    diff = ob.diff_to_diff([0, 0, 50], [0, 0, -1], [0, 0, 0], [0, 0, 1])
    assert poly == pytest.approx(diff / np.pi * np.pi, rel=1e-3)


# --- MakeScales (src/SourceSharp.Tests/MapTools/Rad/Bounce/VisMatrixTests.cs:27-57) ---

def test_make_scales_under_pi_divides_by_pi():
    s = np.float32(1.0 / np.pi)
    np.testing.assert_array_equal(ob.make_scales([1.0, 2.0]), [np.float32(1.0) * s, np.float32(2.0) * s])


def test_make_scales_over_pi_divides_by_total():
    np.testing.assert_array_equal(ob.make_scales([3.0, 5.0]), [3.0 * (1 / 8.0), 5.0 * (1 / 8.0)])
    assert ob.make_scales([]).shape == (0,)


@pytest.fixture(scope="module")
def small_grid():
    g = rooms.room_grid(2, 2, size=(384, 320, 192), patch=128, minchop=64, seed=5)
    rp, src, w, ncand = ob.build_transfers(g)
    return g, rp, src, w, ncand


def test_room_grid_structure(small_grid):
    g = small_grid[0]
    assert np.all(g.child1[g.child1 >= 0] > np.nonzero(g.child1 >= 0)[0])  # children after parents
    # areas: a parent is the sum of its children
    par = np.nonzero(g.child1 >= 0)[0]
    np.testing.assert_allclose(g.area[par], g.area[g.child1[par]] + g.area[g.child2[par]], rtol=1e-6)
    # every leaf's winding is wound against its normal
    w = g.winding
    cross = np.cross(w[:, 1] - w[:, 0], w[:, 2] - w[:, 0])
    assert np.all(np.einsum("ij,ij->i", cross, g.normal) < 0)


def test_transfer_lists_obey_the_build_rules(small_grid):
    g, rp, src, w, ncand = small_grid
    assert 0 < len(src) <= ncand
    leaf = g.leaf
    for r in range(g.npatch):
        a, b = rp[r], rp[r + 1]
        if not leaf[r]:
            assert a == b  # only receivers have lists
            continue
        s = src[a:b]
        if len(s) == 0:
            continue
        assert np.all(g.face[s] != g.face[r])  # never its own face
        total = float(np.sum(w[a:b], dtype=np.float64))
        assert total <= 1.0 + 1e-5  # MakeScales: rows sum to 1, or to total/pi below pi
        assert np.all(w[a:b] > 0)
        # sources are in the receiver's PVS row
        row = set(g.row_faces[g.row_faces_ptr[g.cluster[r]]:g.row_faces_ptr[g.cluster[r] + 1]].tolist())
        assert set(g.face[s].tolist()) <= row
        # in front of the receiver's plane
        assert np.all(g.origin[s] @ g.normal[r] > g.plane_dist[r] + 0.01)


def test_candidates_respect_hierarchical_descent(small_grid):
    g = small_grid[0]
    ps = ob.PatchSet(g)
    r = int(g.receivers[len(g.receivers) // 2])
    c = ps.candidates(r)
    for s in c:
        if g.child1[s] >= 0:  # a parent kept as a source is far enough away
            d = g.origin[r] - g.origin[s]
            assert float(d @ d) * 0.0625 >= g.area[s]


def test_walls_block_transfers_between_rooms_without_a_doorway():
    """In a 1x3 row of rooms, room 0 and room 2 do not see each other (not in
    each other's PVS row) and room 0 reaches room 1 only through the doorway:
    a patch on the far side of room 1's wall with room 0 gets no light."""
    g = rooms.room_grid(3, 1, size=(256, 256, 192), patch=128, minchop=64, seed=2)
    rp, src, w, _ = ob.build_transfers(g)
    for r in g.receivers:
        s = src[rp[r]:rp[r + 1]]
        cl = g.cluster[r]
        assert np.all(np.abs(g.cluster[s] - cl) <= 1)


def dense_bounce(g, rp, src, w, e0):
    """The converged bounced light from the linear system, in float64:
    E_{n+1} = M E_n with M = W diag(rho) U (U pulls leaf emission up to
    parents); total = sum_{n>=1} M^n E_0 = (I - M)^-1 M E_0."""
    P = g.npatch
    leaves = np.nonzero(g.leaf)[0]
    L = len(leaves)
    pos = -np.ones(P, np.int64)
    pos[leaves] = np.arange(L)
    # U: patch emission as a combination of leaf emission (parents area weighted)
    U = np.zeros((P, L))
    U[leaves, pos[leaves]] = 1.0
    for i in range(P - 1, -1, -1):
        if g.child1[i] >= 0:
            a, b = g.child1[i], g.child2[i]
            s1 = g.area[a] / (g.area[a] + g.area[b])
            U[i] = s1 * U[a] + (1 - s1) * U[b]
    sky_leaf = g.sky[leaves].astype(bool)
    U[:, sky_leaf] = 0.0
    W = np.zeros((L, P))
    for r in leaves:
        W[pos[r], src[rp[r]:rp[r + 1]]] += w[rp[r]:rp[r + 1]]
    out = np.zeros((P, 3))
    for c in range(3):
        M = W @ (g.refl[:, c:c + 1].astype(np.float64) * U)
        E0 = e0[leaves, c].astype(np.float64)
        E0[sky_leaf] = 0.0
        tot = np.linalg.solve(np.eye(L) - M, M @ E0)
        full = U @ tot  # parents pulled up
        full[leaves] = tot
        out[:, c] = full
    return out


def test_fp64_bounce_matches_the_dense_solve(small_grid):
    g, rp, src, w, _ = small_grid
    rng = np.random.default_rng(3)
    e0 = np.zeros((g.npatch, 3), np.float32)
    lv = np.nonzero(g.leaf)[0]
    pick = rng.choice(lv, 40, replace=False)
    e0[pick] = rng.uniform(100, 1000, (40, 3))
    e0 = g.pull_up(e0)
    want = dense_bounce(g, rp, src, w, e0)
    tot, res, en, it = ob.bounce(g, rp, src, w, e0, tol=1e-12, max_iter=500, use_double=True)
    assert res[-1] <= 1e-12
    np.testing.assert_allclose(tot, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_fp32_reference_form_is_close_to_fp64(small_grid):
    g, rp, src, w, _ = small_grid
    rng = np.random.default_rng(4)
    e0 = g.pull_up(np.where(g.leaf[:, None], rng.uniform(0, 50, (g.npatch, 3)), 0).astype(np.float32))
    t32, _, _, n32 = ob.bounce(g, rp, src, w, e0, tol=1e-6, max_iter=300)
    t64, _, _, n64 = ob.bounce(g, rp, src, w, e0, tol=1e-6, max_iter=300, use_double=True)
    rel = np.abs(t32 - t64).sum() / np.abs(t64).sum()
    assert rel < 1e-5


def test_reference_stop_rule_and_from_start_form(small_grid):
    g, rp, src, w, _ = small_grid
    rng = np.random.default_rng(5)
    e0 = g.pull_up(np.where(g.leaf[:, None], rng.uniform(0, 5, (g.npatch, 3)), 0).astype(np.float32))
    # the reference's rule: stop once every channel of the added light is under 1
    t, res, en, it = ob.bounce(g, rp, src, w, e0, tol=0.0, max_iter=100, use_double=True)
    assert np.all(en[-1] < 1.0) and (it == 1 or not np.all(en[-2] < 1.0))
    # from a zero start the two forms produce the same totals
    a, _, _, ia = ob.bounce(g, rp, src, w, e0, tol=1e-8, max_iter=300, use_double=True)
    b, _, _, ib = ob.bounce(g, rp, src, w, e0, tol=1e-8, max_iter=300, use_double=True, mode=1,
                           start=np.zeros_like(e0))
    assert ia == ib
    np.testing.assert_allclose(a, b, rtol=1e-7, atol=1e-9)
    # starting from the answer converges at once
    c, res_c, _, ic = ob.bounce(g, rp, src, w, e0, tol=1e-6, max_iter=300, use_double=True, mode=1,
                                start=a.astype(np.float32))
    assert ic <= 2


def _two_patches(src_origin, src_normal):
    from types import SimpleNamespace

    g = SimpleNamespace(npatch=2, refl=np.ones((2, 3), np.float32), sky=np.zeros(2, np.uint8),
                        child1=np.array([-1, -1], np.int32), child2=np.array([-1, -1], np.int32),
                        area=np.ones(2, np.float32),
                        origin=np.array([(0, 0, 0), src_origin], np.float32),
                        normal=np.array([(0, 0, 1), src_normal], np.float32))
    return g


def test_a_bumped_receiver_gathers_along_each_normal_it_faces():
    """GatherBumpedScalar (Rad/Bounce/Radiosity.cs:406-421): the cosine divided
    out and re-applied per normal; a source straight above gives the flat
    gather on the flat slot and cos(bump normal, z) on each bump slot."""
    from hl2sdk_tools_sharp_amd.light import bump_basis

    g = _two_patches((0, 0, 10), (0, 0, -1))
    rp = np.array([0, 1, 1], np.int64)
    src = np.array([1], np.int32)
    w = np.array([0.5], np.float32)
    e0 = np.array([(0, 0, 0), (2, 4, 6)], np.float32)
    bn = np.zeros((2, 3, 3), np.float32)
    bn[0] = bump_basis(np.array([0, 0, 1.0]))
    tb = np.zeros((2, 3, 3))
    flat, *_ = ob.bounce(g, rp, src, w, e0, max_iter=1, tol=1e-9)
    tot, *_ = ob.bounce(g, rp, src, w, e0, max_iter=1, tol=1e-9, bumped=np.array([1, 0], np.uint8),
                        bump_normals=bn, bump_out=tb)
    assert list(tot[0]) == [1.0, 2.0, 3.0] == list(flat[0])
    for k in range(3):
        d = np.float32(bn[0, k, 2])
        assert list(tb[0, k]) == [float(np.float32(v) * d) for v in (1, 2, 3)]
    assert np.all(tb[1] == 0)


def test_a_bump_normal_the_source_is_behind_gets_nothing():
    from hl2sdk_tools_sharp_amd.light import bump_basis

    bn = np.zeros((2, 3, 3), np.float32)
    bn[0] = bump_basis(np.array([0, 0, 1.0]))
    # a source low on the side away from the first bump normal's tilt
    t = bn[0, 0] - np.array([0, 0, bn[0, 0, 2]], np.float32)
    t /= np.linalg.norm(t)
    so = (-10 * t + np.array([0, 0, 1.0])).astype(np.float32)
    g = _two_patches(tuple(so), (0, 0, -1))
    rp = np.array([0, 1, 1], np.int64)
    tb = np.zeros((2, 3, 3))
    tot, *_ = ob.bounce(g, rp, np.array([1], np.int32), np.array([0.5], np.float32),
                        np.array([(0, 0, 0), (2, 4, 6)], np.float32), max_iter=1, tol=1e-9,
                        bumped=np.array([1, 0], np.uint8), bump_normals=bn, bump_out=tb)
    assert np.all(tb[0, 0] == 0)
    assert np.all(tb[0, 1] > 0) and np.all(tb[0, 2] > 0)
    # the flat slot: the cosine removed and put back, the flat gather within rounding
    np.testing.assert_allclose(tot[0], [1, 2, 3], rtol=1e-6)


def test_threaded_gather_light_is_the_bounces_first_pass():
    """vc_gather_light (the bounce's CPU baseline) is exactly the flat rows'
    GatherLight of the bounce: with emission e0, one progressive iteration's
    bounced light is GatherLight(e0 * reflectivity), bit for bit, whatever the
    thread count."""
    from hl2sdk_tools_sharp_amd import rooms

    g = rooms.room_grid(2, 2, size=(384, 320, 192), patch=128, minchop=64, seed=5)
    rp, src, w, _ = ob.build_transfers(g)
    e0 = np.random.default_rng(1).uniform(0, 50, (g.npatch, 3)).astype(np.float32)
    tot, _, _, _ = ob.bounce(g, rp, src, w, e0, max_iter=1)
    shoot = (e0 * g.refl).astype(np.float32)
    leaf = g.leaf & ~g.sky.astype(bool)
    for threads in (1, 5):
        add = ob.gather_light(rp, src, w, shoot, threads=threads)
        assert np.array_equal(tot[leaf].astype(np.float32), add[leaf])
