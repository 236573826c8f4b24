"""The 3D skybox's host logic and the oracle's restatement of its pieces,
against the reference's own facts (CPU only):

- CompiledBspTree.LeafFromPoint (RadWorldTests.cs:400-446), in the host mirror
  (skybox.CompiledBspTree) and in the oracle (vc_bsp_leaf_from_point);
- SkyCameras.Build (GatherTests.cs:361-395: ACameraClaimsTheAreaItStandsIn,
  ACameraWithNoScaleIsIgnored; SkyCameras.cs:96-101: a later camera wins a
  shared area);
- ResolveSkyTest's arithmetic over given answers (SkyRecursionTests,
  GatherTests.cs:410-545: the replayed fractions), in the oracle.

The gather's full recursion is pinned by gather_cases' skybox facts (run on
the oracle in test_oracle_gather and on the device in test_gather_gpu)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.gather as og
from hl2sdk_tools_sharp_amd import rooms
from hl2sdk_tools_sharp_amd.skybox import CompiledBspTree, SkyCameras, Skybox, float_for_key, vector_for_key


def split_tree():
    """RadWorldTests.Split: one node on the plane x = 100, front leaf 1, back leaf 0."""
    return CompiledBspTree(node_plane=[0], node_children=[[-2, -1]], plane_normal=[[1, 0, 0]],
                           plane_dist=[100.0], plane_type=[0])


def no_nodes():
    return CompiledBspTree(node_plane=np.zeros(0), node_children=np.zeros((0, 2)),
                           plane_normal=np.zeros((0, 3)), plane_dist=np.zeros(0), plane_type=np.zeros(0))


def as_skybox(tree, nleaves=8):
    return Skybox(np.zeros((1, 3), np.float32), np.ones(1, np.float32), tree, np.zeros(nleaves, np.uint8))


@pytest.mark.parametrize("where", ["mirror", "oracle"])
def test_leaf_from_point_splits_on_the_sign_with_no_epsilon(where):
    t = split_tree()
    q = t.leaf_from_point if where == "mirror" else (lambda p: og.bsp_leaf_from_point(as_skybox(t), p))
    assert q([100, 0, 0]) == 1  # on the plane: front
    assert q(np.array([99.99, 0, 0], np.float32)) == 0
    assert q([200, 0, 0]) == 1 and q([0, 0, 0]) == 0


@pytest.mark.parametrize("where", ["mirror", "oracle"])
def test_a_tree_with_no_nodes_is_leaf_zero(where):
    t = no_nodes()
    q = t.leaf_from_point if where == "mirror" else (lambda p: og.bsp_leaf_from_point(as_skybox(t), p))
    assert q([5, 5, 5]) == 0


@pytest.mark.parametrize("where", ["mirror", "oracle"])
def test_leaf_from_point_plane_forms(where):
    """A non-axial plane takes the dot product (a point on it goes front); a
    type below 3 reads one coordinate whatever the normal says, and a
    negative type reads z; a child index past the last node is leaf 0."""
    # node 0: 0.6 x + 0.8 y = 300 (dot); node 1: z = 10 (dot); node 2: type 0 with
    # normal (-1, 0, 0), so x = 5 axially where the dot product would give -x - 5;
    # node 3: type -1 with normal (1, 0, 0), so z = 50 where the dot would read x;
    # node 3's front child 7 is past the last node
    t = CompiledBspTree(
        node_plane=[0, 1, 2, 3],
        node_children=[[1, -2], [2, -3], [3, -4], [7, -5]],
        plane_normal=[[0.6, 0.8, 0.0], [0, 0, 1], [-1, 0, 0], [1, 0, 0]],
        plane_dist=[300.0, 10.0, 5.0, 50.0],
        plane_type=[3, 3, 0, -1])
    q = t.leaf_from_point if where == "mirror" else (lambda p: og.bsp_leaf_from_point(as_skybox(t), p))
    assert q(np.array([499.99, 0, 100], np.float32)) == 1  # 299.994 < 300: back
    assert q([500, 0, 100]) == 0  # on the plane: front, and on to node 3's front: past the nodes, leaf 0
    assert q([1000, 0, 0]) == 2  # z = 0 < 10: back
    assert q([4, 400, 20]) == 3  # x = 4 < 5: back
    assert q([10, 400, 20]) == 4  # x = 10 >= 5 axially (the dot would say back); then z = 20 < 50: back
    assert q([1000, 0, 20]) == 4  # type -1 reads z (20 < 50), not x
    assert q([1000, 0, 60]) == 0  # z = 60 >= 50: front child 7, past the nodes: leaf 0


def test_mirror_and_oracle_agree_on_the_skybox_map():
    m = rooms.skybox_grid()
    rng = np.random.default_rng(5)
    pts = np.concatenate([rng.uniform([-9000, -3000, -10], [1100, 3000, 300], size=(3000, 3)),
                          np.array([[500, 0, 5], [0, 375, 5], [-4096, 0, 0], [0, 516, 0]])]).astype(np.float32)
    a = [m.tree.leaf_from_point(p) for p in pts]
    b = [og.bsp_leaf_from_point(m.skybox, p) for p in pts]
    assert a == b and set(a) == {1, 2, 3, 4}


def test_a_camera_claims_the_area_it_stands_in():
    ents = [{"classname": "SKY_CAMERA", "origin": "1 1 1", "scale": "16"}]
    c = SkyCameras.build(ents, no_nodes(), np.array([0]), 1)
    assert c.origin.shape == (1, 3)
    assert (float(c.sky_to_world[0]), float(c.world_to_sky[0]), int(c.area[0])) == (16.0, 1.0 / 16.0, 0)
    assert c.camera_in_area(0) == 0
    assert c.leaf_recurses(np.array([0])).tolist() == [0]


def test_a_camera_with_no_scale_is_ignored():
    ents = [{"classname": "sky_camera", "origin": "1 1 1"}, {"classname": "sky_camera", "scale": "-2"},
            {"classname": "light", "origin": "0 0 0", "scale": "4"}]
    c = SkyCameras.build(ents, no_nodes(), np.array([0]), 1)
    assert len(c.origin) == 0 and c.camera_in_area(0) == -1
    assert c.leaf_recurses(np.array([0, -1, 1, 5])).tolist() == [1, 0, 0, 0]


def test_a_later_camera_wins_a_shared_area():
    t = split_tree()  # leaf 1 (x >= 100) in area 1, leaf 0 in area 0
    ents = [{"classname": "sky_camera", "origin": "200 0 0", "scale": "8"},
            {"classname": "sky_camera", "origin": "300 0 0", "scale": "4"}]
    c = SkyCameras.build(ents, t, np.array([0, 1]), 2)
    assert len(c.origin) == 2 and c.camera_in_area(1) == 1 and c.camera_in_area(0) == -1
    assert c.leaf_recurses(np.array([0, 1])).tolist() == [1, 0]


def test_entity_key_scanning():
    e = {"origin": " 1.5  -2e1 ", "scale": "16x", "bad": "x1"}
    assert vector_for_key(e, "origin").tolist() == [1.5, -20.0, 0.0]
    assert float_for_key(e, "scale") == 16.0 and float_for_key(e, "bad") == 0.0
    assert float_for_key(e, "missing") == 0.0


def test_skybox_grid_tables():
    m = rooms.skybox_grid()
    assert m.skybox.ncameras == 2  # the scale-0 camera is dropped
    assert m.skybox.camera_world_to_sky.tolist() == [1.0 / 16.0, 1.0 / 32.0]
    assert m.skybox.leaf_recurses.tolist() == [1, 0, 1, 1, 0]


# --- ResolveSkyTest over replayed answers (SkyRecursionTests) ---

def test_replayed_fractions_add_each_cameras_occlusion():
    # DeferredAndEagerRecursionGiveTheSameFractions (GatherTests.cs:497-524)
    first, skybox = [1, 0, 0, 1], [[1, 1, 0, 0]]
    assert og.sky_resolve(4, first, skybox, 0b1111).tolist() == [0, 0, 1, 0]


def test_replayed_lane_outside_the_recursion_keeps_its_first_answer():
    # AStockGroupRecursesByLaneZerosArea, the correct side (GatherTests.cs:418-443):
    # lane 0 in the skybox's own area, the others recurse into a blocked skybox
    assert og.sky_resolve(4, [0, 0, 0, 0], [[1, 1, 1, 1]], 0b1110).tolist() == [1, 0, 0, 0]


def test_a_fully_occluded_group_takes_no_skybox_answer():
    assert og.sky_resolve(4, [1, 1, 1, 1], [[0, 0, 0, 0]], 0b1111).tolist() == [0, 0, 0, 0]
    # two cameras: either one's hit occludes
    assert og.sky_resolve(4, [0, 0, 0, 0], [[0, 1, 0, 0], [0, 0, 1, 0]], 0b1111).tolist() == [1, 0, 0, 1]


def test_padding_lanes_copy_the_last_real_lane():
    # three real lanes: lane 3's inputs are ignored and copy lane 2's
    assert og.sky_resolve(3, [0, 0, 1, 0], [[0, 1, 0, 0]], 0b1111).tolist() == [1, 0, 0, 0]
    assert og.sky_resolve(3, [1, 1, 1, 0], [[0, 0, 0, 0]], 0b1111).tolist() == [0, 0, 0, 0]


def test_planted_replace_breaks_the_replay():
    og.set_mutation(17)
    try:
        assert og.sky_resolve(4, [1, 0, 0, 1], [[1, 1, 0, 0]], 0b1111).tolist() != [0, 0, 1, 0]
    finally:
        og.set_mutation(0)
