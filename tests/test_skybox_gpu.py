"""The gather's 3D-skybox stage on the GPU (vh_gather_skybox; TestLineDoesHitSky's
recursion, Rad/Light/DirectLightGatherer.cs:510-714) against the oracle: every
LDR and HDR value bit-equal and the same rays traced, over the three
schedules and through supersampling, on a 3 x 3 room grid lit through sky
ceilings with two skyboxes (rooms.skybox_grid: cameras at scale 16 and 32,
roofs and a tower in the skyboxes, a BSP with axial and slanted planes and a
corner of the grid in the cameras' own area), every light type, a jittered
sun and sky ambient, bumped faces.  The closed forms (gather_cases' skybox
facts) run on the device in test_gather_gpu.  Needs an MI355X."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle.gather as og
import oracle.supersample as oss
from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, light, rooms
from hl2sdk_tools_sharp_amd.light import HipFaceLighter
from hl2sdk_tools_sharp_amd.skybox import CompiledBspTree, Skybox

pytestmark = pytest.mark.gpu


def open_lighter(casters):
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer, HipFaceLighter(offer.tracer)


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def case():
    m = rooms.skybox_grid(3, 3, nlights=9, seed=13)
    fset = rooms.room_faces(m.grid, 16.0, bumped_every=7)
    lights, pvs = rooms.room_lights(m.grid, sun=True, sky_ambient=True, spot_every=3, surface_every=4)
    sun = np.array([0.15, 0.1, 1.0], np.float32)
    lights[len(m.grid.lights)]["normal"] = -sun / np.float32(np.linalg.norm(sun))
    for k in ("intensity", "intensity_hdr"):
        lights[k] *= 100  # bright enough for shadow edges to pass the supersampling gradient
    sky = light.sky_directions()
    kw = dict(light_pvs=pvs, sun_extent=0.05, sun_samples=light.SUN_AREA_LIGHT_SAMPLES, sky_dirs=sky)
    desc = light.pack_desc(fset, lights, pvs, 0.05, light.SUN_AREA_LIGHT_SAMPLES, sky, skybox=m.skybox)
    want = og.gather_styles(m.casters, desc, fset.nsamples)
    plain = og.gather_styles(m.casters, light.pack_desc(fset, lights, pvs, 0.05, light.SUN_AREA_LIGHT_SAMPLES,
                                                        sky), fset.nsamples)
    print(f"[skybox] {fset.nfaces} faces, {fset.nsamples} samples, {len(lights)} lights, "
          f"{want.rays} oracle rays", flush=True)
    return m, fset, lights, kw, want, plain


@pytest.mark.parametrize("schedule", ["queue", "grid", "per_face"])
def test_skybox_gather_matches_oracle(case, schedule):
    m, fset, lights, kw, want, plain = case
    print(f"[skybox] gather, schedule {schedule}", flush=True)
    tr, fl = open_lighter(m.casters)
    with tr:
        fl.load(fset, lights, skybox=m.skybox, **kw)
        before = tr.stats()["rays_gather"]
        got = fl.light_faces(schedule=schedule)
        rays = tr.stats()["rays_gather"] - before
    assert same(got.ldr, want.ldr[0]) and same(got.hdr, want.hdr[0])
    assert rays == want.rays
    # the skyboxes shade part of the map: the result is not the plain gather's
    assert not same(want.ldr[0], plain.ldr[0]) and np.all(want.ldr[0] <= plain.ldr[0])


def test_skybox_removed_on_reload(case):
    """Loading the gather again without a skybox gives the plain gather."""
    m, fset, lights, kw, want, plain = case
    tr, fl = open_lighter(m.casters)
    with tr:
        fl.load(fset, lights, skybox=m.skybox, **kw)
        a = fl.light_faces()
        fl.load(fset, lights, **kw)
        b = fl.light_faces()
    assert same(a.ldr, want.ldr[0]) and same(b.ldr, plain.ldr[0])


@pytest.mark.parametrize("which", [0, 1])
def test_skybox_supersample_matches_oracle(case, which):
    """The resample's points look up their own leaves: a luxel's 4 x 4 points
    can fall on both sides of the slanted plane."""
    m, fset, lights, kw, want, plain = case
    ff, _ = rooms.room_final_inputs(m.grid, fset)
    base = want.hdr[0] if which else want.ldr[0]
    ref, ref_counts = oss.supersample(m.casters, fset, lights, kw["light_pvs"], kw["sun_extent"],
                                      kw["sun_samples"], kw["sky_dirs"], ff.faces, base, which=which,
                                      skybox=m.skybox)
    print(f"[skybox] supersample which={which}: passes {ref_counts}", flush=True)
    tr, fl = open_lighter(m.casters)
    with tr:
        fl.load(fset, lights, skybox=m.skybox, **kw)
        fl.run()
        counts = fl.supersample(ff.faces, which=which)
        got = fl.read()
    assert counts == ref_counts and len(counts) >= 2
    assert same(got.hdr if which else got.ldr, ref)


def _raw(tr, desc):
    return _lib.lib().vh_gather_skybox(tr.handle, None if desc is None else ctypes.byref(desc))


def test_skybox_descriptor_checks(case):
    """vh_gather_skybox refuses what could not be walked or cast: too many
    cameras, a non-finite camera, a plane index out of range, a node reached
    twice from node 0 (a cycle would never end); NULL removes the skybox."""
    m = case[0]
    t = m.tree
    ok = m.skybox

    def variant(**kw):
        tree = CompiledBspTree(kw.get("node_plane", t.node_plane), kw.get("node_children", t.node_children),
                               t.plane_normal, t.plane_dist, t.plane_type)
        return Skybox(kw.get("origin", ok.camera_origin), kw.get("w2s", ok.camera_world_to_sky), tree,
                      ok.leaf_recurses)

    tr, _ = open_lighter(m.casters)
    with tr:
        assert _raw(tr, ok.pack()) == _lib.VH_OK
        many = variant(origin=np.zeros((65, 3), np.float32), w2s=np.ones(65, np.float32))
        bad_cam = variant(origin=np.array([[0, np.inf, 0], [0, 0, 0]], np.float32))
        bad_plane = variant(node_plane=[0, 1, 9])
        cycle = variant(node_children=[[1, -2], [-3, 2], [1, -5]])
        shared = variant(node_children=[[1, 2], [-3, 2], [-4, -5]])
        for sb in (many, bad_cam, bad_plane, cycle, shared):
            assert _raw(tr, sb.pack()) == _lib.VH_EINVAL, _lib.last_error()
        assert _raw(tr, None) == _lib.VH_OK
