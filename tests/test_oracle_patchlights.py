"""The oracle's PatchLighting.BuildPatchLights restatement against closed
forms and a scalar walk of the reference's loops (CPU)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.bounce as ob
from hl2sdk_tools_sharp_amd import rooms


@pytest.fixture(scope="module")
def grid():
    g = rooms.room_grid(2, 1, size=(256, 192, 128), patch=64, minchop=32, sky_ceiling=True, seed=2)
    faces = rooms.room_faces(g, luxel=16.0)
    return g, faces


def scalar_walk(g, faces, light, area):
    """AddSampleToPatch and BuildPatchLights as the reference's loops walk them
    (Rad/Light/PatchLighting.cs:65-249), one sample and one patch at a time."""
    f32 = np.float32
    P = g.npatch
    sa = np.zeros(P, f32)
    sl = np.zeros((P, 3), f32)
    for f in range(g.nface):
        patches = np.nonzero(g.face == f)[0]
        for s in range(faces.face_sample_ptr[f], faces.face_sample_ptr[f + 1]):
            L = light[s].astype(f32)
            if ((L[0] + L[1]) + L[2]) / f32(3) < f32(1):
                continue
            r = f32(np.float64(np.sqrt(f32(area[s]))) / 2.0)
            for p in patches:
                if g.face_sky[f] or g.child1[p] >= 0:
                    continue
                mins, maxs = g.winding[p].min(axis=0), g.winding[p].max(axis=0)
                if any(mins[i] > faces.sample_pos[s, i] + r or maxs[i] < faces.sample_pos[s, i] - r
                       for i in range(3)):
                    continue
                sa[p] = sa[p] + f32(area[s])
                sl[p] = sl[p] + f32(area[s]) * L
    d = np.zeros((P, 3), f32)
    for p in range(P):
        if g.child1[p] < 0 and sa[p] != 0:
            d[p] = sl[p] * f32(1.0 / np.float64(sa[p]))
    return g.pull_up(d)


def test_uniform_light_reaches_every_lit_patch(grid):
    g, faces = grid
    light = np.full((faces.nsamples, 3), 2.0, np.float32)
    d = ob.patch_lights(g, faces, light, faces.sample_area)
    nonsky = ~g.face_sky[g.face].astype(bool)
    np.testing.assert_allclose(d[nonsky], 2.0, rtol=1e-6)
    assert np.all(d[~nonsky & (g.child1 < 0)] == 0)


def test_samples_averaging_under_one_are_dropped(grid):
    g, faces = grid
    light = np.full((faces.nsamples, 3), 0.999, np.float32)
    assert not ob.patch_lights(g, faces, light, faces.sample_area).any()
    light[:, 0] = 1.002  # average (1.002 + 0.999 + 0.999) / 3 = 1.0
    assert ob.patch_lights(g, faces, light, faces.sample_area).any()


def test_dark_samples_do_not_dilute(grid):
    """Samples under 1 are dropped, not averaged in: a patch with any lit
    sample in reach takes exactly the lit value; one with none takes 0."""
    g, faces = grid
    rng = np.random.default_rng(0)
    lit = rng.uniform(size=faces.nsamples) < 0.5
    light = np.where(lit[:, None], 4.0, 0.0).astype(np.float32).repeat(3, axis=1)[:, :3]
    d = ob.patch_lights(g, faces, light, faces.sample_area)
    leaf = (g.child1 < 0) & ~g.face_sky[g.face].astype(bool)
    vals = d[leaf, 0]
    assert np.all((np.abs(vals - 4.0) <= 4e-6) | (vals == 0.0))
    assert (vals == 0.0).any() and (vals > 0).any()


def test_matches_the_scalar_walk(grid):
    g, faces = grid
    rng = np.random.default_rng(1)
    light = rng.uniform(0.0, 3.0, (faces.nsamples, 3)).astype(np.float32)
    d = ob.patch_lights(g, faces, light, faces.sample_area)
    want = scalar_walk(g, faces, light, faces.sample_area)
    assert np.array_equal(d.view(np.uint32), want.view(np.uint32))


def test_parents_are_area_weighted(grid):
    g, faces = grid
    rng = np.random.default_rng(2)
    light = rng.uniform(0.0, 3.0, (faces.nsamples, 3)).astype(np.float32)
    d = ob.patch_lights(g, faces, light, faces.sample_area).astype(np.float64)
    for p in np.nonzero(g.child1 >= 0)[0]:
        c1, c2 = g.child1[p], g.child2[p]
        a1, a2 = float(g.area[c1]), float(g.area[c2])
        np.testing.assert_allclose(d[p], (d[c1] * a1 + d[c2] * a2) / (a1 + a2), rtol=1e-6, atol=1e-7)
