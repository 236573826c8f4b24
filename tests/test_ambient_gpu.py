"""The GPU leaf-ambient fan against the oracle.  Needs an MI355X.

The bar: every cube component bit-equal to the oracle's (oracle/ambient.py:
the oracle's BVH closest hit, pinned to the reference's KD closest hit, and
RayAmbientLighting/AmbientCube restated in float32), on a room map lit by the
GPU gather and final stage, reading the lump where the final stage left it."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.ambient as oa
from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, ambient, rooms
from hl2sdk_tools_sharp_amd.final import HipFinalLighter
from hl2sdk_tools_sharp_amd.light import HipFaceLighter

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lit_rooms():
    """A 2 x 2 room map under a sky ceiling, lit by the GPU gather and final
    stage (a seeded stand-in for the bounced light), with sample points."""
    g = rooms.room_grid(2, 2, sky_ceiling=True, seed=7)
    fset = rooms.room_faces(g, luxel=32.0)
    lights, pvs = rooms.room_lights(g, sun=True)
    ff, fp = rooms.room_final_inputs(g, fset, luxel=32.0)
    pts = rooms.room_ambient_points(g, per_room=24, seed=3)
    print(f"[ambient] {fset.nfaces} faces, {len(pts)} points", flush=True)
    return g, fset, lights, pvs, ff, fp, pts


def cube_lights(g):
    """Surface lights under each room's ceiling, facing down: dim ones the
    cubes take (one with a radius), a bright one they do not."""
    rows = []
    for k, (lo, hi) in enumerate(zip(g.room_lo, g.room_hi)):
        c = (lo + hi) / 2
        for dx, inten, radius in ((-60, 300.0, 0.0), (60, 900.0, 150.0), (0, 5000.0, 0.0)):
            rows.append(((c[0] + dx, c[1], hi[2] - 4), (0, 0, -1), (inten, inten * 0.5, inten * 0.25), radius))
    L = np.zeros(len(rows), ambient.LIGHT_DTYPE)
    for k, (o, n, i, r) in enumerate(rows):
        L[k]["origin"], L[k]["normal"], L[k]["intensity"], L[k]["radius"] = o, n, i, r
    return L


def run_gpu(g, fset, lights, pvs, ff, fp, pts, host_lump=False, sky=None, wl=None):
    offer = HipTracerFactory().try_create(g.casters)
    assert offer.tracer is not None, offer.decline_reason
    with offer.tracer as tr:
        fl = HipFaceLighter(tr)
        fl.load(fset, lights, light_pvs=pvs, sun_samples=1)
        fl.run()
        fin = HipFinalLighter(fl)
        fin.load(ff, fp)
        offs, _ = fin.layout()
        pl = np.zeros((fp.npatch, 4, 3), np.float32)
        pl[:, 0] = np.random.default_rng(2).uniform(0, 30, (fp.npatch, 3))
        lump = fin.run(bounce="host", patch_light=pl)
        faces = rooms.room_ambient_faces(g, ff, offs)
        amb = ambient.HipLeafAmbient(tr)
        amb.load(faces, sky_ambient=sky, lump=lump if host_lump else None, lights=wl)
        cubes = amb.cubes(pts)
    return faces, lump, cubes


@pytest.mark.parametrize("host_lump,sky,with_lights", [(False, None, False), (True, (40.0, 60.0, 90.0), False),
                                                      (False, (40.0, 60.0, 90.0), True)])
def test_cubes_match_oracle(lit_rooms, host_lump, sky, with_lights):
    g, fset, lights, pvs, ff, fp, pts = lit_rooms
    wl = cube_lights(g) if with_lights else None
    faces, lump, cubes = run_gpu(*lit_rooms, host_lump=host_lump, sky=sky, wl=wl)
    want = oa.ambient_cubes(g.casters, faces, sky, lump, pts, threads=8, lights=wl)
    if with_lights:
        base = oa.ambient_cubes(g.casters, faces, sky, lump, pts, threads=8)
        assert np.any(want != base) and len(oa.flagged_lights(wl)) == 2 * len(g.room_lo)
    assert np.any(want > 0)
    same = cubes.view(np.uint32) == want.view(np.uint32)
    if not same.all():
        bad = np.argwhere(~same)[:5]
        raise AssertionError(f"{(~same).sum()} of {same.size} differ, e.g. {bad.tolist()}: "
                             f"{cubes[tuple(bad[0])]} vs {want[tuple(bad[0])]}")


def test_directions_are_the_fan_table():
    assert np.array_equal(ambient.directions(), oa.VERTEX_NORMALS)


def test_errors():
    g = rooms.room_grid(1, 1, seed=1)
    offer = HipTracerFactory().try_create(g.casters)
    with offer.tracer as tr:
        amb = ambient.HipLeafAmbient(tr)
        with pytest.raises(_lib.VradHipError):
            amb.load(np.zeros(0, ambient.FACE_DTYPE))  # no final lump on the device
        with pytest.raises(_lib.VradHipError):
            amb.cubes(np.zeros((1, 3), np.float32))   # nothing loaded
        bad = np.zeros(1, ambient.FACE_DTYPE)
        bad["light_ofs"], bad["smax"], bad["tmax"] = 8, 4, 4
        with pytest.raises(ValueError):
            amb.load(bad, lump=np.zeros(16, np.uint8))  # the face's lightmap overruns the lump
