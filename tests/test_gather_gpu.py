"""The GPU direct-light gather, lightmap encode and AO against the oracle.
Needs an MI355X.

The bar: every sample's LDR and HDR lighting (per normal slot: RGB and the
direct-sun amount) and AO bit-equal to the oracle's (both raise the spot
cone's multiplier with the correctly rounded DetMathF.Pow); the RGBE bytes
equal; every schedule (persistent queue, grid, one launch per face) gives the
same bits.  The reference's closed-form gather facts (GatherTests.cs) hold on
the device itself, ray counts included."""
from __future__ import annotations

import numpy as np
import pytest

import gather_cases as gc
import oracle.gather as og
from hl2sdk_tools_sharp_amd import HipTracerFactory, light, rooms, scenes
from hl2sdk_tools_sharp_amd.light import HipFaceLighter

pytestmark = pytest.mark.gpu


def open_lighter(casters):
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer, HipFaceLighter(offer.tracer)


@pytest.fixture(scope="module")
def c1():
    """A 2x2 room map under a sky ceiling: point, spot and surface lights
    behind PVS rows, a sun with a 30-sample area, sky ambient, every fifth
    face bumped, 32-ray AO."""
    g = rooms.room_grid(2, 2, sky_ceiling=True, seed=7)
    faces = rooms.room_faces(g, luxel=32.0, bumped_every=5)
    lights, pvs = rooms.room_lights(g, sun=True, sky_ambient=True, spot_every=3, surface_every=3)
    kw = dict(light_pvs=pvs, sun_extent=0.02, sun_samples=30, sky_dirs=light.sky_directions(),
              ao_count=32, ao_radius=40.0)
    d = light.pack_desc(faces, lights, pvs, kw["sun_extent"], kw["sun_samples"], kw["sky_dirs"],
                        kw["ao_count"], kw["ao_radius"])
    want = og.gather(g.casters, d, faces.nsamples)
    print(f"[c1] {faces.nfaces} faces, {faces.nsamples} samples, {len(lights)} lights", flush=True)
    return g, faces, lights, kw, want


def check(got, want):
    ldr, hdr, ao = want
    for name, a, b in (("ldr", got.ldr, ldr), ("hdr", got.hdr, hdr)):
        diff = a.view(np.uint32) != b.view(np.uint32)
        assert not diff.any(), (name, int(diff.sum()), np.argwhere(diff)[:5].tolist())
    assert np.array_equal(got.ao, ao)


class _GpuOut:
    def __init__(self, ldr, hdr, ao, face_styles, rays):
        self.ldr, self.hdr, self.ao, self.face_styles, self.rays = ldr, hdr, ao, face_styles, rays


def gpu_run(faces, L, **kw):
    """The gather_cases runner on the device: slot 0 (style 0) and the rays
    the gather traced."""
    tr, fl = open_lighter(kw.get("casters", gc.far_scene()))
    with tr:
        fl.load(faces, L, light_pvs=kw.get("pvs"), sun_extent=kw.get("sun_extent", 0.0),
                sun_samples=kw.get("sun_samples", 1), sky_dirs=kw.get("sky_dirs"),
                ao_count=kw.get("ao_count", 0), ao_radius=kw.get("ao_radius", 40.0),
                skybox=kw.get("skybox"))
        before = tr.stats()["rays_gather"]
        got = fl.light_faces()
        rays = tr.stats()["rays_gather"] - before
    S = faces.nsamples
    ldr = np.zeros((4, 4, S, 4), np.float32)
    hdr = np.zeros((4, 4, S, 4), np.float32)
    ldr[0], hdr[0] = got.ldr, got.hdr
    fs = np.full((faces.nfaces, 4), 255, np.uint8)
    fs[:, 0] = 0
    return _GpuOut(ldr, hdr, got.ao, fs, rays)


@pytest.mark.parametrize("fact", sorted({f for f, _ in gc.FACTS}, key=lambda f: f.__name__),
                         ids=lambda f: f.__name__)
def test_gather_fact_on_the_device(fact):
    """Each of the reference's closed-form gather facts, on the GPU."""
    print(f"[fact] {fact.__name__}", flush=True)
    assert fact(gpu_run)


def test_spot_pow_rows_match_the_oracle_bitwise():
    """The planted spot cases (cone multiplier = the golden row's base) give
    the oracle's bits on every output channel."""
    rows = gc.spot_pow_rows()
    X = np.array([r[0] for r in rows], np.float32)
    Y = np.array([r[1] for r in rows], np.float32)
    faces, L, pvs = gc.spot_pow_case(X, Y)
    got = gpu_run(faces, L, pvs=pvs)
    d = light.pack_desc(faces, L, pvs, 0.0, 1, None)
    want = og.gather(gc.far_scene(), d, faces.nsamples)
    assert np.array_equal(got.ldr[0].view(np.uint32), want[0].view(np.uint32))
    assert np.array_equal(got.ldr[0, 0, :, 0], gc.spot_pow_expected(X, [r[2] for r in rows]))


@pytest.mark.parametrize("schedule", ["queue", "grid", "per_face"])
def test_gather_matches_oracle(c1, schedule):
    g, faces, lights, kw, want = c1
    print(f"[gather] schedule={schedule}", flush=True)
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(faces, lights, **kw)
        got = fl.light_faces(schedule=schedule)
    check(got, want)
    assert np.any(got.ldr[0, :, 3] > 0) and np.any(got.ldr[0, :, 3] == 0)  # sun both seen and shaded
    assert np.any(got.ldr[1:, :, 0] > 0)  # bump slots lit
    assert 0 < got.ao.min() < got.ao.max() <= 1


def test_schedules_bitwise_equal_and_rerun_is_idempotent(c1):
    g, faces, lights, kw, _ = c1
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(faces, lights, **kw)
        outs = [fl.light_faces(schedule=s, waves_per_cu=w, chunks_per_ticket=k)
                for s, w, k in (("queue", 0, 1), ("queue", 1, 1), ("queue", 7, 1), ("grid", 0, 1),
                                ("per_face", 0, 1), ("queue", 0, 1), ("queue", 4, 2), ("queue", 10, 7),
                                ("queue", 2, 64))]
        for k in (0, 65):
            with pytest.raises(ValueError):
                fl.run("queue", 4, k)
    for o in outs[1:]:
        assert np.array_equal(o.ldr.view(np.uint32), outs[0].ldr.view(np.uint32))
        assert np.array_equal(o.hdr.view(np.uint32), outs[0].hdr.view(np.uint32))
        assert np.array_equal(o.ao, outs[0].ao)


def test_lightmap_lump_matches_oracle_encode(c1):
    g, faces, lights, kw, want = c1
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(faces, lights, **kw)
        got = fl.light_faces()
        lump = fl.lightmap(hdr=False)
        lump_hdr = fl.lightmap(hdr=True)
    # the lump's order: face, normal slot, sample (BSP lighting lump layout)
    for which, lm, src in ((0, lump, got.ldr), (1, lump_hdr, got.hdr)):
        parts = []
        for f in range(faces.nfaces):
            a, b = faces.face_sample_ptr[f], faces.face_sample_ptr[f + 1]
            for k in range(faces.face_normal_count[f]):
                parts.append(src[k, a:b, :3])
        exp = og.encode_rgbe(np.concatenate(parts)).reshape(-1)
        assert np.array_equal(lm, exp), which


def test_encode_rgbe_matches_oracle():
    rng = np.random.default_rng(11)
    v = (rng.uniform(0, 1, (100_003, 3)) * 10.0 ** rng.uniform(-8, 8, (100_003, 1))).astype(np.float32)
    v[:50] = [[0, 0, 0], [1, 1, 1], [3, 3, 1], [-1, 2, 0], [np.nan, 1, 1], [np.inf, 0, 0],
              [1e38, 1e38, 1e38], [1e-40, 0, 0], [255, 255, 255], [256, 0, 0]] * 5
    g = rooms.room_grid(1, 1)
    tr, _ = open_lighter(g.casters)
    with tr:
        got = light.encode_rgbe(tr, v)
        wide = np.concatenate([v, np.ones((len(v), 1), np.float32)], axis=1)
        got4 = light.encode_rgbe(tr, wide)
    exp = og.encode_rgbe(v)
    assert np.array_equal(got, exp)
    assert np.array_equal(got4, exp)


def test_gather_batching_scene_subset():
    """Faces with ragged luxel counts (4 to 2,048) of the batching config's scene, sun
    plus sky ambient, queue against per-face launches and the oracle."""
    casters, faces = scenes.batching_workload(nfaces=96, min_luxels=4, max_luxels=2048)
    L = light.make_lights(2)
    L[0]["type"], L[0]["normal"], L[0]["intensity"] = light.EMIT_SKYLIGHT, (0.3, 0.2, -0.93), (1, 1, 1)
    L[1]["type"], L[1]["intensity"] = light.EMIT_SKYAMBIENT, (0.2, 0.2, 0.3)
    kw = dict(sun_extent=0.0, sun_samples=1, sky_dirs=light.sky_directions())
    d = light.pack_desc(faces, L, None, 0.0, 1, kw["sky_dirs"])
    want = og.gather(casters, d, faces.nsamples)
    tr, fl = open_lighter(casters)
    with tr:
        fl.load(faces, L, **kw)
        for s in ("queue", "per_face"):
            check(fl.light_faces(schedule=s), want)


def test_gather_errors():
    g = rooms.room_grid(1, 1)
    tr, fl = open_lighter(g.casters)
    with tr:
        with pytest.raises(Exception):
            fl.run()  # nothing loaded
        faces = rooms.room_faces(g, luxel=64.0)
        lights, pvs = rooms.room_lights(g)
        bad = lights.copy()
        bad[0]["type"] = 9
        with pytest.raises(Exception):
            fl.load(faces, bad, light_pvs=pvs)


def test_ray_count_matches_closed_form():
    """AO traces ao_count rays per sample; the sun traces sun_samples rays per
    sample of every face that faces it (a flat face: one dot for all samples)."""
    g = rooms.room_grid(2, 1, seed=3)
    faces = rooms.room_faces(g, luxel=64.0)
    L = light.make_lights(1)
    sun_dir = np.array([0.2, 0.3, -0.93], np.float32)
    L[0]["type"], L[0]["normal"], L[0]["intensity"] = light.EMIT_SKYLIGHT, sun_dir, (1, 1, 1)
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(faces, L, sun_extent=0.01, sun_samples=16, ao_count=8, ao_radius=40.0)
        _ms, n, rays = fl.timed_run("queue")
    per_face = np.diff(faces.face_sample_ptr)
    facing = (-(faces.face_normals[:, 0] @ sun_dir)) > 0
    assert n == 1
    assert rays == int(per_face.sum()) * 8 + int(per_face[facing].sum()) * 16
