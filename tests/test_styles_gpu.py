"""Light styles on the GPU against the oracle.  Needs an MI355X.

Room maps whose lights use several styles (switchable and animated lights:
styles 0, 3, 7, 12 and 33, more than a face's four slots can hold in places):
* the gather's style slots per face (FaceLight.FindOrAllocateStyle in first-
  touch order) and every slot's light, LDR and HDR, bit-equal to the oracle's;
* supersampling per slot (one StyleSupersample per style) bit-equal, with the
  same number of samples re-lit per pass;
* the final light's lump -- every style's luxels, the medians before lightofs
  in reverse style order, neighbours matched by style number, bounced light
  in style 0 only -- byte-equal to the C restatement's, on the gather's own
  slots, with and without a bounce, and through a whole lighting pass."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.final as of
import oracle.gather as og
import oracle.supersample as oss
from hl2sdk_tools_sharp_amd import HipTracerFactory, light, rooms
from hl2sdk_tools_sharp_amd.final import HipFinalLighter
from hl2sdk_tools_sharp_amd.light import HipFaceLighter
from hl2sdk_tools_sharp_amd.vrad import light_pass

pytestmark = pytest.mark.gpu

STYLES = (0, 3, 7, 0, 12, 33, 3, 7)


def styled_lights(g, **kw):
    lights, pvs = rooms.room_lights(g, **kw)
    npoint = len(g.lights)
    for k in range(npoint):
        lights[k]["style"] = STYLES[k % len(STYLES)]
    return lights, pvs


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def open_lighter(casters):
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer, HipFaceLighter(offer.tracer)


@pytest.fixture(scope="module")
def styled_map():
    g = rooms.room_grid(3, 2, sky_ceiling=True, seed=7, nlights=12)
    faces = rooms.room_faces(g, luxel=32.0, bumped_every=5)
    lights, pvs = styled_lights(g, sun=True, sky_ambient=True, spot_every=3, surface_every=4)
    lights[len(g.lights)]["style"] = 3  # the sun shares a switchable style
    lights[len(g.lights) + 1]["style"] = 45  # sky ambient reaches every face: a fifth style in places
    kw = dict(light_pvs=pvs, sun_extent=0.02, sun_samples=30, sky_dirs=light.sky_directions())
    d = light.pack_desc(faces, lights, pvs, kw["sun_extent"], kw["sun_samples"], kw["sky_dirs"])
    want = og.gather_styles(g.casters, d, faces.nsamples)
    nst = (want.face_styles != 255).sum(1)
    print(f"[styles] {faces.nfaces} faces, {faces.nsamples} samples, {len(lights)} lights; faces with "
          f"1/2/3/4 styles: {[int((nst == k).sum()) for k in (1, 2, 3, 4)]}; dropped {want.dropped}", flush=True)
    assert (nst >= 3).any() and want.dropped > 0  # the case covers full slots and a fifth style
    return g, faces, lights, kw, want


def test_gather_slots_match_oracle(styled_map):
    g, faces, lights, kw, want = styled_map
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(faces, lights, **kw)
        for schedule in ("queue", "grid"):
            fl.run(schedule)
            fs, ldr, hdr, dropped = fl.read_styles()
            print(f"[styles] {schedule}: dropped {dropped}", flush=True)
            assert np.array_equal(fs, want.face_styles), schedule
            assert same(ldr, want.ldr) and same(hdr, want.hdr), schedule
            assert dropped > 0
            base = fl.read()
            assert same(base.ldr, want.ldr[0]) and np.array_equal(base.ao, want.ao)


@pytest.mark.parametrize("which", [0, 1])
def test_supersampling_per_style_matches_oracle(which):
    g = rooms.room_grid(2, 2, tile=64.0, seed=2)
    fset = rooms.room_faces(g, 16.0, bumped_every=5)
    lights, pvs = styled_lights(g, sun=True, sky_ambient=True)
    for k in ("intensity", "intensity_hdr"):
        lights[k] *= 100  # shadow edges past the gradient threshold
    ff, _ = rooms.room_final_inputs(g, fset)
    sky = light.sky_directions()
    d = light.pack_desc(fset, lights, pvs, 0.0, 1, sky)
    base = og.gather_styles(g.casters, d, fset.nsamples)
    want, want_counts = oss.supersample_styles(g.casters, fset, lights, pvs, 0.0, 1, sky, ff.faces,
                                               base.hdr if which else base.ldr, base.face_styles, which=which)
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(fset, lights, light_pvs=pvs, sun_samples=1, sky_dirs=sky)
        fl.run()
        counts = fl.supersample(ff.faces, which=which)
        fs, ldr, hdr, _ = fl.read_styles()
    print(f"[styles] supersample which={which}: passes {counts}, oracle {want_counts}", flush=True)
    assert np.array_equal(fs, base.face_styles)
    assert counts == want_counts and len(counts) >= 1
    assert same(hdr if which else ldr, want)


def _final_inputs_c(fin, faces):
    ff, fp = fin
    pd = {"origin": fp.origin, "face": fp.face, "child1": fp.child1, "winding_ptr": fp.winding_ptr,
          "winding": fp.winding}
    return ff, fp, pd


@pytest.mark.parametrize("bounce", [False, True])
def test_final_lump_per_style_matches_the_c_restatement(styled_map, bounce):
    g, faces, lights, kw, _want = styled_map
    fin = rooms.room_final_inputs(g, faces)
    ff, fp, pd = _final_inputs_c(fin, faces)
    rng = np.random.default_rng(4)
    total = rng.uniform(0, 20, (fp.npatch, 4, 3)).astype(np.float32)
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(faces, lights, **kw)
        fl.run()
        fs, ldr, _hdr, _ = fl.read_styles()
        fr = HipFinalLighter(fl)
        fr.load(ff, fp)
        offs, size = fr.layout()
        lump = fr.run(bounce="host", patch_light=total) if bounce else fr.run()
    want, woffs = of.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                                   faces.face_sample_ptr, faces.face_normal_count, faces.sample_pos, ldr,
                                   pd if bounce else None, total if bounce else None, face_styles=fs)
    print(f"[styles] final lump {size} bytes, bounce={bounce}", flush=True)
    assert list(offs) == list(woffs)
    assert lump.shape == want.shape and np.array_equal(lump, want)


def test_a_lighting_pass_with_styles_matches_the_oracle():
    """The whole pass: gather (styles), patch lights (style 0), transfers,
    bounce, final light with the bounce on the device; the lump equals the C
    final stage run on the oracle's own gather with the GPU's bounce."""
    g = rooms.room_grid(3, 3, size=(512, 512, 256), wall=8, door=128, patch=128, minchop=64, nlights=8,
                        sky_ceiling=True, seed=1)
    faces = rooms.room_faces(g, luxel=16.0)
    lights, pvs = styled_lights(g, sun=True)
    fin = rooms.room_final_inputs(g, faces)
    r = light_pass(g, faces, lights, pvs, sun_extent=0.0, sun_samples=1, tol=1e-4, final_inputs=fin)
    d = light.pack_desc(faces, lights, pvs, 0.0, 1, None)
    o = og.gather_styles(g.casters, d, faces.nsamples)
    ff, fp, pd = _final_inputs_c(fin, faces)
    pl = np.zeros((fp.npatch, 4, 3), np.float32)
    pl[:, 0] = r.bounce.total_light
    want, _ = of.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                               faces.face_sample_ptr, faces.face_normal_count, faces.sample_pos, o.ldr, pd, pl,
                               face_styles=o.face_styles)
    nst = (o.face_styles != 255).sum(1)
    print(f"[styles] pass: lump {r.lump.shape[0]} bytes; faces with 1/2/3/4 styles "
          f"{[int((nst == k).sum()) for k in (1, 2, 3, 4)]}", flush=True)
    assert (nst >= 2).any()
    assert np.array_equal(r.lump, want)
