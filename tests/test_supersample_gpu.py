"""Supersampling on the GPU (vh_gather_supersample) against the oracle
(oracle/supersample.py): every refined value bit-equal, LDR and HDR, flat and
bumped faces, the Correct and stock readings of a luxel without a sample,
and the same number of samples re-lit each pass.  Needs an MI355X."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.gather as og
import oracle.supersample as oss
from hl2sdk_tools_sharp_amd import HipTracerFactory, light, rooms
from hl2sdk_tools_sharp_amd.final import FACE_DTYPE
from hl2sdk_tools_sharp_amd.light import FaceSet, HipFaceLighter
from hl2sdk_tools_sharp_amd.scenes import box_triangles
from hl2sdk_tools_sharp_amd.tracer import TRACE_ID_OPAQUE, ShadowCasterSet
from hl2sdk_tools_sharp_amd.vrad import light_pass

pytestmark = pytest.mark.gpu


def open_lighter(casters):
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer, HipFaceLighter(offer.tracer)


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def rooms_case():
    g = rooms.room_grid(2, 2, tile=64.0, seed=2)
    fset = rooms.room_faces(g, 16.0, bumped_every=5)
    lights, pvs = rooms.room_lights(g, sun=True, sky_ambient=True)
    for k in ("intensity", "intensity_hdr"):
        lights[k] *= 100  # bright enough for shadow edges to pass the gradient test
    ff, _ = rooms.room_final_inputs(g, fset)
    sky = light.sky_directions()
    ldr, hdr, _ = og.gather(g.casters, light.pack_desc(fset, lights, pvs, 0.0, 1, sky), fset.nsamples)
    print(f"[supersample] rooms: {fset.nfaces} faces, {fset.nsamples} samples, {len(lights)} lights", flush=True)
    return g, fset, lights, pvs, ff.faces, sky, ldr, hdr


@pytest.mark.parametrize("which", [0, 1])
def test_rooms_match_oracle(rooms_case, which):
    g, fset, lights, pvs, frames, sky, ldr, hdr = rooms_case
    want, want_counts = oss.supersample(g.casters, fset, lights, pvs, 0.0, 1, sky, frames, hdr if which else ldr,
                                        which=which)
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(fset, lights, light_pvs=pvs, sun_samples=1, sky_dirs=sky)
        fl.run()
        base = fl.read()
        assert same(base.ldr, ldr) and same(base.hdr, hdr)
        counts = fl.supersample(frames, which=which)
        got = fl.read()
    print(f"[supersample] which={which} passes {counts}", flush=True)
    assert counts == want_counts and len(counts) >= 1
    assert same(got.hdr if which else got.ldr, want)
    assert same(got.ldr if which else got.hdr, ldr if which else hdr)  # the other result untouched


def _holey_case():
    """The floor of 8 x 8 luxels of 16 units under a point light and a slab
    (a shadow edge across it), four luxels without a sample, the samples
    listed in reverse luxel order with their luxels given."""
    v9, ids = box_triangles(np.array([[-100.0, -100.0, 50.0]]), np.array([[42.0, 300.0, 60.0]]))
    casters = ShadowCasterSet(v9, (ids & 0xFFFFFF) | TRACE_ID_OPAQUE)
    keep = np.array([k for k in range(63, -1, -1) if k not in (10, 27, 36, 50)])
    st = np.stack([keep % 8, keep // 8], 1).astype(np.int32)
    pos = np.stack([8 + 16 * st[:, 0], 8 + 16 * st[:, 1], np.zeros(len(keep))], 1).astype(np.float32)
    normals = np.zeros((1, 4, 3), np.float32)
    normals[0, 0] = (0, 0, 1)
    n = len(keep)
    fset = FaceSet(np.array([0, n], np.int32), np.array([1], np.int32), normals, pos, np.zeros(n, np.int32))
    L = light.make_lights(1)
    L["type"], L["origin"], L["quadratic_attn"] = light.EMIT_POINT, (64, 64, 200), 1
    L["intensity"] = L["intensity_hdr"] = (5e6, 4e6, 3e6)
    fr = np.zeros(1, FACE_DTYPE)
    fr["origin"] = (8, 8, 0)
    fr["w2l_s"], fr["w2l_t"] = (1 / 16, 0, 0), (0, 1 / 16, 0)
    fr["l2w_s"], fr["l2w_t"] = (16, 0, 0), (0, 16, 0)
    fr["width"] = fr["height"] = 8
    fr["lit"] = 1
    return casters, fset, L, fr, st


@pytest.mark.parametrize("garbage", [False, True])
def test_given_luxels_and_holes_match_oracle(garbage):
    casters, fset, L, fr, st = _holey_case()
    ldr, _, _ = og.gather(casters, light.pack_desc(fset, L, None, 0.0, 1, None), fset.nsamples)
    want, want_counts = oss.supersample(casters, fset, L, None, 0.0, 1, None, fr, ldr, sample_st=st,
                                        garbage=garbage)
    tr, fl = open_lighter(casters)
    with tr:
        fl.load(fset, L, sun_samples=1)
        fl.run()
        counts = fl.supersample(fr, sample_st=st, unset_is_garbage=garbage)
        got = fl.read()
    assert counts == want_counts
    assert same(got.ldr, want)
    if garbage:  # every neighbour of a hole reads infinity: all selected at once
        assert counts[0] >= 20


def test_supersample_errors():
    casters, fset, L, fr, st = _holey_case()
    tr, fl = open_lighter(casters)
    with tr:
        fl.load(fset, L, sun_samples=1)
        with pytest.raises(Exception, match="vh_gather_run"):
            fl.supersample(fr, sample_st=st)
        fl.run()
        with pytest.raises(ValueError):
            fl.supersample(fr, which=2, sample_st=st)
        with pytest.raises(ValueError):
            fl.supersample(fr, extra_passes=65, sample_st=st)
        bad = st.copy()
        bad[0] = (8, 0)
        with pytest.raises(ValueError, match="luxel"):
            fl.supersample(fr, sample_st=bad)
        bad = st.copy()
        bad[1] = bad[0]
        with pytest.raises(ValueError, match="luxel"):
            fl.supersample(fr, sample_st=bad)
        assert fl.supersample(fr, extra_passes=0, sample_st=st) == []


def test_light_pass_supersamples_before_the_patch_handoff(rooms_case):
    """FaceLightJob.Finish builds the patch lights from the supersampled
    light: a pass with extra passes hands the patches what a gather plus
    supersampling gives, after the oracle's passes."""
    from hl2sdk_tools_sharp_amd.bounce import HipBounceStage

    g, fset, lights, pvs, frames, sky, ldr, hdr = rooms_case
    fin = rooms.room_final_inputs(g, fset)
    r = light_pass(g, fset, lights, pvs, sun_samples=1, sky_dirs=sky, tol=1e-3, read_direct=True,
                   final_inputs=fin, extra_passes=4)
    _, counts = oss.supersample(g.casters, fset, lights, pvs, 0.0, 1, sky, frames, ldr, which=0)
    assert r.supersample_passes == counts and counts
    tr, fl = open_lighter(g.casters)
    with tr:
        fl.load(fset, lights, light_pvs=pvs, sun_samples=1, sky_dirs=sky)
        fl.run()
        fl.supersample(frames, which=0)
        st = HipBounceStage(tr, g)
        st.load_patches()
        direct = st.patch_lights(fl, hdr=False, read=True)
    assert same(r.direct, direct)


def _ragged_case(seed=11, nfaces=23):
    """Floor faces of ragged sizes (3 to 700 samples, 8-unit luxels) laid side
    by side under a point light and a slab: shadow edges cross many faces,
    and the faces straddle the selection kernel's 256-sample blocks."""
    rng = np.random.default_rng(seed)
    v9, ids = box_triangles(np.array([[-50.0, 40.0, 60.0]]), np.array([[2000.0, 120.0, 70.0]]))
    casters = ShadowCasterSet(v9, (ids & 0xFFFFFF) | TRACE_ID_OPAQUE)
    fr = np.zeros(nfaces, FACE_DTYPE)
    ptr, pos = [0], []
    x0 = 0.0
    for f in range(nfaces):
        w = int(rng.integers(1, 36))
        h = int(rng.integers(3, 21))
        if f % 5 == 2:
            w, h = 1, 3  # tiny faces among the large ones
        fr[f]["origin"] = (x0 + 4, 4, 0)
        fr[f]["w2l_s"], fr[f]["w2l_t"] = (1 / 8, 0, 0), (0, 1 / 8, 0)
        fr[f]["l2w_s"], fr[f]["l2w_t"] = (8, 0, 0), (0, 8, 0)
        fr[f]["width"], fr[f]["height"], fr[f]["lit"] = w, h, 1
        for t in range(h):
            for s in range(w):
                pos.append((x0 + 4 + 8 * s, 4 + 8 * t, 0.0))
        ptr.append(len(pos))
        x0 += 8 * w + 16
    n = len(pos)
    normals = np.zeros((nfaces, 4, 3), np.float32)
    normals[:, 0] = (0, 0, 1)
    fset = FaceSet(np.array(ptr, np.int32), np.ones(nfaces, np.int32), normals, np.array(pos, np.float32),
                   np.zeros(n, np.int32))
    L = light.make_lights(1)
    L["type"], L["origin"], L["quadratic_attn"] = light.EMIT_POINT, (x0 / 2, 80, 400), 1
    L["intensity"] = L["intensity_hdr"] = (2e7, 1.5e7, 1e7)
    return casters, fset, L, fr


def test_ragged_faces_straddling_blocks_match_oracle():
    casters, fset, L, fr = _ragged_case()
    assert fset.nsamples > 2048 and len(set(np.diff(fset.face_sample_ptr).tolist())) > 10
    ldr, _, _ = og.gather(casters, light.pack_desc(fset, L, None, 0.0, 1, None), fset.nsamples)
    want, want_counts = oss.supersample(casters, fset, L, None, 0.0, 1, None, fr, ldr)
    tr, fl = open_lighter(casters)
    with tr:
        fl.load(fset, L, sun_samples=1)
        fl.run()
        counts = fl.supersample(fr)
        got = fl.read()
    print(f"[supersample] ragged: {fset.nfaces} faces, {fset.nsamples} samples, passes {counts}", flush=True)
    assert counts == want_counts and counts[0] > 20
    assert same(got.ldr, want)


def test_sky_ambient_alone_skips_the_direct_resample_and_matches_oracle(rooms_case):
    """A scene lit only by sky ambient: the direct kind of the resample has no
    light (it is not launched), the ambient points carry everything."""
    g, fset, _lights, pvs, frames, sky, _ldr, _hdr = rooms_case
    L = light.make_lights(1)
    L["type"] = light.EMIT_SKYAMBIENT
    L["intensity"] = L["intensity_hdr"] = (4000.0, 6000.0, 9000.0)
    g2 = rooms.room_grid(2, 2, tile=64.0, seed=2, sky_ceiling=True)
    f2 = rooms.room_faces(g2, 16.0, bumped_every=5)
    ff2, _ = rooms.room_final_inputs(g2, f2)
    ldr, hdr, _ = og.gather(g2.casters, light.pack_desc(f2, L, None, 0.0, 1, sky), f2.nsamples)
    want, want_counts = oss.supersample(g2.casters, f2, L, None, 0.0, 1, sky, ff2.faces, hdr, which=1)
    tr, fl = open_lighter(g2.casters)
    with tr:
        fl.load(f2, L, sun_samples=1, sky_dirs=sky)
        fl.run()
        counts = fl.supersample(ff2.faces, which=1)
        got = fl.read()
    print(f"[supersample] sky ambient alone: passes {counts}", flush=True)
    assert counts == want_counts and len(counts) >= 1
    assert same(got.hdr, want)
