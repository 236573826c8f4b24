"""The patch-light hand-off and a whole lighting pass on the GPU against the
oracle.  Needs an MI355X.

Patch lights: bit-equal to the oracle's BuildPatchLights restatement fed the
GPU gather's own luxel light (the gather itself is pinned separately).  The
pass: the device-resident hand-off gives the same bounce, bit for bit, as
passing the patch lights through the host, and the bounced light agrees with
the oracle's FP32 bounce on the oracle's transfers to 1e-5."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.bounce as ob
import oracle.gather as og
from hl2sdk_tools_sharp_amd import HipTracerFactory, light, rooms
from hl2sdk_tools_sharp_amd.bounce import HipBounceStage
from hl2sdk_tools_sharp_amd.light import HipFaceLighter
from hl2sdk_tools_sharp_amd.vrad import light_pass

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_small():
    """A 2x2 grid of the plain-vrad rooms (512x512x256, 128 doorways, patch
    128) under a sky ceiling: 8 point lights and a sun, luxel 32."""
    g = rooms.room_grid(2, 2, size=(512, 512, 256), patch=128, minchop=64, nlights=8, sky_ceiling=True, seed=1)
    faces = rooms.room_faces(g, luxel=32.0)
    lights, pvs = rooms.room_lights(g, sun=True)
    print(f"[c1_small] {g.npatch} patches, {faces.nsamples} luxels, {len(lights)} lights", flush=True)
    return g, faces, lights, pvs


def test_patch_lights_match_oracle(c1_small):
    g, faces, lights, pvs = c1_small
    offer = HipTracerFactory().try_create(g.casters)
    assert offer.tracer is not None, offer.decline_reason
    with offer.tracer as tr:
        fl = HipFaceLighter(tr)
        fl.load(faces, lights, light_pvs=pvs, sun_samples=1)
        res = fl.light_faces()
        st = HipBounceStage(tr, g)
        got = st.patch_lights(fl)
        got_hdr = st.patch_lights(fl, hdr=True)
    for src, out in ((res.ldr, got), (res.hdr, got_hdr)):
        want = ob.patch_lights(g, faces, src[0], faces.sample_area)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    assert (got > 0).any() and np.all(np.isfinite(got))
    np.testing.assert_array_equal(got_hdr > 0, got > 0)


def test_patch_lights_errors(c1_small):
    g, faces, lights, pvs = c1_small
    offer = HipTracerFactory().try_create(g.casters)
    with offer.tracer as tr:
        st = HipBounceStage(tr, g)
        fl = HipFaceLighter(tr)
        with pytest.raises(Exception):
            st.bounce(None)  # no patch lights and no transfers yet
        other = rooms.room_faces(rooms.room_grid(1, 1), luxel=64.0)
        fl.load(other, lights[:1])
        fl.run()
        with pytest.raises(Exception):
            st.patch_lights(fl)  # different faces


def test_pass_device_handoff_and_oracle_bounce(c1_small):
    g, faces, lights, pvs = c1_small
    r = light_pass(g, faces, lights, pvs, sun_samples=1, tol=1e-4, read_direct=True)
    print(f"[pass] stages {r.stage_s}, {r.transfers} transfers, {r.bounce.iterations} iterations", flush=True)
    assert r.bounce.iterations > 1 and r.bounce.residuals[-1] <= 1e-4
    # host hand-off of the same patch lights: the same bounce
    offer = HipTracerFactory().try_create(g.casters)
    with offer.tracer as tr:
        st = HipBounceStage(tr, g)
        st.build_transfers()
        rp, src, w = st.read_transfers()
        host = st.bounce(r.direct, tol=1e-4)
    assert host.iterations == r.bounce.iterations
    assert np.array_equal(host.total_light.view(np.uint32), r.bounce.total_light.view(np.uint32))
    tot, res, _en, it = ob.bounce(g, rp, src, w, r.direct, tol=1e-4, max_iter=100)
    assert it == r.bounce.iterations
    np.testing.assert_allclose(r.bounce.total_light, tot, rtol=1e-5, atol=1e-6)
    # the lump: the gather's LDR through the encoder, face by face
    d = light.pack_desc(faces, lights, pvs, 0.0, 1, None)
    ldr, _hdr, _ao = og.gather(g.casters, d, faces.nsamples)
    parts = []
    for f in range(faces.nfaces):
        a, b = faces.face_sample_ptr[f], faces.face_sample_ptr[f + 1]
        parts.append(ldr[0, a:b, :3])
    assert np.array_equal(r.lump, og.encode_rgbe(np.concatenate(parts)).reshape(-1))
