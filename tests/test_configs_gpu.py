"""Each workload at its own generator and parameters, on the GPU, against the
oracle.  Needs an MI355X.

* Plain vrad (3 x 3 rooms of 512 x 512 x 256, 8-unit walls, 128 doorways, 8
  point lights of intensity 50-300, one sun, luxel 16, patch 128, FP32,
  residual 1e-4): the patch lights bit-equal to the oracle's; the bounce
  within 1e-5 of the oracle's FP32 bounce and 1e-3 of its FP64 one; the lump
  byte-equal to the oracle's final stage on the same bounce.
* Bounce (12 x 12 rooms, 200,000 +- 2 % patches at patch size 64, about 1,500
  transfers per row): FP32 and FP16(+row scale) weights with FP32
  accumulation, each iterated to residual 1e-4, within 1e-3 (relative L1) of
  the oracle's FP64 bounce on the same transfers, to the same residual and
  converged.
* Batching (4,096 faces with luxel counts log-uniform over 4-16,384; per
  luxel 162 sky-ambient, 64 jittered sun-disk (0.5 degrees) and 32 AO rays):
  the persistent queue and per-face launches bit-equal to each other and, on
  a 1.5 % face subset, to the oracle.
* Learned warm start (maps of the bounce config's generator, 50k patches
  each): the network's gradient at batch 4,096 against the oracle, and on a
  held-out seed from 200-207 after training on three training seeds, fewer
  iterations than the zero start and no more than the direct-only start.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle.bounce as ob
import oracle.final as of
import oracle.gather as og
import oracle.mlp as om
from hl2sdk_tools_sharp_amd import HipTracerFactory, light, rooms, scenes
from hl2sdk_tools_sharp_amd.bounce import HipBounceStage
from hl2sdk_tools_sharp_amd.light import HipFaceLighter
from hl2sdk_tools_sharp_amd.tracer import unpack_bits
from hl2sdk_tools_sharp_amd.vrad import light_pass

pytestmark = pytest.mark.gpu


def rel_l1(a, b, mask=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def final_lump_cpu(fin, faces, ldr, total):
    ff, fp = fin
    pl = np.zeros((fp.npatch, 4, 3), np.float32)
    pl[:, 0] = total
    pd = {"origin": fp.origin, "face": fp.face, "child1": fp.child1, "winding_ptr": fp.winding_ptr,
          "winding": fp.winding}
    lump, _ = of.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                               faces.face_sample_ptr, faces.face_normal_count, faces.sample_pos, ldr, pd, pl)
    return lump


def test_plain_vrad_config():
    g = rooms.room_grid(3, 3, size=(512, 512, 256), wall=8, door=128, patch=128, minchop=64, nlights=8,
                        sky_ceiling=True, seed=1)
    faces = rooms.room_faces(g, luxel=16.0)
    lights, pvs = rooms.room_lights(g, sun=True)
    fin = rooms.room_final_inputs(g, faces)
    print(f"[plain vrad] {g.npatch} patches, {faces.nsamples} luxels, {len(lights)} lights", flush=True)
    assert len(lights) == 9 and np.all((lights["intensity"][:8] >= 50) & (lights["intensity"][:8] <= 300))
    r = light_pass(g, faces, lights, pvs, sun_extent=0.0, sun_samples=1, tol=1e-4, final_inputs=fin,
                   read_direct=True)
    d = light.pack_desc(faces, lights, pvs, 0.0, 1, None)
    ldr, _hdr, _ao = og.gather(g.casters, d, faces.nsamples)
    direct = ob.patch_lights(g, faces, ldr[0], faces.sample_area)
    assert np.array_equal(direct.view(np.uint32), r.direct.view(np.uint32))
    rp, src, w, _nc = ob.build_transfers(g)
    assert int(rp[-1]) == r.transfers
    tot32, _res, _en, it32 = ob.bounce(g, rp, src, w, direct, tol=1e-4, max_iter=100)
    tot64, _res, _en, it64 = ob.bounce(g, rp, src, w, direct, tol=1e-4, max_iter=100, use_double=True)
    gpu = r.bounce.total_light
    print(f"[plain vrad] iterations GPU {r.bounce.iterations}, oracle FP32 {it32}, FP64 {it64}; "
          f"rel L1 vs FP32 {rel_l1(gpu, tot32):.2e}, vs FP64 {rel_l1(gpu, tot64):.2e}", flush=True)
    assert r.bounce.iterations == it32
    assert rel_l1(gpu, tot32) < 1e-5
    assert rel_l1(gpu, tot64) < 1e-3
    assert np.array_equal(final_lump_cpu(fin, faces, ldr, gpu), r.lump)


@pytest.fixture(scope="module")
def bounce_config():
    g = rooms.bounce_grid(12, 12)
    offer = HipTracerFactory().try_create(g.casters)
    assert offer.tracer is not None, offer.decline_reason
    tr = offer.tracer
    st = HipBounceStage(tr, g)
    st.load_patches()
    T = st.build_transfers()
    pts = rooms.sample_points(g)
    bits = tr.trace_point_lights(pts, g.lights)
    e0 = rooms.direct_light(g, unpack_bits(bits, g.npatch * len(g.lights)))
    print(f"[bounce] {g.npatch} patches, {T} transfers, {T / len(g.receivers):.0f} per row", flush=True)
    yield g, tr, st, T, e0
    tr.dispose()


def test_bounce_config_shape(bounce_config):
    g, _tr, _st, T, _e0 = bounce_config
    assert abs(g.npatch / 200_000 - 1) <= 0.02
    assert 1_400 <= T / len(g.receivers) <= 1_600


def test_bounce_config_fp32_and_fp16_against_fp64(bounce_config):
    g, _tr, st, _T, e0 = bounce_config
    r32 = st.bounce(e0, weights="f32", tol=1e-4, max_iter=200)
    r16 = st.bounce(e0, weights="f16", tol=1e-4, max_iter=200)
    rp, src, w = st.read_transfers()
    same, _res, _en, it_same = ob.bounce(g, rp, src, w, e0, tol=1e-4, max_iter=200, use_double=True)
    conv, _res, _en, it_conv = ob.bounce(g, rp, src, w, e0, tol=1e-7, max_iter=400, use_double=True)
    leaf = g.leaf & ~g.sky.astype(bool)
    rows = {}
    for name, r in (("f32", r32), ("f16", r16)):
        rows[name] = (r.iterations, rel_l1(r.total_light, same, leaf), rel_l1(r.total_light, conv, leaf))
    print(f"[bounce] FP64 oracle {it_same} iterations to 1e-4, {it_conv} to 1e-7; "
          + "; ".join(f"{k}: {v[0]} iterations, rel L1 {v[1]:.2e} / {v[2]:.2e}" for k, v in rows.items()),
          flush=True)
    for name, (it, e_same, e_conv) in rows.items():
        assert abs(it - it_same) <= 1, name
        assert e_same < 1e-3 and e_conv < 1e-3, name
    assert rows["f32"][1] < 1e-5  # FP32 is the FP64 iteration up to float rounding


def test_batching_config_queue_per_face_and_oracle():
    casters, faces = scenes.batching_workload(nfaces=4096, seed=4)
    assert 7.0e6 < faces.nsamples < 9.0e6
    rng = np.random.default_rng(4)
    z = rng.uniform(0.0, 1.0)
    phi = rng.uniform(0.0, 2 * math.pi)
    to_sun = np.array([math.sqrt(1 - z * z) * math.cos(phi), math.sqrt(1 - z * z) * math.sin(phi), z])
    L = light.make_lights(2)
    L[0]["type"], L[0]["normal"] = light.EMIT_SKYLIGHT, -to_sun
    L[0]["intensity"], L[0]["intensity_hdr"] = (1.0, 0.9, 0.8), (2.0, 1.8, 1.6)
    L[1]["type"], L[1]["intensity"], L[1]["intensity_hdr"] = light.EMIT_SKYAMBIENT, (0.2, 0.25, 0.3), (0.4, 0.5, 0.6)
    kw = dict(sun_extent=math.sin(math.radians(0.5)), sun_samples=64, sky_dirs=light.sky_directions(),
              ao_count=32, ao_radius=40.0)
    assert len(kw["sky_dirs"]) == 162
    sub = np.sort(np.random.default_rng(5).choice(faces.nfaces, 61, replace=False))
    fs = faces.subset(sub)
    print(f"[batching] {faces.nfaces} faces, {faces.nsamples} luxels; subset {fs.nfaces} faces, "
          f"{fs.nsamples} luxels", flush=True)
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is not None, offer.decline_reason
    with offer.tracer as tr:
        fl = HipFaceLighter(tr)
        fl.load(faces, L, **kw)
        q = fl.light_faces(schedule="queue")
        qs = (q.ldr[:, sub_index(faces, sub)], q.hdr[:, sub_index(faces, sub)], q.ao[sub_index(faces, sub)])
        pf = fl.light_faces(schedule="per_face")
        assert np.array_equal(pf.ldr.view(np.uint32), q.ldr.view(np.uint32))
        assert np.array_equal(pf.hdr.view(np.uint32), q.hdr.view(np.uint32))
        assert np.array_equal(pf.ao, q.ao)
        del q, pf
    d = light.pack_desc(fs, L, None, kw["sun_extent"], kw["sun_samples"], kw["sky_dirs"], kw["ao_count"],
                        kw["ao_radius"])
    ldr, hdr, ao = og.gather(casters, d, fs.nsamples)
    assert np.array_equal(qs[0].view(np.uint32), ldr.view(np.uint32))
    assert np.array_equal(qs[1].view(np.uint32), hdr.view(np.uint32))
    assert np.array_equal(qs[2], ao)
    assert np.any(ldr[0, :, 3] > 0) and np.any(ao < 1)


def sub_index(faces, sub):
    lo = faces.face_sample_ptr[sub].astype(np.int64)
    hi = faces.face_sample_ptr[sub + 1].astype(np.int64)
    return np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)])


def test_warm_start_config_gradient_at_batch_4096():
    from hl2sdk_tools_sharp_amd.mlp import WarmStartMLP, init_params

    g = rooms.room_grid(2, 2, size=(384, 320, 192), patch=128, minchop=64, seed=5)
    offer = HipTracerFactory().try_create(g.casters)
    with offer.tracer as tr:
        net = WarmStartMLP(tr, params=init_params(3, out_bias=[0.2, 0.1, 0.0], out_scale=1.0))
        rng = np.random.default_rng(4)
        x = rng.standard_normal((4096, 25)).astype(np.float32)
        t = rng.standard_normal((4096, 3)).astype(np.float32)
        loss, grad = net.grad(x, t)
        wl, wg = om.loss_and_grad(net.params(), x, t, emulate=True)
    assert loss == pytest.approx(wl, rel=1e-4)
    for (gw, gb), (ww, wb) in zip(om.unpack(grad), om.unpack(wg)):
        assert np.linalg.norm(gw - ww) < 1e-3 * np.linalg.norm(ww)
        assert np.linalg.norm(gb - wb) < 1e-3 * np.linalg.norm(wb)


def test_warm_start_config_held_out_seed():
    from hl2sdk_tools_sharp_amd.mlp import Prior, WarmStartMLP, bounce_map, init_params

    train = [rooms.bounce_grid(6, 6, seed=s) for s in (105, 112, 119)]
    held_out = rooms.bounce_grid(6, 6, seed=203)
    assert all(45_000 <= m.npatch <= 60_000 for m in train + [held_out])
    offer = HipTracerFactory().try_create(train[0].casters)
    with offer.tracer as tr:
        net = WarmStartMLP(tr)
        net.dataset_clear()
        prior = Prior()
        for g in train:
            _st, e0, r = bounce_map(tr, g)
            net.dataset_add()
            prior.add(g, e0, r.total_light)
        net = WarmStartMLP(tr, params=init_params(0), gain=prior.gain, scale=prior.scale)
        losses = net.train(epochs=20, batch=4096, lr=2e-3, seed=1)
        st, e0, zero = bounce_map(tr, held_out, tol=1e-4)
        net.predict(e0)
        warm = st.bounce(e0, tol=1e-4, max_iter=200, from_prediction=True)
        leaf = held_out.leaf & ~held_out.sky.astype(bool)
        direct_only = st.bounce(e0, start=(prior.gain * e0 * leaf[:, None]).astype(np.float32), tol=1e-4,
                                max_iter=200)
    print(f"[warm start] loss {losses[0]:.4f} -> {losses[-1]:.4f}; held-out seed 203 iterations: zero "
          f"{zero.iterations}, direct-only {direct_only.iterations}, network {warm.iterations}", flush=True)
    assert warm.iterations < zero.iterations
    assert warm.iterations <= direct_only.iterations
    assert rel_l1(warm.total_light, zero.total_light, leaf) < 1e-3
