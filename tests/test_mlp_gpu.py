"""The warm-start network's HIP kernels against the oracle, and what the
trained network gains.  Needs an MI355X.

Forward, gradient and Adam are held to the oracle run with bfloat16 rounding
at the points where the GPU stores bfloat16 (what is left is FP32 against
FP64 accumulation): relative L2 error 1e-3 per layer.  Training is
deterministic (a replayed graph gives the same weights as direct launches).
The gain: trained on seeded maps, the network's start reaches the bounce's
residual 1e-4 in fewer iterations than the zero start on a held-out map, in
no more than the direct-only start (the prior alone), and lands on the same
answer (relative L1 1e-3)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.mlp as om
from hl2sdk_tools_sharp_amd import HipTracerFactory, rooms
from hl2sdk_tools_sharp_amd.bounce import HipBounceStage
from hl2sdk_tools_sharp_amd.mlp import Prior, WarmStartMLP, bounce_map, init_params

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def tracer():
    g = rooms.room_grid(2, 2, size=(384, 320, 192), patch=128, minchop=64, seed=5)
    offer = HipTracerFactory().try_create(g.casters)
    assert offer.tracer is not None, offer.decline_reason
    yield g, offer.tracer
    offer.tracer.dispose()


def test_forward_matches_oracle(tracer):
    _g, tr = tracer
    net = WarmStartMLP(tr, params=init_params(1, out_bias=[0.5, 1.0, 1.5], out_scale=1.0))
    rng = np.random.default_rng(2)
    for n in (16, 200, 4099):
        x = rng.standard_normal((n, 25)).astype(np.float32)
        y = net.forward(x)
        want, _ = om.forward(net.params(), x, emulate=True)
        print(f"[forward] n={n} rel {rel(y, want):.2e}", flush=True)
        assert rel(y, want) < 1e-3


def test_gradient_matches_oracle(tracer):
    _g, tr = tracer
    net = WarmStartMLP(tr, params=init_params(3, out_bias=[0.2, 0.1, 0.0], out_scale=1.0))
    rng = np.random.default_rng(4)
    x = rng.standard_normal((256, 25)).astype(np.float32)
    t = rng.standard_normal((256, 3)).astype(np.float32)
    loss, g = net.grad(x, t)
    wl, wg = om.loss_and_grad(net.params(), x, t, emulate=True)
    assert loss == pytest.approx(wl, rel=1e-4)
    for (gw, gb), (ww, wb) in zip(om.unpack(g), om.unpack(wg)):
        print(f"[grad] W rel {rel(gw, ww):.2e}  b rel {rel(gb, wb):.2e}", flush=True)
        assert rel(gw, ww) < 1e-3 and rel(gb, wb) < 1e-3
    # unrounded float64 gradients are close too: the BF16 operands cost little
    _l64, g64 = om.loss_and_grad(net.params(), x, t, emulate=False)
    assert rel(g, g64) < 3e-2


def test_adam_matches_oracle(tracer):
    _g, tr = tracer
    p0 = init_params(5)
    net = WarmStartMLP(tr, params=p0)
    rng = np.random.default_rng(6)
    w, m, v = p0.astype(np.float64), np.zeros(om.N_PARAMS), np.zeros(om.N_PARAMS)
    for t in (1, 2, 3):
        g = (rng.standard_normal(om.N_PARAMS) * 1e-2).astype(np.float32)
        net.adam(g, lr=1e-2)
        w, m, v = om.adam(w, m, v, g.astype(np.float64), t, lr=1e-2)
        np.testing.assert_allclose(net.params(), w, rtol=1e-5, atol=1e-7)


def test_features_match_oracle(tracer):
    g, tr = tracer
    st = HipBounceStage(tr, g)
    st.load_patches()
    net = WarmStartMLP(tr)
    e0 = np.random.default_rng(7).uniform(0, 50, (g.npatch, 3)).astype(np.float32)
    x = net.features(e0)
    o = g.origin.astype(np.float32)
    want = om.features(o, g.normal, e0, o.min(0), o.max(0), om.fourier_matrix(), 1024.0)
    np.testing.assert_allclose(x, want, atol=2e-5, rtol=1e-5)


def small_map(seed, n=3):
    return rooms.room_grid(n, n, size=(1344.0, 1344.0, 384.0), patch=64.0, minchop=64.0, tile=64.0,
                           nlights=2 * n * n, seed=seed)


def test_training_is_deterministic_and_learns():
    g = small_map(101, 2)
    offer = HipTracerFactory().try_create(g.casters)
    with offer.tracer as tr:
        out = []
        # per-layer kernels in a graph, the same launched one by one, then the
        # fused two-launch step (graph and eager)
        for graph, fused in ((True, False), (False, False), (True, True), (False, True)):
            net = WarmStartMLP(tr, params=init_params(0))
            bounce_map(tr, g)
            net.dataset_clear()
            n = net.dataset_add()
            losses = net.train(epochs=6, batch=512, lr=2e-3, seed=3, use_graph=graph, fused=fused)
            out.append((net.params(), losses))
        print(f"[train] {n} rows, losses {out[0][1]}, fused {out[2][1]}", flush=True)
    assert n == int((g.leaf & ~g.sky.astype(bool)).sum())
    for params, losses in out[1:]:
        assert np.array_equal(params.view(np.uint32), out[0][0].view(np.uint32))
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[2][1], out[3][1])
    # the fused step sums the loss (reported only) in another order
    assert np.allclose(out[2][1], out[0][1], rtol=1e-6, atol=0)
    assert out[0][1][-1] < 0.5 * out[0][1][0]


def test_warm_start_gain_on_a_held_out_map():
    maps = [small_map(s) for s in range(110, 118)]
    held_out = small_map(210)
    offer = HipTracerFactory().try_create(maps[0].casters)
    with offer.tracer as tr:
        net = WarmStartMLP(tr)
        net.dataset_clear()
        prior = Prior()
        for g in maps:
            _st, e0, r = bounce_map(tr, g)
            rows = net.dataset_add()
            prior.add(g, e0, r.total_light)
        # new weights and the prior's gain and scale; the data set stays
        net = WarmStartMLP(tr, params=init_params(0), gain=prior.gain, scale=prior.scale)
        losses = net.train(epochs=40, batch=1024, lr=2e-3, seed=1)
        st, e0, zero = bounce_map(tr, held_out, tol=1e-4)
        pred = net.predict(e0)
        warm = st.bounce(e0, tol=1e-4, max_iter=200, from_prediction=True)
        zs = st.bounce(e0, start=np.zeros((held_out.npatch, 3), np.float32), tol=1e-4, max_iter=200)
        leaf = held_out.leaf & ~held_out.sky.astype(bool)
        direct_only = st.bounce(e0, start=(prior.gain * e0 * leaf[:, None]).astype(np.float32), tol=1e-4,
                                max_iter=200)
        # the prediction: prior plus scaled network output, rescaled per
        # cluster to the prior's light
        y = net.forward(net.features(e0)[leaf])
        ev = np.maximum(e0[leaf], 0)
        q = np.maximum(np.float32(prior.gain) * ev + np.float32(prior.scale) * y, 0).astype(np.float64)
        want = np.zeros_like(q)
        cl = held_out.cluster[leaf]
        for c in np.unique(cl):
            m = cl == c
            want[m] = q[m] * (np.float32(prior.gain) * ev[m]).astype(np.float64).sum(0) / q[m].sum(0)
    print(f"[gain] {rows} rows, loss {losses[0]:.4f} -> {losses[-1]:.4f}; iterations zero {zero.iterations}, "
          f"direct-only {direct_only.iterations}, network {warm.iterations}", flush=True)
    np.testing.assert_allclose(pred[leaf], want, rtol=1e-5, atol=1e-12)
    assert zs.iterations == zero.iterations
    assert warm.iterations < zero.iterations
    assert warm.iterations <= direct_only.iterations
    leaf = held_out.leaf & ~held_out.sky.astype(bool)
    a, b = warm.total_light[leaf].astype(np.float64), zero.total_light[leaf].astype(np.float64)
    assert np.abs(a - b).sum() / np.abs(b).sum() < 1e-3
    assert np.all(np.isfinite(pred)) and (pred >= 0).all()
