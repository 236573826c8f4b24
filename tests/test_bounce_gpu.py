"""The GPU transfer build and bounce against the oracle.  Needs an MI355X.

Transfers: the same rows, sources and order as the oracle's VisMatrix
restatement; weights bit-equal except where the arcsine's double rounding
moves a last bit (relative 1e-6 at most).  Bounce: FP32 within 1e-5 of the
oracle's FP32 reference form on the same transfers, and both FP32 and FP16
weights within 1e-3 relative (L1 over all patches and channels) of the FP64
oracle at residual 1e-4."""
from __future__ import annotations

import numpy as np
import pytest

import oracle.bounce as ob
from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, rooms
from hl2sdk_tools_sharp_amd.bounce import HipBounceStage

pytestmark = pytest.mark.gpu


def open_stage(g):
    offer = HipTracerFactory().try_create(g.casters)
    assert offer.tracer is not None, offer.decline_reason
    return offer.tracer, HipBounceStage(offer.tracer, g)


def emission(g, seed=3, lit=40):
    rng = np.random.default_rng(seed)
    e0 = np.zeros((g.npatch, 3), np.float32)
    lv = np.nonzero(g.leaf & ~g.sky.astype(bool))[0]
    pick = rng.choice(lv, min(lit, len(lv)), replace=False)
    e0[pick] = rng.uniform(100, 1000, (len(pick), 3))
    return g.pull_up(e0)


def rel_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / np.abs(b).sum())


@pytest.fixture(scope="module")
def case():
    g = rooms.room_grid(2, 2, size=(384, 320, 192), patch=128, minchop=64, seed=5)
    rp, src, w, ncand = ob.build_transfers(g)
    return g, rp, src, w, ncand


def test_transfer_build_matches_oracle(case):
    g, rp, src, w, ncand = case
    print(f"[transfers] {g.npatch} patches, oracle {len(src)} transfers of {ncand} candidates", flush=True)
    tr, st = open_stage(g)
    with tr:
        n = st.build_transfers()
        grp, gsrc, gw = st.read_transfers()
    assert st.candidate_count == ncand
    assert n == len(src)
    assert np.array_equal(grp, rp)
    assert np.array_equal(gsrc, src)
    same = np.mean(gw.view(np.uint32) == w.view(np.uint32))
    assert same > 0.999, same
    np.testing.assert_allclose(gw, w, rtol=1e-6, atol=0)


def test_transfer_build_config_shape():
    """A 2x2 slice of the bounce config (64-unit tiles, one patch per face)."""
    g = rooms.bounce_grid(2, 2)
    print(f"[transfers 2x2 bounce rooms] {g.npatch} patches", flush=True)
    rp, src, w, ncand = ob.build_transfers(g)
    tr, st = open_stage(g)
    with tr:
        st.build_transfers()
        grp, gsrc, gw = st.read_transfers()
    assert np.array_equal(grp, rp) and np.array_equal(gsrc, src)
    np.testing.assert_allclose(gw, w, rtol=1e-6, atol=0)


def test_trf1_export_is_the_transfer_set(case):
    """The device-written records framed as the reference's transfer cache
    (TransferSetCodec) read back to the built rows, entry for entry, raw and
    packed."""
    from hl2sdk_tools_sharp_amd import trf1

    g = case[0]
    tr, st = open_stage(g)
    with tr:
        st.build_transfers()
        rp, src, w = st.read_transfers()
        idx, chunks = st.export_trf1(max_chunk_bytes=1 << 16)
    print(f"[trf1] {len(src)} transfers in {len(chunks)} chunks", flush=True)
    back = trf1.read(idx, chunks, g.npatch)
    assert back is not None
    counts, bsrc, bw, mx = back
    assert np.array_equal(counts, np.diff(rp)) and mx == int(np.diff(rp).max())
    assert np.array_equal(bsrc, src) and np.array_equal(bw.view(np.uint32), w.view(np.uint32))
    assert all(len(c) == 1 << 16 for c in chunks[:-1])
    unpacked = [trf1.unpack(trf1.pack(c)) for c in chunks]
    assert unpacked == chunks


@pytest.mark.parametrize("graph", [True, False])
def test_bounce_fp32_matches_oracle(case, graph):
    g, rp, src, w, _ = case
    e0 = emission(g)
    tr, st = open_stage(g)
    with tr:
        st.load_transfers(rp, src, w)
        r = st.bounce(e0, weights="f32", tol=1e-6, max_iter=300, use_graph=graph)
    t32, res32, en32, n32 = ob.bounce(g, rp, src, w, e0, tol=1e-6, max_iter=300)
    assert abs(r.iterations - n32) <= 1
    assert rel_l1(r.total_light, t32) < 1e-5
    np.testing.assert_allclose(r.added[: min(n32, r.iterations)], en32[: min(n32, r.iterations)], rtol=1e-4)


def test_bumped_receivers_match_oracle(case):
    """GatherBumpedScalar (Rad/Bounce/Radiosity.cs:406-421): every third face
    bumped; the flat totals and the three bump directions' totals within 1e-5
    (relative L1) of the oracle's FP32 form, parents pulled up by area."""
    g, rp, src, w, _ = case
    fset = rooms.room_faces(g, 16.0, bumped_every=3)
    bumped, bn = rooms.patch_bumps(g, fset)
    assert 0 < bumped.sum() < g.npatch
    e0 = emission(g, seed=11, lit=60)
    tr, st = open_stage(g)
    with tr:
        st.load_patches(bumped, bn)
        st.load_transfers(rp, src, w)
        r = st.bounce(e0, weights="f32", tol=1e-6, max_iter=300)
        st.load_patches()
        st.load_transfers(rp, src, w)
        flat = st.bounce(e0, weights="f32", tol=1e-6, max_iter=300)
    tb = np.zeros((g.npatch, 3, 3))
    t32, _, _, n32 = ob.bounce(g, rp, src, w, e0, tol=1e-6, max_iter=300, bumped=bumped, bump_normals=bn,
                               bump_out=tb)
    print(f"[bumped] {int(bumped.sum())} of {g.npatch} patches bumped, {r.iterations} iterations", flush=True)
    assert abs(r.iterations - n32) <= 1
    assert r.bump_light is not None and flat.bump_light is None
    assert rel_l1(r.total_light, t32) < 1e-5
    assert rel_l1(r.bump_light, tb) < 1e-5
    assert np.all(r.bump_light[bumped == 0] == 0)
    # the flat slot of a bumped receiver divides the cosine out and back in:
    # the same light as the flat gather, to rounding
    assert rel_l1(r.total_light, flat.total_light) < 1e-5
    assert np.any(r.bump_light[bumped == 1] > 0)


@pytest.mark.parametrize("weights", ["f32", "f16"])
def test_bounce_within_1e3_of_fp64(case, weights):
    g, rp, src, w, _ = case
    e0 = emission(g, seed=7, lit=200)
    tr, st = open_stage(g)
    with tr:
        st.build_transfers()
        r = st.bounce(e0, weights=weights, tol=1e-4, max_iter=200)
    t64, _, _, _ = ob.bounce(g, rp, src, w, e0, tol=1e-12, max_iter=1000, use_double=True)
    err = rel_l1(r.total_light, t64)
    print(f"[{weights}] iterations {r.iterations}, relative error vs FP64 {err:.2e}", flush=True)
    assert r.residuals[-1] <= 1e-4
    assert err <= 1e-3


def test_reference_stop_rule(case):
    g, rp, src, w, _ = case
    e0 = emission(g, seed=9, lit=30) * np.float32(0.02)
    tr, st = open_stage(g)
    with tr:
        st.load_transfers(rp, src, w)
        r = st.bounce(e0, tol=0.0, max_iter=100)
    _, _, en, n = ob.bounce(g, rp, src, w, e0, tol=0.0, max_iter=100)
    assert r.iterations == n
    assert np.all(r.added[-1] < 1.0)


def test_from_start_and_determinism(case):
    g, rp, src, w, _ = case
    e0 = emission(g, seed=11)
    tr, st = open_stage(g)
    with tr:
        st.load_transfers(rp, src, w)
        a = st.bounce(e0, tol=1e-5, max_iter=300)
        b = st.bounce(e0, tol=1e-5, max_iter=300)
        z = st.bounce(e0, start=np.zeros_like(e0), tol=1e-5, max_iter=300)
        warm = st.bounce(e0, start=a.total_light, tol=1e-4, max_iter=300)
    assert np.array_equal(a.total_light, b.total_light)  # run to run, bit for bit
    assert abs(z.iterations - a.iterations) <= 1
    assert rel_l1(z.total_light, a.total_light.astype(np.float64)) < 1e-4
    assert warm.iterations <= 2


def test_bounce_errors():
    g = rooms.room_grid(1, 1, size=(256, 256, 128), patch=128, minchop=64, seed=1)
    tr, st = open_stage(g)
    with tr:
        st.load_patches()
        with pytest.raises(_lib.VradHipError) as ei:
            st.bounce(np.zeros((g.npatch, 3), np.float32))
        assert ei.value.status == _lib.VH_ESTATE
        bad = np.zeros(g.npatch + 1, np.int64)
        bad[1:] = 1  # a transfer on patch 0, which is a parent
        with pytest.raises(ValueError):
            st.load_transfers(bad, np.zeros(1, np.int32), np.ones(1, np.float32))
