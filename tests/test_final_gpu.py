"""The GPU final light (luxel radial filter, patch radial, minlight, median,
lump) against the oracle.  Needs an MI355X.

The bar: every lump byte of every face the oracle runs equal to the
oracle's, which restates FinalLightFace.Run in the reference's float order
(oracle/final.py); the sample light both read is the GPU gather's own
result, read back."""
from __future__ import annotations

import numpy as np
import pytest

from final_cases import face_bytes, oracle_inputs, random_case
from oracle import final as fo
from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, rooms
from hl2sdk_tools_sharp_amd.tracer import NotSupportedError
from hl2sdk_tools_sharp_amd.final import HipFinalLighter
from hl2sdk_tools_sharp_amd.light import HipFaceLighter

pytestmark = pytest.mark.gpu


def open_all(casters):
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is not None, offer.decline_reason
    fl = HipFaceLighter(offer.tracer)
    return offer.tracer, fl, HipFinalLighter(fl)


def check_lump(lump, size, want):
    assert lump.shape[0] == size
    for f, (ofs, b, nst) in want.items():
        got = lump[ofs - 4 * nst:ofs - 4 * nst + len(b)].tobytes()
        if got != b:
            g = np.frombuffer(got, np.uint8).reshape(-1, 4)
            e = np.frombuffer(b, np.uint8).reshape(-1, 4)
            bad = np.nonzero(np.any(g != e, 1))[0]
            raise AssertionError(f"face {f}: {len(bad)} of {len(e)} words differ, first {bad[:5]}: "
                                 f"{g[bad[:3]].tolist()} vs {e[bad[:3]].tolist()}")


@pytest.mark.parametrize("seed,hdr,red,grid", [(1, False, False, False), (2, True, True, False),
                                               (3, False, True, True)])
def test_random_faces_match_oracle(seed, hdr, red, grid):
    """grid: the luxels are the frames' grid points, made on the device."""
    casters, L, fset, ff, fp = random_case(seed)
    if grid:
        ff.luxel_pos = None
    print(f"[final] random seed={seed} hdr={hdr} red={red}: {fset.nfaces} faces, {fset.nsamples} samples",
          flush=True)
    tr, fl, fin = open_all(casters)
    with tr:
        fl.load(fset, L, sun_samples=1)
        g = fl.light_faces()
        fin.load(ff, fp)
        offs, size = fin.layout()
        lump = fin.run(hdr=hdr, red_errors=red)
        rng = np.random.default_rng(seed)
        total = rng.uniform(0, 40, (fp.npatch, 4, 3)).astype(np.float32)
        lump_b = fin.run(hdr=hdr, bounce="host", patch_light=total, red_errors=red)
    light = g.hdr if hdr else g.ldr
    assert np.any(light[0, :, :3] > 0)
    faces, lux, sm, nb, _ = oracle_inputs(ff, fset, light)
    want_offs, want_size = fo.layout(faces)
    assert list(offs) == want_offs and size == want_size
    res = fo.final_light(faces, lux, sm, nb, None, red)
    check_lump(lump, size, face_bytes(faces, res, offs))
    faces, lux, sm, nb, pt = oracle_inputs(ff, fset, light, fp, total)
    res = fo.final_light(faces, lux, sm, nb, pt, red)
    check_lump(lump_b, size, face_bytes(faces, res, offs))
    assert lump_b.tobytes() != lump.tobytes()


@pytest.fixture(scope="module")
def rooms_case():
    g = rooms.room_grid(2, 2, tile=64.0, seed=2)
    fset = rooms.room_faces(g, 16.0, bumped_every=5)
    lights, pvs = rooms.room_lights(g, sun=False)
    ff, fp = rooms.room_final_inputs(g, fset)
    print(f"[final] rooms: {fset.nfaces} faces, {fset.nsamples} samples, {g.npatch} patches", flush=True)
    return g, fset, lights, pvs, ff, fp


def test_room_faces_match_oracle(rooms_case):
    g, fset, lights, pvs, ff, fp = rooms_case
    tr, fl, fin = open_all(g.casters)
    with tr:
        fl.load(fset, lights, light_pvs=pvs, sun_samples=1)
        gr = fl.light_faces()
        fin.load(ff, fp)
        offs, size = fin.layout()
        lump = fin.run()
        total = np.zeros((g.npatch, 4, 3), np.float32)
        total[:, :, :] = np.random.default_rng(5).uniform(0, 20, (g.npatch, 4, 3))
        lump_b = fin.run(bounce="host", patch_light=total)
    # a spread of faces: bumped ones, wall, floor and jamb tiles
    only = set(range(0, fset.nfaces, 37)) | {f for f in range(fset.nfaces) if fset.face_normal_count[f] == 4
                                            and f % 40 == 0}
    faces, lux, sm, nb, pt = oracle_inputs(ff, fset, gr.ldr, fp, total)
    res = fo.final_light(faces, lux, sm, nb, None, only=only)
    check_lump(lump, size, face_bytes(faces, res, offs))
    res = fo.final_light(faces, lux, sm, nb, pt, only=only)
    check_lump(lump_b, size, face_bytes(faces, res, offs))


def test_device_bounce_equals_the_same_light_from_the_host(rooms_case):
    """The bounce's device-resident light gives the bytes the same light
    handed in from the host gives; a bumped face owning patches needs the
    bounce's bump totals, and is refused without them."""
    from hl2sdk_tools_sharp_amd.bounce import HipBounceStage

    g, fset, lights, pvs, ff, fp = rooms_case
    flat = rooms.room_faces(g, 16.0)
    tr, fl, fin = open_all(g.casters)
    with tr:
        fl.load(flat, lights, light_pvs=pvs, sun_samples=1)
        fl.run()
        st = HipBounceStage(tr, g)
        st.load_patches()
        st.build_transfers()
        rng = np.random.default_rng(6)
        e = np.where((g.leaf & ~g.sky.astype(bool))[:, None], rng.uniform(0, 50, (g.npatch, 3)), 0)
        b = st.bounce(g.pull_up(e.astype(np.float32)), tol=1e-3, max_iter=30)
        fin.load(ff, fp)
        dev = fin.run(bounce="device")
        total = np.zeros((g.npatch, 4, 3), np.float32)
        total[:, 0] = b.total_light
        host = fin.run(bounce="host", patch_light=total)
        assert np.any(b.total_light > 0)
        assert dev.tobytes() == host.tobytes()
        # bumped faces with patches: the flat light cannot stand in for
        # theirs, so a bounce loaded without bump normals is refused ...
        fl.load(fset, lights, light_pvs=pvs, sun_samples=1)
        fl.run()
        fin.load(ff, fp)
        with pytest.raises(NotSupportedError):
            fin.run(bounce="device")
        # ... and one loaded with them hands over all four directions
        bumped, bn = rooms.patch_bumps(g, fset)
        st.load_patches(bumped, bn)
        st.build_transfers()
        b = st.bounce(g.pull_up(e.astype(np.float32)), tol=1e-3, max_iter=30)
        assert b.bump_light is not None and np.any(b.bump_light > 0)
        fin.load(ff, fp)
        dev = fin.run(bounce="device")
        total[:, 0] = b.total_light
        total[:, 1:] = b.bump_light
        host = fin.run(bounce="host", patch_light=total)
        assert dev.tobytes() == host.tobytes()


def test_stage_errors():
    casters, L, fset, ff, fp = random_case(4)
    tr, fl, fin = open_all(casters)
    with tr:
        with pytest.raises(_lib.VradHipError):
            fin.load(ff, fp)  # no gather loaded
        fl.load(fset, L, sun_samples=1)
        fin.load(ff, fp)
        with pytest.raises(_lib.VradHipError):
            fin.run()  # no gather results yet
        fl.run()
        with pytest.raises(_lib.VradHipError):
            fin.run(bounce="device")  # no bounce
        lump = fin.run()
        assert lump.shape[0] == fin.layout()[1]
