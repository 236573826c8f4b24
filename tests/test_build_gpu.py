"""The scene build on the GPU against the host build.  Needs an MI355X.

Both builders run the same binned SAH (64 bins, the same cost and leaf rule,
the same split predicate), so on scenes without coincident centroids the
device tree is the host tree: the same node count and depth, every node's
child boxes, references and near-child axis equal, and each leaf holding the
same triangles (in an order of its own: the host partitions in place, the
device stably), each converted to the same intersection-format bits.  Every
trace answers the same through either tree."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hl2sdk_tools_sharp_amd import HIT_DTYPE, _lib, rooms, scenes
from hl2sdk_tools_sharp_amd.tracer import HipRayTracer

pytestmark = pytest.mark.gpu


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def open_scene(casters, builder):
    L = _lib.lib()
    opts = _lib.VhOpenOpts(0, 0, 0, 0)
    h = ctypes.c_void_p()
    assert L.vh_open(ctypes.byref(opts), ctypes.byref(h)) == 0, _lib.last_error()
    st = L.vh_scene_load_with(h, _p(casters.vertices), _p(casters.ids), _p(casters.flags), casters.count, builder)
    if st != 0:
        msg = _lib.last_error()
        L.vh_close(h)
        raise ValueError(msg)
    return HipRayTracer(h.value, casters)


def read_scene(tr):
    L = _lib.lib()
    t, n, d = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert L.vh_scene_info(tr.handle, ctypes.byref(t), ctypes.byref(n), ctypes.byref(d)) == 0
    nodes = np.zeros((n.value, 16), np.float32)
    tris = np.zeros((t.value, 12), np.float32)
    assert L.vh_scene_read(tr.handle, _p(nodes), nodes.size, _p(tris), tris.size) == 0, _lib.last_error()
    return nodes, tris, d.value


def leaves(tris):
    """Leaf ranges from the last-in-leaf bit of each triangle's meta word."""
    meta = tris[:, 11].view(np.uint32)
    ends = np.flatnonzero(meta & 16) + 1
    starts = np.concatenate([[0], ends[:-1]])
    return list(zip(starts, ends))


def same_tree(host, dev):
    hn, ht, hd = host
    dn, dt, dd = dev
    assert hn.shape == dn.shape and ht.shape == dt.shape and hd == dd
    # boxes as values (the sign of a zero bound may differ), references and axes as bits
    assert np.array_equal(hn[:, :12], dn[:, :12])
    assert np.array_equal(hn[:, 12:15].view(np.int32), dn[:, 12:15].view(np.int32))
    lh, ld = leaves(ht), leaves(dt)
    assert lh == ld
    rows_h = ht.view(np.uint32).copy()
    rows_d = dt.view(np.uint32).copy()
    rows_h[:, 11] &= ~np.uint32(16)
    rows_d[:, 11] &= ~np.uint32(16)
    for a, b in lh:
        h = rows_h[a:b]
        d = rows_d[a:b]
        h = h[np.lexsort(h.T[::-1])]
        d = d[np.lexsort(d.T[::-1])]
        assert np.array_equal(h, d), (a, b)


@pytest.mark.parametrize("name", ["rooms_6x6", "visibility_24k"])
def test_device_tree_is_the_host_tree(name):
    if name == "rooms_6x6":
        casters = rooms.bounce_grid(6, 6).casters
    else:
        casters = scenes.visibility_workload(nboxes=2000, npoints=1000, nlights=4, seed=2).casters
    print(f"[build] {name}: {casters.count} triangles", flush=True)
    with open_scene(casters, _lib.VH_BUILD_HOST) as th, open_scene(casters, _lib.VH_BUILD_DEVICE) as td:
        host, dev = read_scene(th), read_scene(td)
        print(f"[build] {name}: {host[0].shape[0]} nodes, depth {host[2]}", flush=True)
        same_tree(host, dev)
        # every trace answers the same through either tree: segments between
        # random points of the scene's box
        v = casters.vertices.reshape(-1, 3)
        lo, hi = v.min(0), v.max(0)
        rng = np.random.default_rng(3)
        a = rng.uniform(lo, hi, (20_000, 3)).astype(np.float32)
        b = rng.uniform(lo, hi, (20_000, 3)).astype(np.float32)
        rays = np.zeros((len(a), 7), np.float32)
        rays[:, :3] = a
        rays[:, 3:6] = b - a
        rays[:, 6] = 1.0
        bits = [np.zeros((len(rays) + 63) // 64, np.uint64) for _ in range(2)]
        hits = [np.zeros(len(rays), HIT_DTYPE) for _ in range(2)]
        for k, tr in enumerate((th, td)):
            tr.trace_visibility(rays, bits[k])
            tr.trace_closest(rays, hits[k])
        assert np.array_equal(bits[0], bits[1])
        assert np.array_equal(hits[0]["surface"], hits[1]["surface"])
        assert np.array_equal(hits[0]["fraction"].view(np.uint32), hits[1]["fraction"].view(np.uint32))
        nblocked = int(np.unpackbits(bits[0].view(np.uint8)).sum())
        print(f"[build] {name}: {nblocked} of {len(rays)} segments blocked", flush=True)
        assert 0 < nblocked < len(rays)


def test_the_bake_scene_builds_on_the_device_by_default():
    casters = rooms.bounce_grid(12, 12).casters
    assert casters.count >= _lib.VH_BUILD_DEVICE_MIN
    with open_scene(casters, _lib.VH_BUILD_AUTO) as ta, open_scene(casters, _lib.VH_BUILD_HOST) as th:
        same_tree(read_scene(th), read_scene(ta))


def test_device_build_errors():
    casters = rooms.bounce_grid(2, 2).casters
    v = casters.vertices.copy()
    v[37, 4] = np.nan
    L = _lib.lib()
    opts = _lib.VhOpenOpts(0, 0, 0, 0)
    h = ctypes.c_void_p()
    assert L.vh_open(ctypes.byref(opts), ctypes.byref(h)) == 0
    einval = 2
    try:
        st = L.vh_scene_load_with(h, _p(v), None, None, casters.count, _lib.VH_BUILD_DEVICE)
        assert st == einval and "triangle 37 has a non-finite vertex" in _lib.last_error()
        st = L.vh_scene_load_with(h, _p(casters.vertices), None, None, 8, _lib.VH_BUILD_DEVICE)
        assert st == einval and "at least 16" in _lib.last_error()
        st = L.vh_scene_load_with(h, _p(casters.vertices), None, None, casters.count, 7)
        assert st == einval
        # the context still loads and traces after the refusals
        assert L.vh_scene_load_with(h, _p(casters.vertices), None, None, casters.count, _lib.VH_BUILD_DEVICE) == 0
    finally:
        L.vh_close(h)
