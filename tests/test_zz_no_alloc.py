"""vh_trace_point_lights_noalloc against vh_trace_point_lights_dev through
the C ABI.  Needs an MI355X.

Two calls in a row on one context, the second with fewer points and a light
count that leaves partial words (so the entry point must clear the output
itself): every hit word equal to the allocating entry point's on the same
inputs, and, on the second call, to the oracle's.  A batch over the fixed
maximum is refused before anything runs."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

# torch (the stream and device buffers of the caller) is imported when the
# module is collected: its own HIP runtime must be loaded before a test loads
# libvradhip.so and with it the system's, or torch finds no device.
import torch

import oracle
from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, scenes

pytestmark = pytest.mark.gpu

VH_NOALLOC_MAX_POINTS = 2097152


def _call(fn, h, pts, lts, bits, stream):
    P = ctypes.c_void_p
    return fn(h, P(pts.data_ptr()), pts.shape[0], P(lts.data_ptr()), lts.shape[0], 0.0, P(bits.data_ptr()),
              P(stream.cuda_stream))


def test_noalloc_equals_the_allocating_entry_point_twice():
    dev = torch.device("cuda:0")
    w = scenes.visibility_workload(nboxes=600, npoints=300_000, nlights=64, seed=9, volume=6000.0)
    offer = HipTracerFactory(device=0).try_create(w.casters)
    assert offer.tracer is not None, offer.decline_reason
    L = _lib.lib()
    stream = torch.cuda.Stream(dev)
    with offer.tracer as tr:
        h = tr.handle
        for npts, nl in ((300_000, 64), (70_001, 40)):
            print(f"[noalloc] {npts} points x {nl} lights", flush=True)
            pts = torch.from_numpy(np.ascontiguousarray(w.points[:npts])).to(dev)
            lts = torch.from_numpy(np.ascontiguousarray(w.lights[:nl])).to(dev)
            words = (npts * nl + 63) // 64
            got = torch.full((words,), -1, dtype=torch.int64, device=dev)  # dirty: the call must clear
            want = torch.full((words,), -1, dtype=torch.int64, device=dev)
            torch.cuda.synchronize()
            assert _call(L.vh_trace_point_lights_noalloc, h, pts, lts, got, stream) == 0, _lib.last_error()
            assert _call(L.vh_trace_point_lights_dev, h, pts, lts, want, stream) == 0, _lib.last_error()
            torch.cuda.synchronize()
            g, wnt = got.cpu().numpy(), want.cpu().numpy()
            assert np.array_equal(g, wnt)
            frac = np.unpackbits(g.view(np.uint8)).mean()
            assert 0.05 < frac < 0.95  # both hits and misses
        # the second batch against the oracle as well
        tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
        bvh = oracle.CpuBvh(w.casters.vertices, tris)
        ref = bvh.point_lights(w.points[:70_001], w.lights[:40])
        assert np.array_equal(g.view(np.uint64), ref)
        # over the fixed maximum: refused, nothing launched
        big = VH_NOALLOC_MAX_POINTS + 1
        P = ctypes.c_void_p
        rc = L.vh_trace_point_lights_noalloc(h, P(pts.data_ptr()), big, P(lts.data_ptr()), 40, 0.0,
                                             P(got.data_ptr()), P(stream.cuda_stream))
        assert rc == 2 and b"at most" in _lib.lib().vh_last_error()
