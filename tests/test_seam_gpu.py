"""The seam's threading, deadline and cancellation contract on the GPU.
Needs an MI355X.

* Concurrent callers on one tracer (SourceSharp.Tests/MapTools/Tracing/
  VulkanRayTracerFacts.cs:129-220): four threads, each tracing its own rays
  with TraceVisibility and TraceClosest several times, get exactly the
  answers a lone caller gets.
* A deadline (VulkanRayTracer.cs:39-49 bounds each fence wait): with a
  deadline of 0 ms a normal, bounded launch makes the host give up with
  TimeoutError while the kernel runs to its end; the caller's output is left
  alone, and the context answers correctly afterwards.
* Cancellation (the seam's CancellationToken): a cancel before a call, and
  one from another thread while a slow call waits, raise
  OperationCanceledError; after clear_cancel the context answers correctly.
Every kernel here is an ordinary one that ends by itself."""
from __future__ import annotations

import ctypes
import threading
import time

import numpy as np
import pytest

import oracle
from hl2sdk_tools_sharp_amd import HIT_DTYPE, HipTracerFactory, OperationCanceledError, _lib, scenes
from hl2sdk_tools_sharp_amd.light import HipFaceLighter, make_lights, sky_directions, EMIT_SKYAMBIENT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    w = scenes.visibility_workload(nboxes=600, npoints=300_000, nlights=64, seed=9, volume=6000.0)
    offer = HipTracerFactory().try_create(w.casters)
    assert offer.tracer is not None, offer.decline_reason
    yield w, offer.tracer
    offer.tracer.dispose()


def _batches(w, k, n=40_000):
    rng = np.random.default_rng(100 + k)
    pts = w.points[rng.choice(len(w.points), n // 64, replace=False)]
    return scenes.point_light_rays(pts, w.lights)


def test_concurrent_callers_get_the_lone_callers_answers(scene):
    w, tr = scene
    batches = [_batches(w, k) for k in range(4)]
    want = []
    for rays in batches:
        bits = np.zeros((len(rays) + 63) // 64, np.uint64)
        tr.trace_visibility(rays, bits)
        hits = np.zeros(len(rays), HIT_DTYPE)
        tr.trace_closest(rays, hits)
        want.append((bits, hits))
    errors, got = [], [None] * 4

    def worker(k):
        try:
            out = []
            for _ in range(6):
                bits = np.zeros_like(want[k][0])
                tr.trace_visibility(batches[k], bits)
                hits = np.zeros(len(batches[k]), HIT_DTYPE)
                tr.trace_closest(batches[k], hits)
                out.append((bits, hits))
            got[k] = out
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    for k in range(4):
        for bits, hits in got[k]:
            assert np.array_equal(bits, want[k][0])
            assert np.array_equal(hits["surface"], want[k][1]["surface"])
            assert np.array_equal(hits["fraction"].view(np.uint32), want[k][1]["fraction"].view(np.uint32))


def test_a_zero_deadline_gives_up_and_leaves_the_output_alone(scene):
    w, tr = scene
    words = (w.rays + 63) // 64
    want = tr.trace_point_lights(w.points, w.lights)
    sentinel = np.full(words, 0x5A5A5A5A5A5A5A5A, np.uint64)
    bits = sentinel.copy()
    tr.set_timeout(0)
    try:
        with pytest.raises(TimeoutError):
            st = _lib.lib().vh_trace_point_lights(
                tr.handle, w.points.ctypes.data_as(ctypes.c_void_p), len(w.points),
                w.lights.ctypes.data_as(ctypes.c_void_p), len(w.lights), 0.0,
                bits.ctypes.data_as(ctypes.c_void_p), words)
            tr._raise(st)
    finally:
        tr.set_timeout(120_000)
    assert np.array_equal(bits, sentinel)  # nothing reached the caller
    assert _lib.lib().vh_sync(tr.handle, None) == 0  # the kernel ran to its end
    assert np.array_equal(tr.trace_point_lights(w.points, w.lights), want)


def test_a_cancel_before_the_call_and_one_while_it_waits(scene):
    w, tr = scene
    rays = _batches(w, 7)
    want = np.zeros((len(rays) + 63) // 64, np.uint64)
    tr.trace_visibility(rays, want)
    tr.cancel()
    try:
        bits = np.zeros_like(want)
        with pytest.raises(OperationCanceledError):
            tr.trace_visibility(rays, bits)
        assert not bits.any()
    finally:
        tr.clear_cancel()
    bits = np.zeros_like(want)
    tr.trace_visibility(rays, bits)
    assert np.array_equal(bits, want)

    # a slow call (one gather launch per face, full sky-ambient fans) cancelled
    # from another thread while it waits on the device
    casters, faces = scenes.batching_workload(nfaces=512, seed=4)
    offer = HipTracerFactory().try_create(casters)
    with offer.tracer as t2:
        fl = HipFaceLighter(t2)
        L = make_lights(1)
        L[0]["type"], L[0]["intensity"] = EMIT_SKYAMBIENT, (1.0, 1.0, 1.0)
        fl.load(faces, L, sky_dirs=sky_directions(), ao_count=32)
        t0 = time.perf_counter()
        fl.run("queue")
        quick = time.perf_counter() - t0
        timer = threading.Timer(0.05, t2.cancel)
        timer.start()
        t0 = time.perf_counter()
        try:
            with pytest.raises(OperationCanceledError):
                fl.run("per_face")
        finally:
            timer.join()
            t2.clear_cancel()
        gave_up = time.perf_counter() - t0
        print(f"[seam] queue run {quick * 1e3:.0f} ms; per-face run cancelled after {gave_up * 1e3:.0f} ms",
              flush=True)
        assert _lib.lib().vh_sync(t2.handle, None) == 0
        ref = fl.read()
        fl.run("queue")
        again = fl.read()
        assert np.array_equal(ref.ldr.view(np.uint32), again.ldr.view(np.uint32))


def test_the_oracle_agrees_with_the_concurrent_answers(scene):
    """The lone caller's answers themselves are the oracle's."""
    w, tr = scene
    rays = _batches(w, 0)
    bits = np.zeros((len(rays) + 63) // 64, np.uint64)
    tr.trace_visibility(rays, bits)
    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    want = oracle.CpuBvh(w.casters.vertices, tris).visibility(rays)
    assert np.array_equal(bits, want)
