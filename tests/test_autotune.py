"""The gather autotuner's choice and cache, with a stand-in lighter (CPU)."""
from __future__ import annotations

import json

import numpy as np

from hl2sdk_tools_sharp_amd import autotune, light
from hl2sdk_tools_sharp_amd.light import FaceSet


class FakeTracer:
    tracer_identity = "hip-bvh2-exact-1:gfx950:256cu"


class FakeLighter:
    """Kernel time is a parabola in waves per CU with its minimum at 10, less
    0.3 ms at two chunks per ticket and more beyond."""

    def __init__(self):
        self.tracer = FakeTracer()
        n = 100
        normals = np.zeros((1, 4, 3), np.float32)
        normals[0, 0] = (0, 0, 1)
        self.faces = FaceSet(np.array([0, n], np.int32), np.array([1], np.int32), normals,
                             np.zeros((n, 3), np.float32), np.zeros(n, np.int32))
        L = light.make_lights(1)
        L[0]["type"] = light.EMIT_SKYAMBIENT
        self._keep = light.pack_desc(self.faces, L, None, 0.0, 64, light.sky_directions(), 32, 40.0)
        self.calls = []

    def timed_run(self, schedule, waves_per_cu, chunks_per_ticket=1):
        self.calls.append((schedule, waves_per_cu, chunks_per_ticket))
        g = {1: 0.0, 2: -0.3, 4: 0.2, 8: 0.5}[chunks_per_ticket]
        return 5.0 + (waves_per_cu - 10) ** 2 * 0.1 + g, 1, 1000


def test_picks_the_fastest_and_caches(tmp_path):
    path = str(tmp_path / "tune.json")
    fl = FakeLighter()
    r = autotune.tune_gather(fl, path=path)
    assert r.waves_per_cu == 10 and r.chunks_per_ticket == 2 and not r.cached
    assert {w for _, w, g in fl.calls if g == 1} == set(autotune.CANDIDATE_WAVES_PER_CU)
    # grains are tried at the best wave count and its two neighbours only
    assert {w for _, w, g in fl.calls if g != 1} == {8, 10, 12}
    assert {g for *_, g in fl.calls} == set(autotune.CANDIDATE_CHUNKS_PER_TICKET)
    assert all(s == "queue" for s, *_ in fl.calls)
    data = json.load(open(path))
    (key,) = data
    assert key.startswith("hip-bvh2-exact-1:gfx950:256cu|gather:") and "sun64" in key and "sky162" in key
    fl2 = FakeLighter()
    r2 = autotune.tune_gather(fl2, path=path)
    assert r2.cached and r2.waves_per_cu == 10 and r2.chunks_per_ticket == 2 and fl2.calls == []


def test_ties_go_to_fewer_waves_and_no_cache_mode(tmp_path):
    fl = FakeLighter()
    fl.timed_run = lambda s, w, g=1: (1.0, 1, 1)
    r = autotune.tune_gather(fl, candidates=[8, 4, 12], use_cache=False, path=str(tmp_path / "x.json"))
    assert r.waves_per_cu == 4 and r.chunks_per_ticket == 1
    assert not (tmp_path / "x.json").exists()


def test_a_corrupt_cache_is_ignored(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    r = autotune.tune_gather(FakeLighter(), path=str(path))
    assert r.waves_per_cu == 10
    assert json.load(open(path))


def test_a_cache_entry_without_a_grain_reads_as_one_chunk_per_ticket(tmp_path):
    path = tmp_path / "tune.json"
    fl = FakeLighter()
    key = f"{fl.tracer.tracer_identity}|{autotune.workload_key(fl)}"
    path.write_text(json.dumps({key: {"waves_per_cu": 6, "ms": {"6": 1.0}}}))
    r = autotune.tune_gather(fl, path=str(path))
    assert r.cached and (r.waves_per_cu, r.chunks_per_ticket) == (6, 1) and fl.calls == []
