"""Launch-shape autotuner for the persistent gather queue.

The queue kernel has two free shape parameters: how many persistent waves it
keeps per compute unit, and its grain -- how many consecutive 64-sample
chunks a wave takes per atomic ticket.  Too few waves leave traversal latency
unhidden; more than the register file holds only adds waves that find the
queue empty.  A larger grain keeps a wave on one face longer (its rays share
more of the tree) and takes fewer atomics; a smaller one balances the
queue's tail.  The tuner times the wave counts at grain 1, then the grains at
the best wave count and its neighbours, with HIP events on the library's
stream; it keeps the fastest pair and caches it in a JSON file keyed by the
device identity and a signature of the workload (sample count, bump
fraction, light types, sun/sky/AO ray counts), so a bake pays for the sweep
once per machine and workload shape.

The cache lives at ``$VRADHIP_TUNE_CACHE`` or ``~/.cache/vradhip/tune.json``.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

CANDIDATE_WAVES_PER_CU = (2, 4, 6, 8, 10, 12, 16, 20, 24, 32)
CANDIDATE_CHUNKS_PER_TICKET = (1, 2, 4, 8)


def cache_path() -> str:
    p = os.environ.get("VRADHIP_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "vradhip", "tune.json")


def workload_key(lighter) -> str:
    """A coarse signature of what the loaded gather traces per sample."""
    f = lighter.faces
    d, keep = lighter._keep
    lights = keep["lights"]
    types = ",".join(str(int(t)) for t in sorted(set(lights["type"].tolist())))
    bumped = float((f.face_normal_count == 4).mean()) if f.nfaces else 0.0
    scale = int(round(math.log2(max(f.nsamples, 1))))
    return (f"gather:s2^{scale}:bump{bumped:.2f}:types[{types}]:sun{d.sun_samples}:"
            f"sky{d.sky_dirs_count}:ao{d.ao_count}")


@dataclass
class TuneResult:
    waves_per_cu: int
    ms: Dict[int, float] = field(default_factory=dict)  # per wave count, grain 1
    cached: bool = False
    chunks_per_ticket: int = 1
    grain_ms: Dict[str, float] = field(default_factory=dict)  # "waves:grain" -> ms


def _load(path: str) -> dict:
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: str, data: dict) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_gather(lighter, candidates: Optional[List[int]] = None, repeats: int = 1,
                use_cache: bool = True, path: Optional[str] = None, log=None,
                grains: Optional[List[int]] = None) -> TuneResult:
    """Picks waves_per_cu and chunks_per_ticket for ``lighter`` (a loaded
    HipFaceLighter) and caches the choice.  Each candidate runs the whole
    loaded gather ``repeats`` times; the median kernel time decides."""
    path = path or cache_path()
    key = f"{lighter.tracer.tracer_identity}|{workload_key(lighter)}"
    data = _load(path) if use_cache else {}
    hit = data.get(key)
    if use_cache and isinstance(hit, dict) and isinstance(hit.get("waves_per_cu"), int):
        return TuneResult(hit["waves_per_cu"], {int(k): v for k, v in hit.get("ms", {}).items()}, True,
                          int(hit.get("chunks_per_ticket", 1)), dict(hit.get("grain_ms", {})))
    cands = list(candidates or CANDIDATE_WAVES_PER_CU)

    def timed(w, g):
        runs = sorted(lighter.timed_run("queue", w, g)[0] for _ in range(max(1, repeats)))
        return runs[len(runs) // 2]

    times: Dict[int, float] = {}
    for w in cands:
        times[w] = timed(w, 1)
        if log:
            log(f"[tune] waves_per_cu={w}: {times[w]:.2f} ms")
    best = min(times, key=lambda w: (times[w], w))
    # the grains at the best wave count and its neighbours in the list
    i = cands.index(best)
    grain_ms: Dict[str, float] = {f"{best}:1": times[best]}
    for w in cands[max(0, i - 1):i + 2]:
        for g in grains or CANDIDATE_CHUNKS_PER_TICKET:
            k = f"{w}:{g}"
            if k not in grain_ms:
                grain_ms[k] = times[w] if g == 1 else timed(w, g)
                if log and g != 1:
                    log(f"[tune] waves_per_cu={w} chunks_per_ticket={g}: {grain_ms[k]:.2f} ms")
    pick = min(grain_ms, key=lambda k: (grain_ms[k], int(k.split(":")[1]), int(k.split(":")[0])))
    bw, bg = (int(x) for x in pick.split(":"))
    if use_cache:
        data = _load(path)
        data[key] = {"waves_per_cu": bw, "chunks_per_ticket": bg, "ms": {str(k): v for k, v in times.items()},
                     "grain_ms": grain_ms}
        _store(path, data)
    return TuneResult(bw, times, False, bg, grain_ms)
