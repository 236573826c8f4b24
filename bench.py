"""Benchmark: shadow-ray visibility on one MI355X.

Workload (the visibility config): 2,000 seeded axis-aligned boxes (edges
log-uniform 16-512, centres uniform in an 8192^3 volume) = 24,000 triangles;
2M sample points on box faces, each tested against 64 point lights = 128M
FP32 shadow rays per step, generated on the GPU in vrad's correct-mode
parameterisation.  Inputs are resident in HBM before the timed region.

One JSON line on stdout.  `value` = shadow rays/s over the timed steps.
`roofline` is for the step's trace kernel: k_point_lights_lb (light buffers,
the default at this size: counted cell lookups x 15 + list entries x 1 +
triangle tests x 27 FLOP) or k_point_lights (the BVH walk: node visits x 44 +
triangle tests x 27 FLOP), divided by the kernel's mean duration from HIP
events on its stream.  `bvh_walk` times the walk on the same batch and
compares its bits with the light buffer's.
`cpu_baseline` is the oracle's CPU restatement (BVH + the reference's exact
triangle test, any hit) timed on a bounded sample of the same points.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "shadow rays/s + bounce wall s on 1 MI355X vs plain FP32 GPU at <=1e-3 rel err"
# FP32 operations per BVH node visit (two child boxes: 12 FMA = 24, 16 min/max,
# 2 widened compares = 4) and per triangle test (IntersectTriangle: 11 mul,
# 10 add/sub, 1 div, 5 compares) -- see the kernel in csrc/vh_trace.hip.
FLOP_PER_NODE = 44
FLOP_PER_TRI = 27
# The light-buffer trace: per ray the cell lookup (3 sub, 3 abs, 2 compares for
# the major axis, 1 div, 2 mul, 2 add, 2 mul to cell units = 15) and per list
# entry read the depth compare (1), besides the triangle tests.
FLOP_PER_LOOKUP = 15
FLOP_PER_ENTRY = 1
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 peak (the same figure for the vector and the matrix pipes)
# the pipe the traversal kernels actually issue on: their FP32 work is vector
# FMA, min/max and compares, with no matrix work; the roofline's "bound" names
# the compute roof ("mfma") and this field says which unit meets it
FP32_PIPE = "valu"
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 MFMA peak


def host_threads():
    """(threads the CPU baselines use, cores in this process's affinity mask):
    every core the process may run on, capped by OMP_NUM_THREADS when the host
    sets one (a GPU box's share of its cores)."""
    cores = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    used = min(cores, int(cap)) if cap.isdigit() and int(cap) > 0 else cores
    return used, cores


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def read_traffic(kernel: str, key: str = "hbm_bytes_per_launch"):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(rf"^{key}:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def read_profile_value(name: str, key: str):
    """A figure from a committed rocprofv3 summary (profiles/<name>.txt), if any."""
    path = os.path.join(ROOT, "profiles", f"{name}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(rf"^{key}:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(w, gpu_bits, target_s: float, threads: int):
    import oracle

    tris = oracle.convert_triangles(w.casters.vertices, w.casters.ids)
    bvh = oracle.CpuBvh(w.casters.vertices, tris)
    probe = min(2000, w.points.shape[0])
    t0 = time.perf_counter()
    bvh.point_lights(w.points[:probe], w.lights, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    n = int(min(w.points.shape[0], max(probe, probe * target_s / dt)))
    t0 = time.perf_counter()
    want = bvh.point_lights(w.points[:n], w.lights, threads=threads)
    dt = time.perf_counter() - t0
    rays = n * w.lights.shape[0]
    words = (rays + 63) // 64
    mismatched_words = int(np.count_nonzero(gpu_bits[:words] != want))
    return {
        "value": rays / dt,
        "unit": "rays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} of {w.points.shape[0]} points x {w.lights.shape[0]} lights "
        f"({rays} rays, {dt:.1f} s); CPU restatement of the reference's any-hit "
        "visibility (exact KD triangle test, BVH traversal), pthreads",
    }, mismatched_words


def bounce_leg(a, log):
    """The bounce config: 12 x 12 tiled room grid, ~200k patches.  GPU transfer
    build, direct light from the rooms' point lights, then the bounce to
    relative residual 1e-4 with FP32 and with FP16 transfer weights."""
    import torch

    from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, rooms
    from hl2sdk_tools_sharp_amd.bounce import HipBounceStage
    from hl2sdk_tools_sharp_amd.tracer import unpack_bits

    t0 = time.perf_counter()
    g = rooms.bounce_grid(a.bounce_rooms, a.bounce_rooms)
    log(f"[bench] bounce grid {a.bounce_rooms}x{a.bounce_rooms}: {g.npatch} patches "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    offer = HipTracerFactory(device=0).try_create(g.casters)
    if offer.tracer is None:
        raise RuntimeError(f"HIP tracer declined: {offer.decline_reason}")
    tr = offer.tracer
    st = HipBounceStage(tr, g)
    st.load_patches()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T = st.build_transfers()
    build_s = time.perf_counter() - t0
    log(f"[bench] transfer build: {T} transfers ({T / len(g.receivers):.0f}/row) from "
        f"{st.candidate_count} candidate rays ({st.blocked_count} blocked) in {build_s:.3f} s")
    pts = rooms.sample_points(g)
    bits = tr.trace_point_lights(pts, g.lights)
    e0 = rooms.direct_light(g, unpack_bits(bits, g.npatch * len(g.lights)))
    st.bounce(e0, max_iter=2, tol=0.0)  # first-use allocations outside the timing
    out = {"patches": g.npatch, "transfers": T, "transfers_per_row": T / len(g.receivers),
           "candidate_rays": st.candidate_count, "transfer_build_s": build_s}
    results = {}
    for wname in ("f32", "f16"):
        t0 = time.perf_counter()
        r = st.bounce(e0, weights=wname, tol=1e-4, max_iter=200)
        wall = time.perf_counter() - t0
        results[wname] = r
        out[f"wall_s_{wname}"] = wall
        out[f"iterations_{wname}"] = r.iterations
        log(f"[bench] bounce {wname}: {r.iterations} iterations to residual "
            f"{r.residuals[-1]:.2e} in {wall * 1e3:.2f} ms")
    # the gather kernel alone, from HIP events, without the graph
    L = _lib.lib()
    _lib.check(L.vh_timing_enable(tr.handle, 1))
    for wname in ("f32", "f16"):
        st.bounce(e0, weights=wname, tol=0.0, max_iter=a.bounce_timed_iters, use_graph=False)
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _lib.check(L.vh_timing_read(tr.handle, _lib.VH_KERNEL_GATHER, ctypes.byref(ms), ctypes.byref(n)))
        per = ms.value / max(n.value, 1)
        wbytes = 4 if wname == "f32" else 2
        algo = T * (4 + wbytes) + g.npatch * 24 + (wname == "f16") * 4 * g.npatch
        out[f"gather_ms_{wname}"] = per
        out[f"gather_GBps_{wname}"] = algo / (per / 1e3) / 1e9
        out[f"gather_bytes_{wname}"] = algo
        log(f"[bench] gather {wname}: {per:.4f} ms, {algo / (per / 1e3) / 1e9:.0f} GB/s algorithmic")
    _lib.check(L.vh_timing_enable(tr.handle, 0))
    a32 = results["f32"].total_light.astype(np.float64)
    out["f16_vs_f32_rel_l1"] = float(np.abs(results["f16"].total_light - a32).sum() / np.abs(a32).sum())
    # the HBM roofline of the gather in its FP16 form (the optimised bounce's)
    ach = out["gather_GBps_f16"]
    out["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": read_traffic("k_gather", "hbm_bytes_f16"),
                       "kernel": "k_gather<true>", "kernel_ms": round(out["gather_ms_f16"], 4),
                       "algorithmic_bytes": out["gather_bytes_f16"],
                       "fp32": {"achieved": round(out["gather_GBps_f32"], 1),
                                "frac": round(out["gather_GBps_f32"] / PEAK_HBM_GBS, 4),
                                "traffic": read_traffic("k_gather", "hbm_bytes_f32")}}
    if not a.no_cpu:
        # the CPU restatement of Radiosity.GatherLight on the same FP32
        # transfers, row-parallel, for the FP32 bounce's iteration count
        import oracle.bounce as ob

        threads, cores = host_threads()
        rp, src, w = st.read_transfers()
        shoot = (np.asarray(e0, np.float32) * g.refl).astype(np.float32)
        iters = results["f32"].iterations
        first = ob.gather_light(rp, src, w, shoot, threads)  # pages the arrays in
        t0 = time.perf_counter()
        for _ in range(iters):
            ob.gather_light(rp, src, w, shoot, threads)
        dt = time.perf_counter() - t0
        per = dt / iters
        out["cpu_baseline"] = {
            "value": dt, "unit": "s", "cores": threads, "affinity_cores": cores, "kind": "port",
            "sample": f"{iters} GatherLight passes (the FP32 bounce's iteration count) over all {T} transfers "
            f"of the {g.npatch}-patch map; CPU restatement of Radiosity.GatherLight, FP32, rows over pthreads",
            "per_iteration_ms": per * 1e3, "GBps": out["gather_bytes_f32"] / per / 1e9}
        log(f"[bench] bounce CPU GatherLight {per * 1e3:.1f} ms per pass on {threads} threads "
            f"({out['gather_bytes_f32'] / per / 1e9:.1f} GB/s); {iters} passes {dt:.2f} s")
        del rp, src, w, first
    tr.dispose()
    return out


SUN_EXTENT_HALF_DEGREE = math.sin(math.radians(0.5))


def batching_leg(a, log):
    """The batching config: 4,096 faces of the visibility scene with luxel
    counts log-uniform over 4-16,384; per luxel 162 sky-ambient rays, 64
    jittered sun-disk rays (extent 0.5 degrees) and 32 AO rays (radius 40).
    One persistent-queue launch (waves per CU from the autotuner) against one
    launch per face; rays counted by the kernel.  The CPU baseline is the
    oracle on a seeded face subsample (10 % by default), whose GPU result it
    also checks bit for bit."""
    import oracle.gather as og
    from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, autotune, light, scenes

    t0 = time.perf_counter()
    casters, faces = scenes.batching_workload(nfaces=a.batch_faces, seed=4)
    rng = np.random.default_rng(4)
    z = rng.uniform(0.0, 1.0)
    phi = rng.uniform(0.0, 2 * math.pi)
    to_sun = np.array([math.sqrt(1 - z * z) * math.cos(phi), math.sqrt(1 - z * z) * math.sin(phi), z])
    L = light.make_lights(2)
    L[0]["type"], L[0]["normal"] = light.EMIT_SKYLIGHT, -to_sun
    L[0]["intensity"], L[0]["intensity_hdr"] = (1.0, 0.9, 0.8), (2.0, 1.8, 1.6)
    L[1]["type"], L[1]["intensity"], L[1]["intensity_hdr"] = light.EMIT_SKYAMBIENT, (0.2, 0.25, 0.3), (0.4, 0.5, 0.6)
    kw = dict(sun_extent=SUN_EXTENT_HALF_DEGREE, sun_samples=64, sky_dirs=light.sky_directions(),
              ao_count=32, ao_radius=40.0)
    nominal = faces.nsamples * (162 + 64 + 32)
    log(f"[bench] batching: {faces.nfaces} faces, {faces.nsamples} luxels, {nominal} nominal rays "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    offer = HipTracerFactory(device=0).try_create(casters)
    if offer.tracer is None:
        raise RuntimeError(f"HIP tracer declined: {offer.decline_reason}")
    tr = offer.tracer
    fl = light.HipFaceLighter(tr)
    fl.load(faces, L, **kw)
    fl.run("queue")  # first-use allocations outside the timing
    tune = autotune.tune_gather(fl, use_cache=False, log=log)
    wpc, grain = tune.waves_per_cu, tune.chunks_per_ticket
    q_ms, rays = [], 0
    for _ in range(a.batch_steps):
        ms, _n, r = fl.timed_run("queue", wpc, grain)
        q_ms.append(ms)
        rays = r
    q = float(np.median(q_ms))
    # one counted launch (outside the timing): node visits, triangle tests and
    # the lane slots the packets occupied
    L_ = _lib.lib()
    _lib.check(L_.vh_count_enable(tr.handle, 1))
    fl.run("queue", wpc, grain)
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L_.vh_count_read(tr.handle, ctypes.byref(nv), ctypes.byref(nt)))
    ns, ts = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(L_.vh_count_read_slots(tr.handle, ctypes.byref(ns), ctypes.byref(ts)))
    _lib.check(L_.vh_count_enable(tr.handle, 0))
    flops = nv.value * FLOP_PER_NODE + nt.value * FLOP_PER_TRI
    ach = flops / (q / 1e3) / 1e12
    lanes = (nv.value + nt.value) / max(64 * (ns.value + ts.value), 1)
    roofline = {"bound": "mfma", "pipe": FP32_PIPE, "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s",
                "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": None, "kernel": "k_facelight<true>",
                "kernel_ms": round(q, 3), "flop_per_launch": flops,
                "node_visits_per_ray": round(nv.value / max(rays, 1), 3),
                "triangle_tests_per_ray": round(nt.value / max(rays, 1), 3),
                "active_lane_fraction": round(lanes, 4),
                "node_packet_efficiency": round(nv.value / max(64 * ns.value, 1), 4)}
    log(f"[bench] batching roofline: {ach:.2f} TFLOP/s ({ach / PEAK_FP32_TFLOPS:.1%} of FP32 peak); "
        f"active-lane fraction {lanes:.3f} over node and triangle slots")
    ref = fl.read()
    pf_ms, _n, pf_rays = fl.timed_run("per_face")
    same = bool(np.array_equal(fl.read().ldr.view(np.uint32), ref.ldr.view(np.uint32)))
    log(f"[bench] batching queue (waves/CU {wpc}, {grain} chunks/ticket): {q:.1f} ms, {rays / (q / 1e3) / 1e9:.2f} Grays/s; "
        f"per-face launches {pf_ms:.1f} ms; bitwise equal: {same}")
    del ref
    out = {
        "faces": faces.nfaces, "luxels": faces.nsamples, "rays_nominal": nominal, "rays_traced": rays,
        "queue_ms": q, "queue_rays_per_s": rays / (q / 1e3), "waves_per_cu": wpc, "chunks_per_ticket": grain,
        "tune_ms": {str(k): round(v, 3) for k, v in tune.ms.items()},
        "tune_grain_ms": {k: round(v, 3) for k, v in tune.grain_ms.items()},
        "per_face_ms": pf_ms, "per_face_rays_per_s": pf_rays / (pf_ms / 1e3),
        "queue_speedup": pf_ms / q, "queue_equals_per_face": same, "roofline": roofline,
    }
    # a seeded face subsample: the CPU baseline and a parity check at the config's scale
    nsub = max(1, int(round(faces.nfaces * a.batch_cpu_fraction)))
    sub = np.sort(np.random.default_rng(5).choice(faces.nfaces, nsub, replace=False))
    fs = faces.subset(sub)
    fl.load(fs, L, **kw)
    _ms, _n, sub_rays = fl.timed_run("queue", wpc, grain)
    got = fl.read()
    tr.dispose()
    if not a.no_cpu:
        threads, cores = host_threads()
        d = light.pack_desc(fs, L, None, kw["sun_extent"], kw["sun_samples"], kw["sky_dirs"],
                            kw["ao_count"], kw["ao_radius"])
        t0 = time.perf_counter()
        ldr, hdr, ao = og.gather(casters, d, fs.nsamples, threads=threads)
        dt = time.perf_counter() - t0
        diff = int(np.count_nonzero(ldr.view(np.uint32) != got.ldr.view(np.uint32))
                   + np.count_nonzero(hdr.view(np.uint32) != got.hdr.view(np.uint32))
                   + np.count_nonzero(ao != got.ao))
        out["cpu_baseline"] = {
            "value": sub_rays / dt, "unit": "rays/s", "cores": threads, "affinity_cores": cores, "kind": "port",
            "sample": f"{len(sub)} of {faces.nfaces} faces ({fs.nsamples} luxels, {sub_rays} rays, "
            f"{dt:.1f} s); CPU restatement of the reference's gather (sun, sky ambient) plus AO, pthreads",
        }
        out["subset_values_differing"] = diff
        log(f"[bench] batching CPU {sub_rays / dt / 1e6:.2f} Mrays/s on {threads} threads; "
            f"subset values differing from GPU: {diff}")
    return out


def decode_lump(lump: np.ndarray) -> np.ndarray:
    """ColorRGBExp32 words back to linear RGB: byte * 2^exponent."""
    w = np.asarray(lump, np.uint8).reshape(-1, 4)
    return w[:, :3].astype(np.float64) * np.exp2(w[:, 3].view(np.int8).astype(np.float64))[:, None]


def lump_rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    """Relative L1 distance of two lumps' decoded HDR values (b the reference)."""
    x, y = decode_lump(a), decode_lump(b)
    return float(np.abs(x - y).sum() / max(np.abs(y).sum(), 1e-300))


def final_lump_cpu(fin, faces, light, total, threads):
    """The final stage restated on the CPU (vradcpu_final.c) with bounced
    light ``total`` (P x 3, flat)."""
    import oracle.final as of

    ff, fp = fin
    pl = np.zeros((fp.npatch, 4, 3), np.float32)
    pl[:, 0] = total
    pd = {"origin": fp.origin, "face": fp.face, "child1": fp.child1, "winding_ptr": fp.winding_ptr,
          "winding": fp.winding}
    return of.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                            faces.face_sample_ptr, faces.face_normal_count, faces.sample_pos, light, pd, pl,
                            threads=threads)


def plain_vrad_leg(a, log):
    """The plain-vrad config: a 3 x 3 grid of 512 x 512 x 256 rooms (8-unit
    walls, 128 x 128 doorways) under a sky ceiling, 8 point lights
    (intensity 50-300, 1/d^2) and one sun, luxel 16, patch 128, FP32, bounce
    to relative residual 1e-4.  The GPU pass (scene, gather, patch lights,
    transfer build, bounce, lump) is timed stage by stage, after one untimed
    pass; the CPU baseline is the oracle's restatement of the same stages on
    the host's cores, and the bounced light of the two is compared."""
    import oracle.bounce as ob
    import oracle.gather as og
    from hl2sdk_tools_sharp_amd import light, rooms
    from hl2sdk_tools_sharp_amd.vrad import light_pass

    g = rooms.room_grid(3, 3, size=(512, 512, 256), wall=8, door=128, patch=128, minchop=64,
                        nlights=8, sky_ceiling=True, seed=1)
    faces = rooms.room_faces(g, luxel=16.0)
    lights, pvs = rooms.room_lights(g, sun=True)
    fin = rooms.room_final_inputs(g, faces)
    kw = dict(sun_extent=0.0, sun_samples=1, tol=1e-4, final_inputs=fin)
    log(f"[bench] plain vrad: {g.npatch} patches, {faces.nsamples} luxels, {len(lights)} lights")
    light_pass(g, faces, lights, pvs, **kw)  # first-use costs outside the timing
    runs = [light_pass(g, faces, lights, pvs, read_direct=True, **kw) for _ in range(3)]
    best = min(runs, key=lambda r: r.wall_s)
    out = {"patches": g.npatch, "luxels": faces.nsamples, "lights": len(lights),
           "transfers": best.transfers, "iterations": best.bounce.iterations,
           "wall_s": best.wall_s, "stage_s": {k: round(v, 5) for k, v in best.stage_s.items()}}
    log(f"[bench] plain vrad GPU pass {best.wall_s * 1e3:.1f} ms: "
        + ", ".join(f"{k} {v * 1e3:.1f}" for k, v in best.stage_s.items()))
    if not a.no_cpu:
        threads, cores = host_threads()
        t0 = time.perf_counter()
        d = light.pack_desc(faces, lights, pvs, 0.0, 1, None)
        ldr, _hdr, _ao = og.gather(g.casters, d, faces.nsamples, threads=threads)
        direct = ob.patch_lights(g, faces, ldr[0], faces.sample_area)
        rp, src, w, _nc = ob.build_transfers(g, threads=threads)
        tot, _res, _en, it = ob.bounce(g, rp, src, w, direct, tol=1e-4, max_iter=100)
        lump, _ = final_lump_cpu(fin, faces, ldr, tot, threads)
        dt = time.perf_counter() - t0
        # the GPU's final stage on the GPU's own bounce, restated on the CPU
        lump_same_bounce, _ = final_lump_cpu(fin, faces, ldr, best.bounce.total_light, threads)
        b = best.bounce.total_light.astype(np.float64)
        out["cpu_baseline"] = {
            "value": dt, "unit": "s", "cores": threads, "affinity_cores": cores, "kind": "port",
            "sample": "the whole pass (gather, patch lights, transfer build, bounce, final light "
            "and lump); CPU restatement of the reference's stages, pthreads where the stage is parallel",
        }
        out["speedup_vs_cpu"] = dt / best.wall_s
        out["lump_equal"] = bool(np.array_equal(lump_same_bounce, best.lump))
        out["lump_rel_l1_vs_cpu"] = lump_rel_l1(best.lump, lump)
        out["direct_equal"] = bool(np.array_equal(direct.view(np.uint32), best.direct.view(np.uint32)))
        out["bounce_rel_l1_vs_cpu"] = float(np.abs(b - tot).sum() / max(np.abs(tot).sum(), 1e-30))
        out["cpu_iterations"] = it
        log(f"[bench] plain vrad CPU {dt:.2f} s on {threads} threads; lump equal on the same bounce "
            f"{out['lump_equal']}, direct equal {out['direct_equal']}, bounce rel L1 "
            f"{out['bounce_rel_l1_vs_cpu']:.2e}, lump rel L1 {out['lump_rel_l1_vs_cpu']:.2e}")
    return out


def _grid_for_seed(seed: int):
    """The warm-start maps: the bounce config's rooms and patches on an
    n x n grid, n = 6 + seed mod 7 (50k-200k patches)."""
    from hl2sdk_tools_sharp_amd import rooms

    return rooms.bounce_grid(6 + seed % 7, 6 + seed % 7, seed=seed)


def warm_start_leg(a, log):
    """The learned warm start: bounces of seeded maps (the bounce config's
    generator, seeds 100.., 50k-200k patches) become the network's data set
    on the GPU; the network trains there (BF16 MFMA, hipGraph steps, Adam);
    on held-out maps (seeds 200..) the bounce to residual 1e-4 runs from a
    zero start, from the direct-only prior and from the network, with FP32
    and FP16 transfer weights.  Maps are generated on the host in worker
    processes."""
    import multiprocessing as mp

    from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib
    from hl2sdk_tools_sharp_amd.mlp import Prior, WarmStartMLP, bounce_map, init_params

    train_seeds = list(range(100, 100 + a.mlp_maps))
    test_seeds = list(range(200, 200 + a.mlp_test_maps))
    t0 = time.perf_counter()
    # at most 8 worker processes, ended by close and join (no terminate), so a
    # profiled run never holds more than a handful of processes
    pool = mp.get_context("spawn").Pool(min(8, host_threads()[0]))
    try:
        maps = pool.map(_grid_for_seed, train_seeds + test_seeds)
    finally:
        pool.close()
        pool.join()
    gen_s = time.perf_counter() - t0
    train, test = maps[:len(train_seeds)], maps[len(train_seeds):]
    log(f"[bench] warm start: {len(train)} training maps ({sum(g.npatch for g in train)} patches), "
        f"{len(test)} held out, generated in {gen_s:.1f} s")
    offer = HipTracerFactory(device=0).try_create(train[0].casters)
    if offer.tracer is None:
        raise RuntimeError(f"HIP tracer declined: {offer.decline_reason}")
    tr = offer.tracer
    net = WarmStartMLP(tr)
    net.dataset_clear()
    prior = Prior()
    t0 = time.perf_counter()
    rows = 0
    for g in train:
        _st, e0, r = bounce_map(tr, g, tol=1e-6)
        rows = net.dataset_add()
        prior.add(g, e0, r.total_light)
    data_s = time.perf_counter() - t0
    net = WarmStartMLP(tr, params=init_params(0), gain=prior.gain, scale=prior.scale)
    L = _lib.lib()
    _lib.check(L.vh_timing_enable(tr.handle, 1))
    t0 = time.perf_counter()
    losses = net.train(epochs=a.mlp_epochs, batch=a.mlp_batch, lr=a.mlp_lr, seed=1)
    train_s = time.perf_counter() - t0
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _lib.check(L.vh_timing_read(tr.handle, _lib.VH_KERNEL_MLP, ctypes.byref(ms), ctypes.byref(n)))
    _lib.check(L.vh_timing_enable(tr.handle, 0))
    steps = (rows // a.mlp_batch) * a.mlp_epochs
    flop = 6.0 * 36352 * a.mlp_batch * steps  # 2 forward + 4 backward FLOP per weight per row
    params = net.params()
    # the per-layer kernels on the same data and seed: the same parameters, bit for bit
    net2 = WarmStartMLP(tr, params=init_params(0), gain=prior.gain, scale=prior.scale)
    _lib.check(L.vh_timing_enable(tr.handle, 1))
    t0 = time.perf_counter()
    net2.train(epochs=a.mlp_epochs, batch=a.mlp_batch, lr=a.mlp_lr, seed=1, fused=False)
    layer_s = time.perf_counter() - t0
    ms2 = ctypes.c_double()
    _lib.check(L.vh_timing_read(tr.handle, _lib.VH_KERNEL_MLP, ctypes.byref(ms2), ctypes.byref(n)))
    _lib.check(L.vh_timing_enable(tr.handle, 0))
    same_params = bool(np.array_equal(params.view(np.uint32), net2.params().view(np.uint32)))
    del net2
    net = WarmStartMLP(tr, params=params, gain=prior.gain, scale=prior.scale)
    tflops = flop / (ms.value / 1e3) / 1e12
    us_step = ms.value * 1e3 / max(steps, 1)
    log(f"[bench] warm start: {rows} rows from {len(train)} maps in {data_s:.1f} s; {steps} steps in "
        f"{train_s:.2f} s, {us_step:.1f} us per step on the device ({tflops:.1f} TFLOP/s); per-layer kernels "
        f"{ms2.value * 1e3 / max(steps, 1):.1f} us per step, parameters bit-equal: {same_params}; "
        f"loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    out = {"train_maps": len(train), "test_maps": len(test), "rows": rows, "epochs": a.mlp_epochs,
           "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": PEAK_BF16_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(tflops / PEAK_BF16_TFLOPS, 5), "traffic": None,
                        "kernel": "the fused training step (k_step_rows + k_step_weights, one hipGraph per step)",
                        "kernel_ms_per_step": round(ms.value / max(steps, 1), 5),
                        "flop_per_step": 6.0 * 36352 * a.mlp_batch},
           "us_per_step": us_step, "per_layer_us_per_step": ms2.value * 1e3 / max(steps, 1),
           "per_layer_train_s": layer_s, "fused_equals_per_layer": same_params,
           "batch": a.mlp_batch, "train_s": train_s, "train_tflops": tflops,
           "data_s": data_s, "generate_s": gen_s, "prior_gain": prior.gain, "prior_scale": prior.scale,
           "loss_first": float(losses[0]), "loss_last": float(losses[-1]), "maps": []}
    for g in test:
        st, e0, _r = bounce_map(tr, g, tol=1e-4, max_iter=2)
        leaf = (g.leaf & ~g.sky.astype(bool))[:, None]
        res = {"patches": g.npatch}
        ref = st.bounce(e0, tol=1e-7, max_iter=300).total_light.astype(np.float64)
        for wname in ("f32", "f16"):
            for start in ("zero", "direct", "network"):
                t0 = time.perf_counter()
                if start == "network":
                    net.predict(e0, read=False)
                    r = st.bounce(e0, tol=1e-4, max_iter=300, weights=wname, from_prediction=True)
                else:
                    x0 = np.zeros_like(e0) if start == "zero" else (prior.gain * e0 * leaf).astype(np.float32)
                    r = st.bounce(e0, start=x0, tol=1e-4, max_iter=300, weights=wname)
                wall = time.perf_counter() - t0
                err = float(np.abs(r.total_light - ref).sum() / np.abs(ref).sum())
                res[f"{wname}_{start}"] = {"iterations": r.iterations, "wall_s": wall, "rel_l1_vs_tight": err}
        out["maps"].append(res)
        log(f"[bench] held-out map {g.npatch} patches: iterations f32 zero/direct/network "
            f"{res['f32_zero']['iterations']}/{res['f32_direct']['iterations']}/{res['f32_network']['iterations']}, "
            f"f16 {res['f16_zero']['iterations']}/{res['f16_direct']['iterations']}/{res['f16_network']['iterations']}")
    for key in ("f32_zero", "f32_direct", "f32_network", "f16_zero", "f16_direct", "f16_network"):
        out[f"mean_iterations_{key}"] = float(np.mean([m[key]["iterations"] for m in out["maps"]]))
        out[f"mean_wall_s_{key}"] = float(np.mean([m[key]["wall_s"] for m in out["maps"]]))
    model = {"params": net.params(), "gain": prior.gain, "scale": prior.scale}
    tr.dispose()
    return out, model


def final_leg(a, log):
    """The final light at the bake's size: the bounce config's map (12 x 12
    rooms, 202,752 patches, 3,168,000 luxels on 198,000 lit 4 x 4 faces, 288
    point lights), the HDR gather on the GPU, then FinalLightFace.Run for
    every face with the patch radial (each patch's TotalLight a seeded
    stand-in: the stage's work does not depend on the values), timed over
    ``--final-reps`` calls by HIP events on the library's stream.  The CPU
    baseline is vradcpu_final.c on the same samples and patch light over the
    host's cores, and its lump is compared with the GPU's byte for byte.
    Then the leaf-ambient fan: ``--ambient-per-room`` seeded points in each
    room, 162 closest-hit rays each, shaded from that lump on the device; its
    CPU baseline is the oracle on the first ``--ambient-cpu-points`` points,
    whose cubes are compared bit for bit."""
    import oracle.final as of
    from hl2sdk_tools_sharp_amd import HipTracerFactory, rooms
    from hl2sdk_tools_sharp_amd.ambient import HipLeafAmbient
    from hl2sdk_tools_sharp_amd.final import HipFinalLighter
    from hl2sdk_tools_sharp_amd.light import HipFaceLighter

    g = rooms.bounce_grid(12, 12)
    faces = rooms.room_faces(g, luxel=16.0)
    lights, pvs = rooms.room_lights(g, sun=False)
    ff, fp = rooms.room_final_inputs(g, faces)
    rng = np.random.default_rng(11)
    pl = np.zeros((g.npatch, 4, 3), np.float32)
    pl[:, 0] = rng.uniform(0, 50, (g.npatch, 3))
    cnt = np.diff(faces.face_sample_ptr).astype(np.int64)
    lit = ff.faces["lit"] == 1
    n = (ff.faces["width"].astype(np.int64) * ff.faces["height"])
    seg = np.repeat(np.arange(len(cnt)), np.diff(ff.nbr_ptr))
    nbr_items = np.bincount(seg, weights=cnt[ff.nbr], minlength=len(cnt)).astype(np.int64)
    tiles = (n + 63) // 64
    steps = int(np.sum(np.where(lit, tiles * (cnt + nbr_items), 0)))
    log(f"[bench] final light: {int(lit.sum())} lit faces, {int(n[lit].sum())} luxels, {g.npatch} patches; "
        f"{steps} wave-item steps of the direct radial")
    offer = HipTracerFactory(device=0).try_create(g.casters)
    with offer.tracer as tr:
        fl = HipFaceLighter(tr)
        fl.load(faces, lights, light_pvs=pvs, sun_samples=1)
        fl.run()
        fr = HipFinalLighter(fl)
        t0 = time.perf_counter()
        fr.load(ff, fp)
        load_s = time.perf_counter() - t0
        lump = fr.run(hdr=True, bounce="host", patch_light=pl)  # first-use costs
        kms = []
        for _ in range(a.final_reps):
            lump_t, ms = fr.timed_run(hdr=True, bounce="host", patch_light=pl)
            kms.append(ms)
        t0 = time.perf_counter()
        fr.run(hdr=True, bounce="host", patch_light=pl)
        wall_s = time.perf_counter() - t0
        light = fl.read().hdr
        # the leaf-ambient fan over the lump the final stage left on the device
        offs, _ = fr.layout()
        afaces = rooms.room_ambient_faces(g, ff, offs)
        pts = rooms.room_ambient_points(g, per_room=a.ambient_per_room, seed=5)
        amb = HipLeafAmbient(tr)
        amb.load(afaces)
        amb.cubes(pts[:64])  # first-use costs
        ams = []
        for _ in range(a.final_reps):
            cubes, ms = amb.timed_cubes(pts)
            ams.append(ms)
        # supersampling (FaceLightJob's gradient passes) of an HDR gather of
        # the same map with every light 10^4 times as bright: the map's own
        # lights leave luxels well under 1 and every gradient under 0.0625;
        # these put them where lightmaps sit (median R + G + B about 16,
        # 90th percentile about 100)
        bright = lights.copy()
        for key in ("intensity", "intensity_hdr"):
            bright[key] *= np.float32(1e4)
        fl.load(faces, bright, light_pvs=pvs, sun_samples=1)
        fl.run()
        ss_base = fl.read().hdr
        ss_counts, ss_ms, ss_rays = fl.timed_supersample(ff.faces, which=1)
        ss_light = fl.read().hdr
    kernel_ms = float(np.mean(kms))
    nlux = int(n[lit].sum())
    out = {"lit_faces": int(lit.sum()), "luxels": nlux, "patches": g.npatch, "wave_item_steps": steps,
           "kernel_ms": kernel_ms, "kernel_ms_runs": [round(x, 4) for x in kms], "call_wall_ms": wall_s * 1e3,
           "load_ms": load_s * 1e3, "luxels_per_s": nlux / (kernel_ms * 1e-3),
           "lump_bytes": int(lump.shape[0]), "repeat_equal": bool(np.array_equal(lump, lump_t))}
    log(f"[bench] final light: kernels {kernel_ms:.3f} ms ({out['luxels_per_s'] / 1e9:.2f} G luxels/s), "
        f"call {wall_s * 1e3:.1f} ms, load {load_s * 1e3:.1f} ms")
    amb_ms = float(np.mean(ams))
    nrays = len(pts) * 162
    out["ambient"] = {"points": len(pts), "rays": nrays, "kernel_ms": amb_ms,
                      "kernel_ms_runs": [round(x, 4) for x in ams], "rays_per_s": nrays / (amb_ms * 1e-3),
                      "points_per_s": len(pts) / (amb_ms * 1e-3)}
    log(f"[bench] leaf ambient: {len(pts)} points x 162 rays in {amb_ms:.3f} ms "
        f"({out['ambient']['rays_per_s'] / 1e9:.2f} G rays/s)")
    out["supersample"] = {"lights": "the map's 288 lights x 10^4", "samples": faces.nsamples,
                          "passes": ss_counts, "kernel_ms": ss_ms, "rays": int(ss_rays),
                          "rays_per_s": ss_rays / (ss_ms * 1e-3) if ss_ms > 0 else None,
                          "samples_changed": int(np.count_nonzero(np.any(ss_light != ss_base, axis=(0, 2))))}
    log(f"[bench] supersampling: passes re-lit {ss_counts} samples, {ss_rays} rays in {ss_ms:.3f} ms")
    if not a.no_cpu:
        threads, cores = host_threads()
        pd = {"origin": fp.origin, "face": fp.face, "child1": fp.child1, "winding_ptr": fp.winding_ptr,
              "winding": fp.winding}
        t0 = time.perf_counter()
        lump_c, _ = of.final_light_c(ff.faces, ff.luxel_pos, ff.sample_bounds, ff.nbr_ptr, ff.nbr,
                                     faces.face_sample_ptr, faces.face_normal_count, faces.sample_pos, light,
                                     pd, pl, threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": dt * 1e3, "unit": "ms", "cores": threads, "affinity_cores": cores, "kind": "port",
            "sample": "the whole map's final light (radial filter, patch radial, minlight, medians, RGBE); "
            "CPU restatement of FinalLightFace.Run, faces over pthreads"}
        out["lump_equal_cpu"] = bool(np.array_equal(lump_c, lump))
        log(f"[bench] final light CPU {dt * 1e3:.0f} ms on {threads} threads; lump equal: {out['lump_equal_cpu']}")
        import oracle.ambient as oamb

        k = min(len(pts), a.ambient_cpu_points)
        t0 = time.perf_counter()
        want = oamb.ambient_cubes(g.casters, afaces, None, lump, pts[:k], threads=threads)
        dt = time.perf_counter() - t0
        out["ambient"]["cpu_baseline"] = {
            "value": k * 162 / dt, "unit": "rays/s", "cores": threads, "affinity_cores": cores, "kind": "port",
            "sample": f"the first {k} of {len(pts)} points (x 162 rays); the oracle's BVH closest hit "
            "(pthreads) and the float32 restatement of Accumulate and Project"}
        out["ambient"]["cubes_differing_cpu"] = int(np.sum(cubes[:k].view(np.uint32) != want.view(np.uint32)))
        log(f"[bench] leaf ambient CPU {k * 162 / dt / 1e6:.2f} Mrays/s on {threads} threads; "
            f"cube components differing: {out['ambient']['cubes_differing_cpu']}")
        # supersampling works face by face, so the oracle over the leading
        # faces gives those faces' answer on the whole map
        import oracle.supersample as oss

        nf = min(faces.nfaces, a.supersample_cpu_faces)
        ns = int(faces.face_sample_ptr[nf])
        sub = faces.subset(np.arange(nf))
        t0 = time.perf_counter()
        want, want_counts = oss.supersample(g.casters, sub, bright, pvs, 0.0, 1, None, ff.faces[:nf],
                                            ss_base[:, :ns], which=1, threads=threads)
        dt = time.perf_counter() - t0
        out["supersample"]["cpu_baseline"] = {
            "value": dt * 1e3, "unit": "ms", "cores": threads, "affinity_cores": cores, "kind": "port",
            "sample": f"the first {nf} of {faces.nfaces} faces ({ns} samples): the oracle's passes, "
            "its gather over pthreads"}
        out["supersample"]["values_differing_cpu"] = int(np.sum(
            ss_light[:, :ns].view(np.uint32) != want.view(np.uint32)))
        log(f"[bench] supersampling CPU {dt * 1e3:.0f} ms for {nf} faces (passes {want_counts}); "
            f"values differing: {out['supersample']['values_differing_cpu']}")
    return out


def bake_leg(a, log, model):
    """The 200k-patch bake end to end, both forms on the bounce config's map
    (12 x 12 rooms, 202,752 patches, 288 point lights behind room PVS rows,
    luxel 16): scene, gather (HDR pass), patch lights, transfer build, bounce to
    residual 1e-4, final light (radial filter of the luxels and the bounced
    light) and lump.
      plain:     one gather launch per face, FP32 weights, bounce from zero;
      optimised: the persistent gather queue, FP16 weights, bounce from the
                 warm-start network's prediction (trained in the warm-start leg).
    Each form's bounced light is compared with the oracle's FP64 bounce on the
    same transfers and emission."""
    import oracle.bounce as ob
    from hl2sdk_tools_sharp_amd import HipTracerFactory, rooms
    from hl2sdk_tools_sharp_amd.bounce import HipBounceStage
    from hl2sdk_tools_sharp_amd.vrad import light_pass

    g = rooms.bounce_grid(12, 12)
    faces = rooms.room_faces(g, luxel=16.0)
    lights, pvs = rooms.room_lights(g, sun=False)
    fin = rooms.room_final_inputs(g, faces)
    log(f"[bench] bake: {g.npatch} patches, {faces.nsamples} luxels, {len(lights)} lights")
    kw = dict(hdr=True, sun_samples=1, tol=1e-4, read_direct=True, final_inputs=fin, extra_passes=4)
    light_pass(g, faces, lights, pvs, **kw)  # first-use costs outside the timing
    plain = light_pass(g, faces, lights, pvs, schedule="per_face", weights="f32", **kw)
    batched = light_pass(g, faces, lights, pvs, schedule="grid", weights="f32", **kw)
    fast = light_pass(g, faces, lights, pvs, schedule="queue", weights="f16", warm_start=model, **kw)
    out = {"patches": g.npatch, "luxels": faces.nsamples, "lights": len(lights)}
    forms = (("plain_f32", plain), ("plain_batched_f32", batched), ("optimised_f16_warm", fast))
    for name, r in forms:
        out[name] = {"wall_s": r.wall_s, "iterations": r.bounce.iterations,
                     "supersample_passes": r.supersample_passes,
                     "stage_s": {k: round(v, 5) for k, v in r.stage_s.items()}}
        log(f"[bench] bake {name}: {r.wall_s:.3f} s, {r.bounce.iterations} iterations: "
            + ", ".join(f"{k} {v * 1e3:.1f}" for k, v in r.stage_s.items()))
    out["speedup"] = plain.wall_s / fast.wall_s
    out["speedup_vs_plain_batched"] = batched.wall_s / fast.wall_s
    out["lump_rel_l1_optimised_vs_plain"] = lump_rel_l1(fast.lump, plain.lump)
    out["plain_batched_lump_equals_plain"] = bool(np.array_equal(batched.lump, plain.lump))
    # the bounce alone on this map from each start, FP16 weights as in the
    # optimised form: zero, the direct-light prior (gain x direct light), the
    # network's prediction
    from hl2sdk_tools_sharp_amd.mlp import WarmStartMLP

    offer = HipTracerFactory(device=0).try_create(g.casters)
    with offer.tracer as tr:
        st = HipBounceStage(tr, g)
        st.build_transfers()
        net = WarmStartMLP(tr, params=model["params"], gain=model["gain"], scale=model["scale"])
        e0 = np.ascontiguousarray(plain.direct, np.float32)
        leaf = (g.leaf & ~g.sky.astype(bool))[:, None]
        starts = {}
        for start in ("zero", "direct", "network"):
            t0 = time.perf_counter()
            if start == "network":
                net.predict(e0, read=False)
                r = st.bounce(e0, tol=1e-4, max_iter=300, weights="f16", from_prediction=True)
            else:
                x0 = np.zeros_like(e0) if start == "zero" else (model["gain"] * e0 * leaf).astype(np.float32)
                r = st.bounce(e0, start=x0, tol=1e-4, max_iter=300, weights="f16")
            starts[start] = {"iterations": r.iterations, "wall_s": time.perf_counter() - t0}
        out["bounce_starts_f16"] = starts
        log("[bench] bake map bounce (FP16) from zero / direct prior / network: "
            + " / ".join(f"{v['iterations']} it {v['wall_s'] * 1e3:.1f} ms" for v in starts.values()))
    if not a.no_cpu:
        offer = HipTracerFactory(device=0).try_create(g.casters)
        with offer.tracer as tr:
            st = HipBounceStage(tr, g)
            st.build_transfers()
            rp, src, w = st.read_transfers()
        t0 = time.perf_counter()
        ref, _res, _en, it = ob.bounce(g, rp, src, w, plain.direct, tol=1e-6, max_iter=300, use_double=True)
        out["fp64_oracle_s"] = time.perf_counter() - t0
        out["fp64_oracle_iterations"] = it
        leaf = (g.leaf & ~g.sky.astype(bool))[:, None]
        for name, r in forms:
            d = np.abs((r.bounce.total_light.astype(np.float64) - ref) * leaf).sum()
            out[name]["rel_l1_vs_fp64"] = float(d / max(np.abs(ref * leaf).sum(), 1e-300))
        # the lightmaps: each form's lump against the lump the same final
        # stage makes from the FP64 bounce
        from hl2sdk_tools_sharp_amd.final import HipFinalLighter
        from hl2sdk_tools_sharp_amd.light import HipFaceLighter

        offer = HipTracerFactory(device=0).try_create(g.casters)
        with offer.tracer as tr:
            fl = HipFaceLighter(tr)
            fl.load(faces, lights, light_pvs=pvs, sun_samples=1)
            fl.run()
            fl.supersample(fin[0].faces, which=1, extra_passes=kw["extra_passes"])  # as both forms did
            fr = HipFinalLighter(fl)
            fr.load(*fin)
            pl = np.zeros((g.npatch, 4, 3), np.float32)
            pl[:, 0] = ref
            ref_lump, out["final_kernel_ms"] = fr.timed_run(hdr=True, bounce="host", patch_light=pl)
        for name, r in forms:
            out[name]["lump_hdr_rel_l1_vs_fp64"] = lump_rel_l1(r.lump, ref_lump)
        log(f"[bench] bake vs FP64 oracle ({it} iterations, {out['fp64_oracle_s']:.1f} s): bounced light "
            f"plain {out['plain_f32']['rel_l1_vs_fp64']:.2e}, optimised "
            f"{out['optimised_f16_warm']['rel_l1_vs_fp64']:.2e}; lightmap HDR plain "
            f"{out['plain_f32']['lump_hdr_rel_l1_vs_fp64']:.2e}, optimised "
            f"{out['optimised_f16_warm']['lump_hdr_rel_l1_vs_fp64']:.2e}")
    log(f"[bench] bake speed-up {out['speedup']:.1f}x over per-face launches, "
        f"{out['speedup_vs_plain_batched']:.2f}x over the batched plain form; lightmap HDR rel L1 optimised vs "
        f"plain {out['lump_rel_l1_optimised_vs_plain']:.2e}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--npoints", type=int, default=2_000_000)
    ap.add_argument("--nlights", type=int, default=64)
    ap.add_argument("--nboxes", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bounce", action="store_true")
    ap.add_argument("--bounce-rooms", type=int, default=12)
    ap.add_argument("--bounce-timed-iters", type=int, default=10)
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--no-bake", action="store_true")
    ap.add_argument("--no-final", action="store_true")
    ap.add_argument("--final-reps", type=int, default=5)
    ap.add_argument("--ambient-per-room", type=int, default=128)
    ap.add_argument("--ambient-cpu-points", type=int, default=2048)
    ap.add_argument("--supersample-cpu-faces", type=int, default=20000)
    ap.add_argument("--mlp-maps", type=int, default=64)
    ap.add_argument("--mlp-test-maps", type=int, default=8)
    ap.add_argument("--mlp-epochs", type=int, default=8)
    ap.add_argument("--mlp-batch", type=int, default=4096)
    ap.add_argument("--mlp-lr", type=float, default=2e-3)
    ap.add_argument("--batch-faces", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-cpu-fraction", type=float, default=0.1)
    ap.add_argument("--point-light-method", choices=["auto", "bvh", "lb"], default="auto",
                    help="light buffers (lb) or BVH walks (bvh); auto picks by batch size")
    ap.add_argument("--lb-cells", type=int, default=0, help="light-buffer cells per cube-face edge (0: 64)")
    ap.add_argument("--no-walk", action="store_true", help="skip the BVH-walk comparison leg")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a device copy of this many MiB (profiler calibration)")
    a = ap.parse_args(argv)

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py measures one GPU: every box in this setup has one MI355X and the path "
            "is not sharded across GPUs (the multi-GPU design is written up, not run). "
            "Run with --gpus 1.",
            file=sys.stderr,
        )
        return 2

    import torch

    from hl2sdk_tools_sharp_amd import HipTracerFactory, _lib, scenes

    if not torch.cuda.is_available():
        print("bench.py needs an MI355X (no GPU visible)", file=sys.stderr)
        return 2
    dev = torch.device("cuda:0")
    log(f"[bench] generating workload: {a.nboxes} boxes, {a.npoints} points x {a.nlights} lights")
    w = scenes.visibility_workload(nboxes=a.nboxes, npoints=a.npoints, nlights=a.nlights, seed=2)
    offer = HipTracerFactory(device=0).try_create(w.casters)
    if offer.tracer is None:
        raise RuntimeError(f"HIP tracer declined: {offer.decline_reason}")
    tr = offer.tracer
    info = tr.scene_info()
    log(f"[bench] scene {info}")
    L = _lib.lib()
    h = tr.handle

    pts = torch.from_numpy(w.points).to(dev)
    lts = torch.from_numpy(w.lights).to(dev)
    nrays = w.rays
    words = (nrays + 63) // 64
    bits = torch.zeros(words, dtype=torch.int64, device=dev)
    # A stream of our own (default priority): the library launches on it and
    # the events below are recorded on it, so they bracket exactly its work.
    stream = torch.cuda.Stream(dev)
    torch.cuda.synchronize()
    sptr = ctypes.c_void_p(stream.cuda_stream)
    if a.copy_probe_mib > 0:
        # a plain copy of known size (an elementwise kernel reading and writing
        # the bytes once), the scale for rocprofv3's FETCH_SIZE/WRITE_SIZE
        src = torch.empty(a.copy_probe_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        with torch.cuda.stream(stream):
            for _ in range(3):
                torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    P = ctypes.c_void_p

    def step():
        st = L.vh_trace_point_lights_dev(
            h, P(pts.data_ptr()), w.points.shape[0], P(lts.data_ptr()), w.lights.shape[0], 0.0,
            P(bits.data_ptr()), sptr,
        )
        if st != 0:
            raise RuntimeError(_lib.last_error())

    def counted():
        """Work counts of one launch (outside any timed region)."""
        _lib.check(L.vh_count_enable(h, 1))
        step()
        torch.cuda.synchronize()
        nv, nt = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.vh_count_read(h, ctypes.byref(nv), ctypes.byref(nt)))
        ns, ts = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(L.vh_count_read_slots(h, ctypes.byref(ns), ctypes.byref(ts)))
        _lib.check(L.vh_count_enable(h, 0))
        return nv.value, nt.value, ns.value, ts.value

    def timed(steps):
        """(ms per step, trace kernel ms per launch, build ms per step) over `steps` steps."""
        _lib.check(L.vh_timing_enable(h, 1))
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        total_ms = e0.elapsed_time(e1)
        kms, kl = ctypes.c_double(), ctypes.c_int64()
        _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_POINT_LIGHTS, ctypes.byref(kms), ctypes.byref(kl)))
        bms, bl = ctypes.c_double(), ctypes.c_int64()
        _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_LIGHT_BUFFER, ctypes.byref(bms), ctypes.byref(bl)))
        _lib.check(L.vh_timing_enable(h, 0))
        return total_ms / steps, kms.value / max(kl.value, 1), bms.value / steps

    methods = {"auto": _lib.VH_POINT_LIGHTS_AUTO, "bvh": _lib.VH_POINT_LIGHTS_BVH,
               "lb": _lib.VH_POINT_LIGHTS_LIGHT_BUFFER}
    tr.set_point_light_method(methods[a.point_light_method], a.lb_cells)
    for i in range(a.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{a.warmup}")
    torch.cuda.synchronize()
    _lib.check(L.vh_timing_enable(h, 1))
    step()
    torch.cuda.synchronize()
    bms, bl = ctypes.c_double(), ctypes.c_int64()
    _lib.check(L.vh_timing_read(h, _lib.VH_KERNEL_LIGHT_BUFFER, ctypes.byref(bms), ctypes.byref(bl)))
    _lib.check(L.vh_timing_enable(h, 0))
    light_buffer = bl.value > 0  # which form the method picked for this batch
    kernel = "k_point_lights_lb" if light_buffer else "k_point_lights"
    nv, nt, ns, ts = counted()
    node_eff = nv / max(ns, 1)
    if light_buffer:
        log(f"[bench] light buffers ({a.lb_cells or 64} cells a face edge); per ray: {nv / nrays:.2f} list "
            f"entries read, {nt / nrays:.2f} triangle tests; lane use {node_eff:.3f} (entries per wave step slot)")
    else:
        log(f"[bench] per ray: {nv / nrays:.2f} node visits, {nt / nrays:.2f} triangle tests; "
            f"packet efficiency {node_eff:.3f} (node visits per lane slot)")

    ms_per_step, kernel_ms, build_ms = timed(a.steps)
    rays_per_s = nrays * a.steps / (ms_per_step * a.steps / 1e3)
    log(f"[bench] {ms_per_step:.3f} ms/step, {kernel} {kernel_ms:.3f} ms, light-buffer build {build_ms:.3f} ms, "
        f"{rays_per_s / 1e9:.3f} Grays/s")

    if light_buffer:
        flops = nrays * FLOP_PER_LOOKUP + nv * FLOP_PER_ENTRY + nt * FLOP_PER_TRI
    else:
        flops = nv * FLOP_PER_NODE + nt * FLOP_PER_TRI
    achieved = flops / (kernel_ms / 1e3) / 1e12
    roofline = {
        "bound": "mfma",
        "pipe": FP32_PIPE,
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": read_traffic(kernel),
        "traffic_source": f"profiles/{kernel}_traffic.txt (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)",
        # the limit the kernel meets: vector instruction issue (at most 1 wave-instruction per CU-cycle),
        # from the SQ counters of the committed profile
        "valu_issue_per_cu_cycle": read_profile_value(f"{kernel}_visibility_128M", "valu_issue_per_cu_cycle"),
        "kernel": kernel,
        "kernel_ms": round(kernel_ms, 4),
        "flop_per_launch": flops,
    }
    if light_buffer:
        roofline.update({
            "lookup_flop_per_ray": FLOP_PER_LOOKUP,
            "entries_read_per_ray": round(nv / nrays, 3),
            "triangle_tests_per_ray": round(nt / nrays, 3),
            "wave_step_slots_per_ray": round(ns / nrays, 3),
            "lane_use": round(node_eff, 4),
            "build_ms_per_step": round(build_ms, 4),
            "cells_per_face_edge": a.lb_cells or 64,
        })
    else:
        roofline.update({
            "node_visits_per_ray": round(nv / nrays, 3),
            "triangle_tests_per_ray": round(nt / nrays, 3),
            "node_slots_per_ray": round(ns / nrays, 3),
            "triangle_slots_per_ray": round(ts / nrays, 3),
            "packet_efficiency": round(node_eff, 4),
        })
    main_bits = bits.clone()

    walk = None
    if light_buffer and not a.no_walk:
        # the BVH walk on the same batch: its time and its bits (both forms are
        # exact, so every word must agree)
        tr.set_point_light_method(_lib.VH_POINT_LIGHTS_BVH)
        step()
        torch.cuda.synchronize()
        wv, wt, wsl, _ = counted()
        w_ms, w_kms, _ = timed(max(2, a.steps // 2))
        differ = int((bits != main_bits).sum().item())
        walk = {"ms_per_step": round(w_ms, 4), "kernel_ms": round(w_kms, 4),
                "rays_per_s": nrays / (w_ms / 1e3), "node_visits_per_ray": round(wv / nrays, 3),
                "triangle_tests_per_ray": round(wt / nrays, 3), "packet_efficiency": round(wv / max(wsl, 1), 4),
                "words_differing_from_light_buffer": differ}
        log(f"[bench] BVH walk on the same batch: {w_ms:.3f} ms/step (kernel {w_kms:.3f} ms), "
            f"{nrays / (w_ms / 1e3) / 1e9:.3f} Grays/s; words differing from the light buffer: {differ}")
        tr.set_point_light_method(methods[a.point_light_method], a.lb_cells)
        bits.copy_(main_bits)

    gpu_bits = bits.cpu().numpy().view(np.uint64)
    cpu = None
    parity = None
    if not a.no_cpu:
        threads, cores = host_threads()
        log(f"[bench] CPU baseline on {threads} threads of {cores} cores in the affinity mask "
            f"(~{a.cpu_seconds:.0f} s)")
        cpu, mism = cpu_baseline(w, gpu_bits, a.cpu_seconds, threads)
        cpu["affinity_cores"] = cores
        parity = {"sample_words_differing": mism}
        log(f"[bench] CPU {cpu['value'] / 1e6:.2f} Mrays/s; sample words differing from GPU: {mism}")
    tr.dispose()
    bounce = None if a.no_bounce else bounce_leg(a, log)
    batching = None if a.no_batching else batching_leg(a, log)
    plain = None if a.no_plain else plain_vrad_leg(a, log)
    final = None if a.no_final else final_leg(a, log)
    warm, model = (None, None) if a.no_warm_start else warm_start_leg(a, log)
    bake = None if (a.no_bake or model is None) else bake_leg(a, log, model)

    out = {
        "metric": METRIC,
        "value": rays_per_s,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, seed 2)",
        "config": {
            "workload": f"visibility: {a.nboxes} boxes = {w.casters.count} triangles, "
            f"{w.points.shape[0]} points x {w.lights.shape[0]} point lights = {nrays} shadow rays/step",
            "rays_per_step": nrays,
            "bvh_nodes": info["nodes"],
            "bvh_depth": info["depth"],
            "method": "light buffers" if light_buffer else "BVH walk",
        },
        "roofline": roofline,
        "bvh_walk": walk,
        "cpu_baseline": cpu,
        "parity": parity,
        "bounce_wall_s": None if bounce is None else bounce["wall_s_f32"],
        "bounce": bounce,
        "batching": batching,
        "plain_vrad": plain,
        "final_light": final,
        "warm_start": warm,
        "bake_200k": bake,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
