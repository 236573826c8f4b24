"""The light buffers' cell geometry, on the CPU: vh_lightbuf.h's host/device
functions (the ones the build and trace kernels call) must list, in every
ray's cell before its depth, every triangle that blocks the ray, so the light
buffer's answer equals the oracle's brute-force exact test on every ray
(boxes and a triangle soup; lights inside, on and at triangle boxes' faces,
edges, corners and vertices; points on cell and face boundaries; 5 to 120
cells a face edge; tmin 0, 1e-3 and 0.25).  lightbuf_geometry.cpp is compiled
here with g++ against the oracle's libvradcpu.so."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _hip_include():
    for d in ("/opt/rocm/include",):
        if os.path.exists(os.path.join(d, "hip", "hip_runtime.h")):
            return d
    return None


def test_cell_lists_hold_every_blocker(oracle_lib, tmp_path):
    gxx = shutil.which("g++")
    inc = _hip_include()
    if gxx is None or inc is None:
        pytest.skip("needs g++ and the HIP headers (the header's __host__ __device__ markers)")
    oracle_dir = os.path.join(ROOT, "oracle")
    exe = str(tmp_path / "lbg")
    cmd = [gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I" + inc,
           "-I" + os.path.join(ROOT, "hl2sdk_tools_sharp_amd", "csrc"), "-I" + oracle_dir,
           os.path.join(HERE, "lightbuf_geometry.cpp"), "-L" + oracle_dir, "-lvradcpu",
           "-Wl,-rpath," + oracle_dir, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    lines = [ln for ln in r.stdout.splitlines() if "mismatches" in ln]
    assert len(lines) == 5, r.stdout + r.stderr
    for ln in lines:
        fields = ln.split()
        rays = int(fields[fields.index("rays") + 1])
        blocked = int(fields[fields.index("blocked") + 1])
        assert rays > 10000 and 0 < blocked < rays, ln
        assert " mismatches 0 " in ln, ln
    assert r.returncode == 0
