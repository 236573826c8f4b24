"""Loader for the reference's KD fixtures under tests/golden/ (layout in
tests/golden/README.md; the reference's loader is
src/SourceSharp.Tests/MapTools/Tracing/StockKdScene.cs:73-203)."""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Correct-mode digests: SHA-256 over every triangle's intersection format, then
# every ray's closest hit (VradCpuIndependenceTests.cs:79-83, 119-148).
CORRECT_DIGESTS = {
    "kd-scene": "45088531120712BDD4436079A4E0ED5605F56BB3847084D33B8A01E36F8334E9",
    "kd-boxes": "43A7BF66182DD274ADE62C49DE3D725EF907F10ABD584E7F02E069DF952F891E",
}


@dataclass
class KdScene:
    name: str
    v9: np.ndarray  # (n, 9) float32
    rays: np.ndarray  # (m, 7) float32: start, end - start, 1
    stock_hit: np.ndarray  # (m,) int32 triangle index, -1 = miss
    stock_dist: np.ndarray  # (m,) float32
    stock_normals: np.ndarray  # (n, 3)
    stock_d: np.ndarray  # (n,)
    stock_edges: np.ndarray  # (n, 6)
    stock_cs: np.ndarray  # (n, 2)

    @property
    def stock_blocked(self) -> np.ndarray:
        """Stock's hit strictly short of the end (KdRayTracerStockParityTests.cs:285-302)."""
        return (self.stock_hit >= 0) & (self.stock_dist < 1.0)


def load(name: str) -> KdScene:
    def raw(suffix):
        return np.fromfile(os.path.join(GOLDEN, f"{name}.{suffix}.bin"), dtype=np.uint8)

    tri = raw("tris")
    n = int(tri[:4].view(np.int32)[0])
    v9 = tri[4 : 4 + n * 36].view(np.float32).reshape(n, 9).copy()
    ray = raw("rays")
    m = int(ray[:4].view(np.int32)[0])
    se = ray[4 : 4 + m * 24].view(np.float32).reshape(m, 6)
    rays = np.empty((m, 7), dtype=np.float32)
    rays[:, :3] = se[:, :3]
    rays[:, 3:6] = se[:, 3:6] - se[:, :3]
    rays[:, 6] = 1.0
    ans = raw("answers")
    k = int(ans[:4].view(np.int32)[0])
    assert k == m
    hit = ans[4 : 4 + 4 * k].view(np.int32).copy()
    dist = ans[4 + 4 * k : 4 + 8 * k].view(np.float32).copy()

    tree = raw("tree")
    p = 0

    def i32():
        nonlocal p
        v = int(tree[p : p + 4].view(np.int32)[0])
        p += 4
        return v

    nnodes = i32()
    p += nnodes * 8
    nidx = i32()
    p += nidx * 4
    p += 24
    ntri = i32()
    assert ntri == n
    normals = np.empty((n, 3), np.float32)
    dd = np.empty(n, np.float32)
    edges = np.empty((n, 6), np.float32)
    cs = np.empty((n, 2), np.int32)
    for i in range(n):
        f = tree[p : p + 20].view(np.float32)
        normals[i] = f[:3]
        dd[i] = f[3]
        p += 20
        edges[i] = tree[p : p + 24].view(np.float32)
        p += 24
        cs[i] = tree[p], tree[p + 1]
        p += 2
    return KdScene(name, v9, rays, hit, dist, normals, dd, edges, cs)
