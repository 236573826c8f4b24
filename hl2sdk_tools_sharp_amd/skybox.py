"""The 3D skybox behind the gather's sky rays: the sky-camera table and the
compiled-BSP point query that decide which sky tests continue into a skybox,
and their hand-off to the device (vh_gather_skybox).

Host logic, mirroring the reference:

- ``SkyCameras.build``: SkyCameras.Build (SourceSharp.MapTools/Rad/Light/
  SkyCameras.cs:103-145): every ``sky_camera`` entity (classname compared
  case-insensitively) with ``scale`` > 0 is a camera, world_to_sky = 1 / scale;
  the camera claims the area of the leaf its origin lies in, a later camera
  winning a shared area;
- ``CompiledBspTree.leaf_from_point``: CompiledBspTree.LeafFromPoint
  (Rad/Light/CompiledBspTree.cs): the descent the device repeats per sample;
- ``SkyCameras.leaf_recurses``: DirectLightGatherer.LaneRecurses
  (Rad/Light/DirectLightGatherer.cs:1611-1621) folded into a flag per leaf: the
  leaf's area is one of the map's and no camera stands in it.

The sky tests themselves (TestLineDoesHitSky's second stage) run in the
gather kernel, vh_gather.hip.
"""
from __future__ import annotations

import ctypes
import re
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

_NUMBER = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")


def _scan_numbers(value: str, count: int) -> List[float]:
    """Up to ``count`` leading numbers of ``value``, as a scanf of "%f %f %f"
    reads them, zero-filled where the text runs out (EntityKeys.GetVectorForKey,
    Rad/Light/EntityKeys.cs:119-134)."""
    out, pos = [], 0
    for _ in range(count):
        m = _NUMBER.match(value, pos)
        if not m:
            break
        out.append(float(m.group(1)))
        pos = m.end()
    return out + [0.0] * (count - len(out))


def float_for_key(entity: dict, key: str) -> np.float32:
    """FloatForKey: the key's leading number, 0 for an absent or empty key."""
    return np.float32(_scan_numbers(str(entity.get(key, "")), 1)[0])


def vector_for_key(entity: dict, key: str) -> np.ndarray:
    return np.array(_scan_numbers(str(entity.get(key, "")), 3), np.float64).astype(np.float32)


@dataclass
class CompiledBspTree:
    """The compiled map's nodes and planes (dnodes, dplanes).

    node_plane (N,), node_children (N, 2): >= 0 a node, < 0 the leaf -1 - c;
    plane_normal (P, 3), plane_dist (P,), plane_type (P,): below 3 the plane is
    axial and the query reads one coordinate (0 x, 1 y, any other z)."""

    node_plane: np.ndarray
    node_children: np.ndarray
    plane_normal: np.ndarray
    plane_dist: np.ndarray
    plane_type: np.ndarray

    def __post_init__(self):
        self.node_plane = np.ascontiguousarray(self.node_plane, np.int32).reshape(-1)
        self.node_children = np.ascontiguousarray(self.node_children, np.int32).reshape(-1, 2)
        self.plane_normal = np.ascontiguousarray(self.plane_normal, np.float32).reshape(-1, 3)
        self.plane_dist = np.ascontiguousarray(self.plane_dist, np.float32).reshape(-1)
        self.plane_type = np.ascontiguousarray(self.plane_type, np.int32).reshape(-1)

    def leaf_from_point(self, p) -> int:
        """CompiledBspTree.LeafFromPoint: from node 0, front child (children[0])
        when the point's signed distance is not below zero; a child past the
        last node is leaf 0 (a tree with no nodes answers leaf 0)."""
        p = np.asarray(p, np.float32).reshape(3)
        node = 0
        while node >= 0:
            if node >= len(self.node_plane):
                return 0
            k = int(self.node_plane[node])
            t = int(self.plane_type[k])
            if t < 3:
                dist = np.float32(p[t if t in (0, 1) else 2] - self.plane_dist[k])
            else:
                n = self.plane_normal[k]
                dot = np.float32(np.float32(np.float32(n[0] * p[0]) + np.float32(n[1] * p[1])) + np.float32(n[2] * p[2]))
                dist = np.float32(dot - self.plane_dist[k])
            node = int(self.node_children[node, 1] if dist < 0.0 else self.node_children[node, 0])
        return -1 - node


@dataclass
class SkyCameras:
    """The map's sky cameras (origin, sky_to_world = scale, world_to_sky =
    1 / scale, area) and the camera standing in each area (-1 none)."""

    origin: np.ndarray  # (C, 3)
    sky_to_world: np.ndarray  # (C,)
    world_to_sky: np.ndarray  # (C,)
    area: np.ndarray  # (C,)
    area_camera: np.ndarray  # (areas,)

    @staticmethod
    def build(entities: Sequence[dict], tree: CompiledBspTree, leaf_area: np.ndarray,
              area_count: int) -> "SkyCameras":
        leaf_area = np.asarray(leaf_area, np.int32)
        area_camera = np.full(max(int(area_count), 0), -1, np.int32)
        cams = []
        for ent in entities:
            if str(ent.get("classname", "")).lower() != "sky_camera":
                continue
            origin = vector_for_key(ent, "origin")
            leaf = tree.leaf_from_point(origin)
            area = int(leaf_area[leaf]) if 0 <= leaf < len(leaf_area) else -1
            scale = float_for_key(ent, "scale")
            if scale > 0.0:
                if 0 <= area < len(area_camera):
                    area_camera[area] = len(cams)
                cams.append((origin, scale, np.float32(np.float32(1.0) / scale), area))
        return SkyCameras(
            np.array([c[0] for c in cams], np.float32).reshape(-1, 3),
            np.array([c[1] for c in cams], np.float32),
            np.array([c[2] for c in cams], np.float32),
            np.array([c[3] for c in cams], np.int32),
            area_camera)

    def camera_in_area(self, area: int) -> int:
        return int(self.area_camera[area]) if 0 <= area < len(self.area_camera) else -1

    def leaf_recurses(self, leaf_area: np.ndarray) -> np.ndarray:
        """Per leaf, 1 when a sky test starting there continues into the
        skyboxes: its area is one of the map's and no camera stands in it."""
        leaf_area = np.asarray(leaf_area, np.int64).reshape(-1)
        ok = (leaf_area >= 0) & (leaf_area < len(self.area_camera))
        cam = np.full(leaf_area.shape, -1, np.int64)
        cam[ok] = self.area_camera[leaf_area[ok]]
        return (ok & (cam < 0)).astype(np.uint8)


class VhSkyboxDesc(ctypes.Structure):
    """vh_skybox_desc (include/vradhip.h); the oracle's vc_skybox has the same layout."""
    _fields_ = [
        ("ncameras", ctypes.c_int32), ("camera_origin", ctypes.c_void_p),
        ("camera_world_to_sky", ctypes.c_void_p),
        ("nnodes", ctypes.c_int32), ("node_plane", ctypes.c_void_p), ("node_children", ctypes.c_void_p),
        ("nplanes", ctypes.c_int32), ("plane_normal", ctypes.c_void_p), ("plane_dist", ctypes.c_void_p),
        ("plane_type", ctypes.c_void_p),
        ("nleaves", ctypes.c_int32), ("leaf_recurses", ctypes.c_void_p),
    ]


@dataclass
class Skybox:
    """What the gather needs of a map's 3D skyboxes: the cameras (origin and
    world_to_sky), the BSP and each leaf's recursion flag."""

    camera_origin: np.ndarray
    camera_world_to_sky: np.ndarray
    tree: CompiledBspTree
    leaf_recurses: np.ndarray
    _keep: dict = field(default_factory=dict, repr=False)

    @staticmethod
    def from_map(entities: Sequence[dict], tree: CompiledBspTree, leaf_area: np.ndarray,
                 area_count: int) -> "Skybox":
        cams = SkyCameras.build(entities, tree, leaf_area, area_count)
        return Skybox(cams.origin, cams.world_to_sky, tree, cams.leaf_recurses(leaf_area))

    @property
    def ncameras(self) -> int:
        return int(np.asarray(self.camera_origin).reshape(-1, 3).shape[0])

    def pack(self) -> VhSkyboxDesc:
        """The descriptor (shared by the product and the oracle); the arrays
        it points into live as long as this object."""
        k = self._keep
        k["o"] = np.ascontiguousarray(self.camera_origin, np.float32).reshape(-1, 3)
        k["w"] = np.ascontiguousarray(self.camera_world_to_sky, np.float32).reshape(-1)
        k["r"] = np.ascontiguousarray(self.leaf_recurses, np.uint8).reshape(-1)
        t = self.tree

        def ptr(a):
            return a.ctypes.data if a.size else None

        return VhSkyboxDesc(len(k["o"]), ptr(k["o"]), ptr(k["w"]),
                            len(t.node_plane), ptr(t.node_plane), ptr(t.node_children),
                            len(t.plane_dist), ptr(t.plane_normal), ptr(t.plane_dist), ptr(t.plane_type),
                            len(k["r"]), ptr(k["r"]))
