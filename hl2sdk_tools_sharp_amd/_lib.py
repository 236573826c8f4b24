"""ctypes binding of libvradhip.so (the C ABI declared in include/vradhip.h).

The library is the only compute path: if it is missing this module raises,
it never falls back to anything on the CPU.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvradhip.so")

VH_OK = 0
VH_DECLINED = 1
VH_EINVAL = 2
VH_EDEVICE = 3
VH_ETIMEOUT = 4
VH_ECANCELED = 5
VH_ENOTSUP = 6
VH_ESTATE = 7

STATUS_NAMES = {
    VH_OK: "VH_OK",
    VH_DECLINED: "VH_DECLINED",
    VH_EINVAL: "VH_EINVAL",
    VH_EDEVICE: "VH_EDEVICE",
    VH_ETIMEOUT: "VH_ETIMEOUT",
    VH_ECANCELED: "VH_ECANCELED",
    VH_ENOTSUP: "VH_ENOTSUP",
    VH_ESTATE: "VH_ESTATE",
}


VH_BUILD_AUTO, VH_BUILD_HOST, VH_BUILD_DEVICE = 0, 1, 2
VH_BUILD_DEVICE_MIN = 65536
VH_POINT_LIGHTS_AUTO, VH_POINT_LIGHTS_BVH, VH_POINT_LIGHTS_LIGHT_BUFFER = 0, 1, 2


class VhOpenOpts(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("self_test", ctypes.c_int32),
        ("fail_stage", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class VhStats(ctypes.Structure):
    _fields_ = [
        ("rays_visibility", ctypes.c_int64),
        ("rays_closest", ctypes.c_int64),
        ("calls", ctypes.c_int64),
        ("rays_gather", ctypes.c_int64),
        ("reserved", ctypes.c_int64 * 4),
    ]


class VhPatchDesc(ctypes.Structure):
    _fields_ = [
        ("npatch", ctypes.c_int32), ("nface", ctypes.c_int32), ("ncluster", ctypes.c_int32),
        ("nreceivers", ctypes.c_int32),
        ("origin", ctypes.c_void_p), ("normal", ctypes.c_void_p), ("plane_dist", ctypes.c_void_p),
        ("area", ctypes.c_void_p), ("child1", ctypes.c_void_p), ("child2", ctypes.c_void_p),
        ("next_parent", ctypes.c_void_p), ("face", ctypes.c_void_p), ("cluster", ctypes.c_void_p),
        ("winding", ctypes.c_void_p), ("reflectivity", ctypes.c_void_p),
        ("face_sky", ctypes.c_void_p), ("face_parent", ctypes.c_void_p),
        ("row_faces_ptr", ctypes.c_void_p), ("row_faces", ctypes.c_void_p),
        ("receivers", ctypes.c_void_p), ("bumped", ctypes.c_void_p), ("bump_normals", ctypes.c_void_p),
    ]


class VhBounceOpts(ctypes.Structure):
    _fields_ = [
        ("weights", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("tol", ctypes.c_double),
    ]


class VradHipError(RuntimeError):
    """A libvradhip call failed; ``status`` is the VH_* code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


_lock = threading.Lock()
_lib = None

P = ctypes.c_void_p
I32, I64, F32, F64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double

# name -> (restype, argtypes); the single table of what this module binds.
class VhMlpOpts(ctypes.Structure):
    _fields_ = [("length", ctypes.c_float), ("floor", ctypes.c_float), ("gain", ctypes.c_float),
                ("scale", ctypes.c_float)]


class VhMlpTrainOpts(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int32), ("batch", ctypes.c_int32), ("seed", ctypes.c_int32),
        ("use_graph", ctypes.c_int32), ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("fused", ctypes.c_int32),
    ]


SIGNATURES = {
    "vh_open": (ctypes.c_int, [ctypes.POINTER(VhOpenOpts), ctypes.POINTER(P)]),
    "vh_close": (None, [P]),
    "vh_last_error": (ctypes.c_char_p, []),
    "vh_tracer_identity": (ctypes.c_char_p, [P]),
    "vh_supports": (ctypes.c_int, [P, F32, I32, I32]),
    "vh_scene_load": (ctypes.c_int, [P, P, P, P, I32]),
    "vh_scene_info": (ctypes.c_int, [P, P, P, P]),
    "vh_scene_load_with": (ctypes.c_int, [P, P, P, P, I32, I32]),
    "vh_scene_read": (ctypes.c_int, [P, P, ctypes.c_int64, P, ctypes.c_int64]),
    "vh_trace_visibility": (ctypes.c_int, [P, P, I32, F32, I64, F32, I32, I32, P, I64]),
    "vh_trace_closest": (ctypes.c_int, [P, P, I32, F32, I64, F32, I32, P, P, I64]),
    "vh_trace_point_lights": (ctypes.c_int, [P, P, I64, P, I32, F32, P, I64]),
    "vh_trace_visibility_dev": (ctypes.c_int, [P, P, I32, F32, I64, F32, I32, I32, P, P]),
    "vh_trace_closest_dev": (ctypes.c_int, [P, P, I32, F32, I64, F32, I32, P, P, P]),
    "vh_trace_point_lights_dev": (ctypes.c_int, [P, P, I64, P, I32, F32, P, P]),
    "vh_trace_point_lights_noalloc": (ctypes.c_int, [P, P, I64, P, I32, F32, P, P]),
    "vh_sync": (ctypes.c_int, [P, P]),
    "vh_set_timeout": (ctypes.c_int, [P, I64]),
    "vh_set_point_light_method": (ctypes.c_int, [P, I32, I32]),
    "vh_cancel": (ctypes.c_int, [P]),
    "vh_cancel_clear": (ctypes.c_int, [P]),
    "vh_get_stats": (ctypes.c_int, [P, ctypes.POINTER(VhStats)]),
    "vh_patches_load": (ctypes.c_int, [P, ctypes.POINTER(VhPatchDesc)]),
    "vh_transfers_build": (ctypes.c_int, [P, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]),
    "vh_transfers_read": (ctypes.c_int, [P, P, P, P, I64]),
    "vh_transfers_export": (ctypes.c_int, [P, P, P, I64]),
    "vh_transfers_load": (ctypes.c_int, [P, P, P, P]),
    "vh_bounce": (ctypes.c_int, [P, ctypes.POINTER(VhBounceOpts), P, P, P, P, P, ctypes.POINTER(I32)]),
    "vh_gather_load": (ctypes.c_int, [P, P]),
    "vh_gather_skybox": (ctypes.c_int, [P, P]),
    "vh_gather_run": (ctypes.c_int, [P, I32, I32]),
    "vh_gather_run_shaped": (ctypes.c_int, [P, I32, I32, I32]),
    "vh_gather_read": (ctypes.c_int, [P, P, P, P]),
    "vh_gather_styles": (ctypes.c_int, [P, P, ctypes.POINTER(I32), ctypes.POINTER(I64)]),
    "vh_gather_read_slot": (ctypes.c_int, [P, I32, P, P]),
    "vh_gather_supersample": (ctypes.c_int, [P, P, P]),
    "vh_encode_rgbe": (ctypes.c_int, [P, P, I64, I32, P]),
    "vh_lightmap_encode": (ctypes.c_int, [P, I32, P, I64]),
    "vh_patch_lights": (ctypes.c_int, [P, I32, P, P]),
    "vh_bounce_bump_totals": (ctypes.c_int, [P, P]),
    "vh_final_load": (ctypes.c_int, [P, P]),
    "vh_final_layout": (ctypes.c_int, [P, P, ctypes.POINTER(I64)]),
    "vh_final_light": (ctypes.c_int, [P, P, P, P, I64]),
    "vh_ambient_load": (ctypes.c_int, [P, P]),
    "vh_ambient_cubes": (ctypes.c_int, [P, P, I64, P]),
    "vh_ambient_directions": (ctypes.c_int, [P]),
    "vh_mlp_init": (ctypes.c_int, [P, P, P, ctypes.POINTER(VhMlpOpts)]),
    "vh_mlp_params": (ctypes.c_int, [P, P]),
    "vh_mlp_forward": (ctypes.c_int, [P, P, I64, P]),
    "vh_mlp_grad": (ctypes.c_int, [P, P, P, I64, P, ctypes.POINTER(F32)]),
    "vh_mlp_adam": (ctypes.c_int, [P, P, F32, F32, F32, F32]),
    "vh_mlp_features": (ctypes.c_int, [P, P, P]),
    "vh_mlp_dataset": (ctypes.c_int, [P, I32, ctypes.POINTER(I64)]),
    "vh_mlp_train": (ctypes.c_int, [P, ctypes.POINTER(VhMlpTrainOpts), P]),
    "vh_mlp_predict": (ctypes.c_int, [P, P, P]),
    "vh_timing_enable": (ctypes.c_int, [P, I32]),
    "vh_timing_read": (ctypes.c_int, [P, I32, ctypes.POINTER(F64), ctypes.POINTER(I64)]),
    "vh_count_enable": (ctypes.c_int, [P, I32]),
    "vh_count_read": (ctypes.c_int, [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
    "vh_count_read_slots": (ctypes.c_int, [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]),
}

VH_KERNEL_POINT_LIGHTS = 0
VH_KERNEL_VISIBILITY = 1
VH_KERNEL_CLOSEST = 2
VH_KERNEL_GATHER = 3
VH_KERNEL_TRANSFERS = 4
VH_KERNEL_FACELIGHT = 5
VH_KERNEL_MLP = 6
VH_KERNEL_FINAL = 7
VH_KERNEL_AMBIENT = 8
VH_KERNEL_SUPERSAMPLE = 9
VH_KERNEL_LIGHT_BUFFER = 10

VH_WEIGHTS_F32 = 0
VH_WEIGHTS_F16 = 1
VH_BOUNCE_PROGRESSIVE = 0
VH_BOUNCE_FROM_START = 1
VH_BOUNCE_FROM_PREDICTION = 2
VH_MLP_PARAMS = 36739
VH_MLP_INPUTS = 25
VH_MLP_DATA_CLEAR = 0
VH_MLP_DATA_ADD = 1


def lib():
    """Loads libvradhip.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m hl2sdk_tools_sharp_amd.build` "
                "(there is no CPU fallback)"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        return L


def last_error() -> str:
    s = lib().vh_last_error()
    return s.decode("utf-8", "replace") if s else ""


def check(status: int) -> None:
    if status != VH_OK:
        raise VradHipError(status, last_error())
