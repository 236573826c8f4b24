"""The C ABI without a GPU: the library loads, exports every symbol the
header declares, and declines cleanly when no device can be used."""
from __future__ import annotations

import ctypes
import glob
import os
import re

import pytest

from hl2sdk_tools_sharp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(vh_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_seam():
    names = declared_functions()
    for must in ("vh_open", "vh_close", "vh_trace_visibility", "vh_trace_closest", "vh_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(vh_lib):
    missing = [n for n in declared_functions() if not hasattr(vh_lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_fault_hook_declines_before_touching_the_device(vh_lib):
    opts = _lib.VhOpenOpts(-1, 1, 1, 0)
    h = ctypes.c_void_p()
    st = vh_lib.vh_open(ctypes.byref(opts), ctypes.byref(h))
    assert st == _lib.VH_DECLINED
    assert h.value is None
    assert "stage 1" in _lib.last_error()


def _has_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present: the no-device decline cannot happen here")
def test_factory_declines_without_a_device(vh_lib):
    """A factory that cannot open a device declines with a reason, never raises
    (Rad/GpuTracing.cs:43-49)."""
    import numpy as np

    from hl2sdk_tools_sharp_amd import HipTracerFactory, ShadowCasterSet

    casters = ShadowCasterSet(np.zeros((1, 9), np.float32), np.zeros(1, np.int32))
    offer = HipTracerFactory().try_create(casters)
    assert offer.tracer is None
    assert offer.decline_reason


def test_null_arguments_are_rejected(vh_lib):
    assert vh_lib.vh_open(None, None) == _lib.VH_EINVAL
    assert vh_lib.vh_trace_visibility(None, None, 7, 1.0, 0, 0.0, -1, 0, None, 0) == _lib.VH_EINVAL
    assert "NULL" in _lib.last_error()
    vh_lib.vh_close(None)  # ignored


def test_last_error_is_per_thread(vh_lib):
    import threading

    vh_lib.vh_trace_closest(None, None, 7, 1.0, 0, 0.0, -1, None, None, 0)
    seen = {}

    def other():
        seen["msg"] = _lib.last_error()

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen["msg"] == ""
    assert _lib.last_error() != ""


def test_deadline_and_cancel_entry_points_reject_a_null_context(vh_lib):
    assert vh_lib.vh_set_timeout(None, 1000) == 2
    assert vh_lib.vh_cancel(None) == 2
    assert vh_lib.vh_cancel_clear(None) == 2
