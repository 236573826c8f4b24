"""Shared test set-up.

Markers: ``gpu`` -- needs an MI355X and libvradhip.so; everything else runs on
the CPU (oracle checks, host logic, the C ABI's exports and decline paths).
"""
from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libvradhip.so")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def vh_lib():
    """libvradhip.so must exist (build it with __graft_entry__.build()); no fallback."""
    from hl2sdk_tools_sharp_amd import _lib

    return _lib.lib()


@pytest.fixture(scope="session")
def kd_scenes():
    import kd_fixtures

    return {name: kd_fixtures.load(name) for name in ("kd-scene", "kd-boxes")}
