"""Shared fixtures. GPU tests are marked `gpu`; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


_scenes = {}


def load_scene(name: str, seed: int = 2007):
    """(Scene with kd-tree, arrays, kd arrays, spec) — cached per session."""
    key = (name, seed)
    if key not in _scenes:
        from beamtrace2007_amd.host import Scene
        s = Scene.generate(name, seed)
        s.build_kd(0.4)
        _scenes[key] = (s, s.arrays(), s.kd_arrays(), s.spec())
    return _scenes[key]


@pytest.fixture(scope="session")
def scene_loader():
    return load_scene


def gpu_available() -> bool:
    try:
        from beamtrace2007_amd import gpu
        return gpu.device_info() is not None
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    from beamtrace2007_amd import gpu
    info = gpu.device_info()  # raises if libbtg.so is missing: the HIP path must be the one that runs
    if info is None:
        pytest.fail("no HIP device visible to libbtg.so")
    return info
