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


VIS_TOL = 1e-5   # absolute, visible fraction
E_TOL = 1e-5     # relative, exact Lambert integral


def check_soft(g, o, vis_tol: float = VIS_TOL, e_tol: float = E_TOL):
    """GPU soft shadows against the oracle: primary hit ids and the
    unoccluded irradiance bit-exact, visible fraction within vis_tol absolute
    (in practice bit-exact: piece areas are summed exactly), exact irradiance
    within e_tol relative (the double acos of device libm against glibc)."""
    assert np.array_equal(g["tri"], o["tri"])
    assert np.array_equal(g["Eu"], o["Eu"])
    dv = np.abs(g["vis"].astype(np.float64) - o["vis"])
    assert dv.max() <= vis_tol, f"max |dvis| {dv.max()}"
    assert np.mean(g["vis"] == o["vis"]) > 0.9999
    rel = np.abs(g["E"].astype(np.float64) - o["E"]) / np.maximum(np.abs(o["E"]), 1e-30)
    assert np.all((rel <= e_tol) | (np.abs(o["E"]) < 1e-30)), f"max rel dE {rel.max()}"


def canonical_equal(gpu_polys, oracle_polys) -> bool:
    """Polygon lists equal bit for bit in canonical order."""
    from beamtrace2007_amd.gpu import polys_to_dict
    from refpy import canon
    return bool(np.array_equal(canon(polys_to_dict(gpu_polys)), canon(oracle_polys)))
