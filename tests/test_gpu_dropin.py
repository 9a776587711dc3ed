"""The drop-in boundary end to end: the reference's own bt_render, built with
-DBT_WITH_CUDA against this project's GpuRenderer (beamtrace2007_amd/shim/
gpu_render_hip.cpp over libbtg.so; oracle/Makefile target `apps`), renders on
the MI355X with `--device cuda` and must reproduce its own CPU path.

The binaries are built in the development container from the reference's
sources (test infrastructure, oracle/_ref); the test skips where they are
absent."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "oracle", "_ref", "bt_render_hip")

# A scene config in the reference's .cfg grammar (src/scene/scene.cpp:104-191):
# the reference's procedural plant, our own view and lights.
CFG = """name plant_dropin
mesh procedural plant
up 0 1 0
view 2.2 1.5 2.8  0 0.5 0  45
point_light 1.0 3.0 0.8
area_light 0.9 2.8 0.7  0.4 0 0  0 0 0.4
"""


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, 3)[::-1, :, 0]


@pytest.mark.skipif(not os.path.exists(BIN), reason="bt_render_hip not built (needs the reference sources)")
def test_reference_cli_renders_soft_shadows_on_the_gpu_exactly(require_gpu, tmp_path):
    cfg = tmp_path / "plant.cfg"
    cfg.write_text(CFG)
    outs = {}
    for dev in ("cuda", "cpu"):
        prefix = tmp_path / dev
        cmd = [BIN, "--config", str(cfg), "--mode", "softshadow", "--res", "80x64", "--exact", "--no-mailbox",
               "--trail", "--threads", "4", "--device", dev, "--out", str(prefix), "--pfm", "--csv",
               str(tmp_path / f"{dev}.csv"), "--quiet"]
        print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, timeout=120)
        outs[dev] = read_pfm(str(prefix) + "_vis.pfm")
    assert outs["cuda"].shape == (64, 80)
    assert np.array_equal(outs["cuda"], outs["cpu"])
    row = (tmp_path / "cuda.csv").read_text().splitlines()[1].split(",")
    assert row[2] == "beam_cuda"
