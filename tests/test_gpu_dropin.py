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


def read_png(path):
    """8-bit RGB / RGBA PNG (the reference's writer, src/util/image.cpp) as an
    (H, W, C) uint8 array: zlib stream, per-row filters 0-4."""
    import struct
    import zlib
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[0], hdr[1], hdr[2], hdr[3]
    assert depth == 8 and ctype in (2, 6)
    c = 3 if ctype == 2 else 4
    raw = zlib.decompress(idat)
    stride = w * c
    out = np.zeros((h, stride), np.int32)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        row = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for x in range(stride):
            a = cur[x - c] if x >= c else 0
            b = prev[x]
            d = prev[x - c] if x >= c else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - d), abs(a - d), abs(a + b - 2 * d)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else d)
            cur[x] = (row[x] + p) & 255
        out[y] = cur
        prev = cur
    return out.reshape(h, w, c).astype(np.uint8)


@pytest.mark.skipif(not os.path.exists(BIN), reason="bt_render_hip not built (needs the reference sources)")
@pytest.mark.parametrize("mode", ["primary", "pointshadow"])
def test_reference_cli_renders_primary_and_point_shadows_on_the_gpu(require_gpu, tmp_path, mode):
    """bt_render --mode primary / pointshadow (apps/bt_render.cpp:217-247):
    with --device cuda the reference's own CLI calls GpuRenderer::primary and
    ::pointShadows, which the shim runs on the MI355X (16 x 16 tile roots, as
    the reference's GPU path, cuda/gpu_render.h:41), then rasterizes and
    shades on the host. --device cpu traces the whole image as one root
    (src/render/render.cpp:150-158), whose fuzzy epsilon scales with the root's
    extent, so pixels whose centre lies within that epsilon of an edge may
    differ; everything else must be equal."""
    cfg = tmp_path / "plant.cfg"
    cfg.write_text(CFG)
    imgs, rows = {}, {}
    for dev in ("cuda", "cpu"):
        prefix = tmp_path / dev
        cmd = [BIN, "--config", str(cfg), "--mode", mode, "--res", "128x96", "--aa", "1", "--no-mailbox", "--trail",
               "--threads", "4", "--device", dev, "--out", str(prefix), "--csv", str(tmp_path / f"{dev}.csv"),
               "--quiet"]
        print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, timeout=120)
        imgs[dev] = read_png(str(prefix) + ".png")
        head, row = (tmp_path / f"{dev}.csv").read_text().splitlines()[:2]
        rows[dev] = dict(zip(head.split(","), row.split(",")))
    assert imgs["cuda"].shape[:2] == (96, 128)
    same = np.all(imgs["cuda"] == imgs["cpu"], axis=2)
    assert same.mean() >= 0.995, f"{(~same).sum()} pixels differ"
    assert rows["cuda"]["method"] == "beam_cuda" and int(rows["cuda"]["hit_beams"]) > 0
    assert rows["cuda"]["visible_tris"] == rows["cpu"]["visible_tris"] or \
        abs(int(rows["cuda"]["visible_tris"]) - int(rows["cpu"]["visible_tris"])) <= 2
