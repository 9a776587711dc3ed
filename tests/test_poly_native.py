"""The GPU library's polygon primitives (beamtrace2007_amd/csrc/beam_dev.h:
area, sidesOf, splitPoly, cleanPoly, splitFive, reverse, extent) compiled for
the host and compared bit for bit with the oracle's restatement of
src/beam/beam_geom.h and beam_core.h on adversarial random polygons
(tests/native/poly_parity.cpp). No device is used."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_polygon_primitives_match_the_oracle_bitwise(tmp_path):
    exe = tmp_path / "poly_parity"
    src = os.path.join(ROOT, "tests", "native", "poly_parity.cpp")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-fno-fast-math", "-Wno-unused-function", src, "-o", str(exe)], check=True,
                   capture_output=True, timeout=300)
    out = subprocess.run([str(exe), "200000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok 200000")
