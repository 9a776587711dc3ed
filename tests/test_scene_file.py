"""The resident scene format (.btgs, beamtrace2007_amd/host/scene_file.h):
round trips keep every array bit for bit, the kd-tree needs no rebuild, and
damaged files are refused."""
import os

import numpy as np
import pytest

from conftest import load_scene


def arrays_equal(a, b):
    return all(np.array_equal(a[k], b[k]) for k in a)


@pytest.mark.parametrize("name", ["room", "plant", "sphere"])
def test_round_trip_keeps_scene_tree_and_spec(tmp_path, name):
    from beamtrace2007_amd.host import Scene
    s, a, k, sp = load_scene(name)
    path = str(tmp_path / f"{name}.btgs")
    s.save(path)
    t = Scene.load(path)
    assert arrays_equal(t.arrays(), a)
    kt = t.kd_arrays()
    for key in ("triIndices", "parent", "nodeBox", "bounds"):
        assert np.array_equal(kt[key], k[key]), key
    assert np.array_equal(kt["nodes"], k["nodes"]) and kt["depth"] == k["depth"]
    ts = t.spec()
    for key in ("view", "point_light", "area_light"):
        assert np.array_equal(ts[key], sp[key])
    assert (ts["has_point"], ts["has_area"], ts["cull"]) == (sp["has_point"], sp["has_area"], sp["cull"])
    # saving the loaded scene writes the same bytes
    path2 = str(tmp_path / "again.btgs")
    t.save(path2)
    assert open(path, "rb").read() == open(path2, "rb").read()


def test_scene_without_tree_round_trips(tmp_path):
    from beamtrace2007_amd.host import Scene
    s = Scene.generate("analytic")
    path = str(tmp_path / "a.btgs")
    s.save(path)
    t = Scene.load(path)
    assert arrays_equal(t.arrays(), s.arrays())
    with pytest.raises(RuntimeError):
        t.kd_arrays()


def test_damaged_files_are_refused(tmp_path):
    from beamtrace2007_amd.host import Scene
    s, a, k, sp = load_scene("room")
    path = str(tmp_path / "r.btgs")
    s.save(path)
    raw = bytearray(open(path, "rb").read())
    cases = {
        "flip": lambda b: b.__setitem__(300, b[300] ^ 1),      # payload bit: hash mismatch
        "magic": lambda b: b.__setitem__(0, ord("X")),
        "count": lambda b: b.__setitem__(16, b[16] ^ 1),       # vertex count: size mismatch
        "truncated": lambda b: b.__delitem__(slice(len(b) - 8, len(b))),
        "trailing": lambda b: b.extend(b"\0" * 8),
    }
    for name, damage in cases.items():
        b = bytearray(raw)
        damage(b)
        bad = str(tmp_path / f"{name}.btgs")
        open(bad, "wb").write(bytes(b))
        with pytest.raises(RuntimeError):
            Scene.load(bad)
    with pytest.raises(RuntimeError):
        Scene.load(str(tmp_path / "missing.btgs"))


def test_header_layout(tmp_path):
    """Magic, version and counts sit where the format says."""
    s, a, k, sp = load_scene("room")
    path = str(tmp_path / "r.btgs")
    s.save(path)
    raw = open(path, "rb").read()
    assert raw[:8] == b"BTGSCENE"
    assert np.frombuffer(raw[8:12], "<u4")[0] == 1
    counts = np.frombuffer(raw[16:48], "<i8")
    assert list(counts) == [len(a["pos"]), len(a["tris"]), len(k["nodes"]), len(k["triIndices"])]
    payload = np.frombuffer(raw[192:200], "<u8")[0]
    assert len(raw) == 256 + payload
    assert np.array_equal(np.frombuffer(raw[256:256 + 12 * len(a["pos"])], "<f4").reshape(-1, 3), a["pos"])
