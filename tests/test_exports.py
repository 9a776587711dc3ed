"""The C-ABI libraries load and export every entry point their headers
declare (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(bt[gh]_[a-z_]+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("btg.h", "libbtg.so"), ("bth.h", "libbt_host.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "beamtrace2007_amd", "lib", lib)
    assert os.path.exists(path), f"{lib} not built"
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_python_binding_lists_the_header():
    from beamtrace2007_amd import gpu
    assert sorted(gpu.EXPORTS) == declared("btg.h")


def test_device_query_without_device_is_graceful():
    from beamtrace2007_amd import gpu
    info = gpu.device_info()  # None here (no GPU in the build container), a dict on the GPU box
    assert info is None or info["wavefront"] == 64


def test_create_rejects_bad_scene_without_touching_the_device():
    """btg_create validates the arrays on the host before any device work."""
    from beamtrace2007_amd import gpu
    L = gpu.lib()
    d = gpu.btg_scene_desc()
    h = ctypes.c_void_p()
    assert L.btg_create(ctypes.byref(h), ctypes.byref(d)) < 0
    assert b"incomplete" in L.btg_last_error(None)


def test_entry_points_reject_null_arguments_without_a_device():
    """Argument checks come before any device work: null contexts, servers
    and outputs are errors (negative codes), never crashes."""
    from beamtrace2007_amd import gpu
    L = gpu.lib()
    q = gpu.btg_request()
    q.cam = gpu.make_camera([0, 0, 0, 0, 0, 1, 1, 0, 0, 0, -1, 0, 0.5, 0.5], 8, 8)
    b = ctypes.c_int64(0)
    assert L.btg_footprint(None, ctypes.byref(q), ctypes.byref(b)) < 0
    t = ctypes.c_uint64(0)
    assert L.btg_server_submit(None, ctypes.byref(q), ctypes.byref(t)) < 0
    st = gpu.btg_request_status()
    assert L.btg_server_wait(None, 1, 0.0, ctypes.byref(st)) < 0
    assert L.btg_server_error(None) == b""
    L.btg_server_destroy(None)
    assert L.btg_scene_bytes(None, 0) == -1


_keep = []


def _chain_tree(inner_levels: int):
    """A kd-tree whose every inner node has a leaf on its left and the next
    inner node on its right (depth-first order), so its depth is
    inner_levels; all leaves are empty."""
    import numpy as np
    from beamtrace2007_amd.host import Scene
    n = 2 * inner_levels + 1
    nodes = np.zeros(n, dtype=[("split", "<f4"), ("data", "<u4"), ("offset", "<u4")])
    parent = np.full(n, 0xFFFFFFFF, np.uint32)
    for lev in range(inner_levels):
        i = 2 * lev
        nodes[i] = (0.0, ((i + 2) << 2) | (lev % 3), 0)  # right child i + 2, axis lev % 3
        nodes[i + 1] = (0.0, 3, 0)                       # empty leaf
        parent[i + 1] = parent[i + 2] = i
    nodes[n - 1] = (0.0, 3, 0)
    s = Scene.generate("analytic")
    _keep.append(s)  # the arrays view the scene's memory
    return s.arrays(), nodes, parent


def _create_with(arrays, nodes, parent):
    import numpy as np
    from beamtrace2007_amd import gpu
    L = gpu.lib()
    d = gpu.btg_scene_desc()
    box = np.zeros((len(nodes), 6), np.float32)
    idx = np.zeros(1, np.uint32)
    d.pos, d.numVerts = arrays["pos"].ctypes.data, len(arrays["pos"])
    d.tri3, d.triN, d.numTris = arrays["tris"].ctypes.data, arrays["triN"].ctypes.data, len(arrays["tris"])
    d.nodes, d.numNodes = nodes.ctypes.data, len(nodes)
    d.triIndices, d.numRefs = idx.ctypes.data, 0
    d.parent, d.nodeBox = parent.ctypes.data, box.ctypes.data
    h = ctypes.c_void_p()
    rc = L.btg_create(ctypes.byref(h), ctypes.byref(d))
    msg = L.btg_last_error(None)
    if rc == 0:
        L.btg_destroy(h)
    return rc, msg


def test_create_rejects_kd_trees_deeper_than_the_restart_trail():
    """The restart trail keeps one bit per inner level in 64 bits: a tree with
    64 inner levels is refused before any device work, one with 63 passes
    the host checks (and then needs a device)."""
    from beamtrace2007_amd import gpu
    arrays, nodes, parent = _chain_tree(65)
    rc, msg = _create_with(arrays, nodes, parent)
    assert rc < 0 and b"deeper than 63" in msg
    arrays, nodes, parent = _chain_tree(64)
    rc, msg = _create_with(arrays, nodes, parent)
    assert rc == 0 or b"deeper" not in msg
    if gpu.device_info() is None:
        assert rc < 0 and b"device" in msg


def test_create_rejects_parent_links_that_do_not_match_the_children():
    arrays, nodes, parent = _chain_tree(5)
    parent[4] = 0  # node 4 is node 2's right child
    rc, msg = _create_with(arrays, nodes, parent)
    assert rc < 0 and b"parent links" in msg
    arrays, nodes, parent = _chain_tree(5)
    nodes[2]["data"] = (1 << 2) | 0  # node 2's right child points back at node 1 (not after its left child)
    rc, msg = _create_with(arrays, nodes, parent)
    assert rc < 0 and b"child index" in msg


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every public struct the Python binding mirrors has the header's size and
    field offsets (compiled with the host C compiler)."""
    import shutil
    import subprocess
    from beamtrace2007_amd import gpu
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = [v for k, v in vars(gpu).items()
               if k.startswith("btg_") and isinstance(v, type) and issubclass(v, ctypes.Structure)]
    assert len(structs) >= 10
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "btg.h"', "int main(void) {"]
    for st in structs:
        lines.append(f'  printf("{st.__name__} size %zu\\n", sizeof({st.__name__}));')
        for f in st._fields_:
            lines.append(f'  printf("{st.__name__} {f[0]} %zu\\n", offsetof({st.__name__}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for st in structs:
        assert got[(st.__name__, "size")] == ctypes.sizeof(st), st.__name__
        for f in st._fields_:
            assert got[(st.__name__, f[0])] == getattr(st, f[0]).offset, (st.__name__, f[0])
