"""The GPU library's FP64 soft-shadow tracer (btg_options.precision = 2,
beamtrace2007_amd/csrc/trace_fp64.h) compiled for the host with bounds checks
that trap (tests/native/fp64_host.cpp) and compared with the reference's
double build (oracle/_ref/libbtref_d.so, the reference compiled with
Real = double):

* at the double build's own receiver points (its double primary rays), the
  visible fraction equals beamLightVisibility's (render.cpp:279-303) to 1e-10
  and the exact irradiance to 1e-6 of the largest, while the scene's kd-tree is
  this library's float one and the busiest points are traced as cells of
  their light: other splits, so the slivers below the double build's minimum
  area differ (up to ~2e-11 of the light) and so do the pieces whose edge
  angles acos takes from cosines near 1 (~1e-7 of the irradiance);
* at the receiver points the FP64 mode's primary pass forms (the float
  camera's pixel ray summed in double, the hit on the triangle's plane in
  double), every pixel is within 1e-5 of the double build's image: what is
  left is the float camera against the double build's double camera.

No device is used."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_scene
from refpy import Ref, ref_available

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
pytestmark = [
    pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available"),
    pytest.mark.skipif(not ref_available("d"), reason="reference double build (oracle/_ref) not present"),
]

dp, fp, ip = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int32)
CASES = [("plant", 1.0, 24), ("blob", 2.0, 24), ("atrium", 4.0, 24), ("conference", 1.0, 24), ("room", 1.0, 24)]


@pytest.fixture(scope="module")
def fp64lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("fp64") / "libfp64_host.so"
    src = os.path.join(ROOT, "tests", "native", "fp64_host.cpp")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-fno-fast-math", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"),
                    "-Xarch_host", "-fsanitize=array-bounds,local-bounds",
                    "-Xarch_host", "-fsanitize-trap=array-bounds,local-bounds", src, "-o", str(out)],
                   check=True, capture_output=True, timeout=300)
    L = C.CDLL(str(out))
    for f in (L.fp64_soft, L.fp64_soft_tiny):
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_void_p, fp, C.POINTER(C.c_uint32), fp, fp, C.c_double, C.c_int, dp, ip, fp,
                      C.c_int, C.c_uint32, dp, dp, dp, C.POINTER(C.c_uint64), C.POINTER(C.c_int64)]
    return L


def leaf_refs(a, k) -> np.ndarray:
    """The library's 64-byte leaf references (LeafTri) in leaf order."""
    ti = np.asarray(k["triIndices"], np.int64)
    tris = np.asarray(a["tris"], np.int64)
    rec = np.zeros((len(ti), 16), np.uint32)
    rec[:, 0:9] = np.asarray(a["pos"], np.float32)[tris[ti]].reshape(len(ti), 9).view(np.uint32)
    rec[:, 9:12] = np.asarray(a["triN"], np.float32)[ti].view(np.uint32)
    rec[:, 12:15] = tris[ti].astype(np.uint32)
    rec[:, 15] = ti.astype(np.uint32)
    return np.ascontiguousarray(rec)


class Fp64Host:
    def __init__(self, L, name):
        self.L = L
        s, self.a, self.k, self.sp = load_scene(name)
        self.refs = leaf_refs(self.a, self.k)
        self.nodes = np.ascontiguousarray(self.k["nodes"])
        self.box = np.ascontiguousarray(self.k["nodeBox"], np.float32)
        self.par = np.ascontiguousarray(self.k["parent"], np.uint32)
        self.kb = np.ascontiguousarray(self.k["bounds"], np.float32)
        first = {}
        for i, t in enumerate(self.k["triIndices"]):
            first.setdefault(int(t), i)
        self.first_ref = first
        sb = np.asarray(self.a["bounds"], np.float32)
        # the library's guarded apex offset: float(|scene diagonal| x 1e-9)
        self.offset = float(np.float32(np.linalg.norm(sb[3:] - sb[:3])) * np.float32(1e-9))

    def shade(self, light, X, tri, nf, tiny=False, work_cap=65536):
        """(vis, E, counters, [cells sent to level 1, to level 2])."""
        n = len(tri)
        ri = np.array([self.first_ref[int(t)] for t in tri], np.int32)
        vis, E, Eu = np.zeros(n), np.zeros(n), np.zeros(n)
        cnt = np.zeros(16, np.uint64)
        levels = np.zeros(2, np.int64)
        X = np.ascontiguousarray(X, np.float64)
        nf = np.ascontiguousarray(nf, np.float32)
        light = np.ascontiguousarray(light, np.float32)
        f = self.L.fp64_soft_tiny if tiny else self.L.fp64_soft
        rc = f(self.refs.ctypes.data, self.nodes.ctypes.data, self.box.ctypes.data_as(fp),
                              self.par.ctypes.data_as(C.POINTER(C.c_uint32)), self.kb.ctypes.data_as(fp),
                              light.ctypes.data_as(fp), self.offset, 1, X.ctypes.data_as(dp), ri.ctypes.data_as(ip),
                              nf.ctypes.data_as(fp), n, int(work_cap), vis.ctypes.data_as(dp), E.ctypes.data_as(dp),
                              Eu.ctypes.data_as(dp), cnt.ctypes.data_as(C.POINTER(C.c_uint64)),
                              levels.ctypes.data_as(C.POINTER(C.c_int64)))
        assert rc == 0, f"a level-2 cell of point {-1 - rc} overflowed its pools"
        return vis, E, cnt, levels


@pytest.mark.parametrize("name,scale,side", CASES)
def test_fp64_tracer_equals_the_double_build_at_its_points(fp64lib, name, scale, side):
    from beamtrace2007_amd.workloads import scaled_light
    h = Fp64Host(fp64lib, name)
    light = scaled_light(h.sp["area_light"], scale)
    refd = Ref("d").load(h.a["pos"], h.a["tris"])
    try:
        pr = refd.primary_rays(h.sp["view"], side, side)
        tri = pr["tri"].ravel()
        hit = np.nonzero(tri >= 0)[0]
        X, N = pr["X"].reshape(-1, 3)[hit], pr["N"].reshape(-1, 3)[hit]
        vis, E, cnt, _ = h.shade(light, X, tri[hit], N.astype(np.float32))
        rv, rG = np.zeros(len(hit)), np.zeros(len(hit))
        for j in range(len(hit)):
            rv[j], rG[j], _ = refd.light_vis(X[j], N[j], int(tri[hit[j]]), light, h.offset)
    finally:
        refd.free()
    assert len(hit) > 0.2 * side * side
    assert np.sum((rv > 1e-3) & (rv < 1 - 1e-3)) > 0  # some penumbra in the sample
    assert cnt[2] > 0  # triangle tests ran
    assert np.abs(vis - rv).max() <= 1e-10
    assert (np.abs(E - rG) / max(np.abs(rG).max(), 1e-30)).max() <= 1e-6


def fp64_primary_points(h, cam14, W, H, tri, hit):
    """The FP64 mode's receiver points: the float camera's pixel ray summed in
    double, the hit on the plane of the triangle's float corners in double,
    the triangle normal turned to face the eye by the double direction."""
    cam = np.asarray(cam14, np.float32).astype(np.float64)
    eye, fwd, right, down, halfW, halfH = cam[0:3], cam[3:6], cam[6:9], cam[9:12], cam[12], cam[13]
    px, py = (hit % W).astype(np.float64), (hit // W).astype(np.float64)
    qx = (2 * (px + 0.5) / W - 1) * halfW
    qy = (2 * (py + 0.5) / H - 1) * halfH
    d = (fwd[None, :] + right[None, :] * qx[:, None]) + down[None, :] * qy[:, None]
    pos = np.asarray(h.a["pos"], np.float32).astype(np.float64)
    tris = np.asarray(h.a["tris"], np.int64)[tri]
    v0, v1, v2 = pos[tris[:, 0]], pos[tris[:, 1]], pos[tris[:, 2]]
    n = np.cross(v1 - v0, v2 - v0)
    t = np.sum(n * (v0 - eye), 1) / np.sum(n * d, 1)
    X = eye[None, :] + d * t[:, None]
    nf = np.asarray(h.a["triN"], np.float32)[tri].copy()
    away = np.sum(nf.astype(np.float64) * d, 1) > 0
    nf[away] = -nf[away]
    return X, nf


@pytest.mark.parametrize("name,scale,side", CASES)
def test_fp64_mode_points_stay_within_1e5_of_the_double_image(fp64lib, name, scale, side):
    from beamtrace2007_amd.host import camera
    from beamtrace2007_amd.workloads import scaled_light
    h = Fp64Host(fp64lib, name)
    light = scaled_light(h.sp["area_light"], scale)
    refd = Ref("d").load(h.a["pos"], h.a["tris"])
    try:
        img = refd.soft(h.sp["view"], side, side, light, exact=True, mailbox=False, trail=True, threads=8)
    finally:
        refd.free()
    tri = img["tri"].ravel()
    hit = np.nonzero(tri >= 0)[0]
    X, nf = fp64_primary_points(h, camera(h.sp["view"], side, side), side, side, tri[hit], hit)
    vis, E, _, _ = h.shade(light, X, tri[hit], nf)
    err = np.abs(vis.astype(np.float32).astype(np.float64) - img["vis"].ravel()[hit])
    assert err.max() <= 1e-5, f"{name}: max {err.max():.3e}"


def test_fp64_cells_add_up_to_the_whole_root(fp64lib):
    """Pools of 4 beams and a work cap of 256 send most lit points through
    the 4 x 4 and 16 x 16 cells of their light; the cells' visible fractions
    add up to the whole root's to 1e-12 (other splits, other rounding; the
    irradiance to 1e-6, acos's rounding on short edges) and stay within 1e-10
    of the double build."""
    from beamtrace2007_amd.workloads import scaled_light
    h = Fp64Host(fp64lib, "blob")
    light = scaled_light(h.sp["area_light"], 2.0)
    refd = Ref("d").load(h.a["pos"], h.a["tris"])
    try:
        pr = refd.primary_rays(h.sp["view"], 16, 16)
        tri = pr["tri"].ravel()
        hit = np.nonzero(tri >= 0)[0]
        X, N = pr["X"].reshape(-1, 3)[hit], pr["N"].reshape(-1, 3)[hit]
        rv = np.array([refd.light_vis(X[j], N[j], int(tri[hit[j]]), light, h.offset)[0] for j in range(len(hit))])
    finally:
        refd.free()
    nf = N.astype(np.float32)
    v0, e0, _, l0 = h.shade(light, X, tri[hit], nf)
    v1, e1, _, l1 = h.shade(light, X, tri[hit], nf, tiny=True, work_cap=256)
    assert l1[0] > l0[0] and l1[0] >= 10 and l1[1] >= 10
    assert np.abs(v0 - v1).max() <= 1e-12
    assert np.abs(e0 - e1).max() <= 1e-6 * max(np.abs(e0).max(), 1e-30)  # acos of cosines near 1
    assert np.abs(v1 - rv).max() <= 1e-10


@pytest.mark.parametrize("name", ["plant", "blob", "atrium", "conference", "room"])
def test_double_camera_equals_the_double_builds(name):
    """bth_camera_d forms Camera::make (render.cpp:14-25) in double exactly as
    the reference's double build does from the same view (bit for bit), the
    camera FP64 soft shadows shoot from; the float camera differs from it by
    its rounding."""
    from beamtrace2007_amd.host import camera, camera_d
    _, a, _, sp = load_scene(name)
    ref = Ref("d")
    for W, H in ((64, 48), (512, 512), (1280, 720)):
        d = ref.camera(sp["view"], W, H)
        assert np.array_equal(camera_d(sp["view"], W, H), d)
    f = camera(sp["view"], 512, 512).astype(np.float64)
    assert np.abs(f - ref.camera(sp["view"], 512, 512)).max() < 1e-6
