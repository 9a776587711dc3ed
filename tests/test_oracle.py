"""Pins the oracle (oracle/bt_oracle.cpp, the CPU restatement the GPU path is
checked against):

* bit for bit against the reference's own float build — golden fixtures
  written from oracle/_ref (tests/golden/make_golden.py), and the live build
  when present;
* against closed forms that do not come from the reference: soft shadows of a
  square occluder under a square light, the Lambert integral of an unoccluded
  light against midpoint quadrature, and area conservation of primary beams.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import load_scene
from refpy import Oracle, Ref, canon, ref_available

SCENES = ["analytic", "room", "random_tris", "sphere", "plant"]
STAT_KEYS = ["kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "rootBeams", "presplitBeams",
             "fiveSplits", "climbs"]


def golden(name):
    return np.load(os.path.join(os.path.dirname(__file__), "golden", f"{name}.npz"))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


_oracles = {}


def oracle(name):
    if name not in _oracles:
        s, a, k, sp = load_scene(name)
        _oracles[name] = Oracle(a, k)
    return _oracles[name]


@pytest.mark.parametrize("name", SCENES)
def test_soft_shadows_match_reference_bitwise(name):
    g = golden(name)
    o = oracle(name).soft(g["cam48"], 48, 48, g["light9"], exact=True, threads=4)
    assert np.array_equal(o["tri"], g["soft_tri"])
    assert np.array_equal(o["vis"], g["soft_vis"])
    assert np.array_equal(o["E"], g["soft_E"])
    assert np.array_equal(o["Eu"], g["soft_Eu"])
    assert np.array_equal(np.array([o["stats"][k] for k in STAT_KEYS]), g["soft_stats"])
    assert o["stats"]["droppedArea"] == pytest.approx(float(g["soft_dropped"]), rel=1e-12, abs=1e-30)


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("tile", [0, 8])
def test_primary_matches_reference_bitwise(name, tile):
    g = golden(name)
    p = oracle(name).primary(g["cam64"], 64, 64, tile=tile)
    c = canon(p)
    assert len(c) == int(g[f"prim{tile}_count"])
    assert digest(c) == str(g[f"prim{tile}_digest"])
    assert np.array_equal(np.array([p["stats"][k] for k in STAT_KEYS]), g[f"prim{tile}_stats"])


@pytest.mark.parametrize("name", SCENES)
def test_point_shadows_match_reference_bitwise(name):
    g = golden(name)
    o = oracle(name)
    prim = o.primary(g["cam64"], 64, 64, tile=8)
    pt = o.point(g["cam64"], 64, 64, prim, g["point3"])
    c = canon(pt)
    assert len(c) == int(g["point_count"])
    assert digest(c) == str(g["point_digest"])
    assert np.array_equal(np.array([pt["stats"][k] for k in STAT_KEYS]), g["point_stats"])


@pytest.mark.skipif(not ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", ["room", "plant"])
def test_oracle_matches_live_reference_at_other_sizes(name):
    """A second resolution and aspect ratio, straight against the reference build."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 40, 27
    ref = Ref("f").load(a["pos"], a["tris"])
    r = ref.soft(sp["view"], W, H, sp["area_light"], exact=True, mailbox=False, trail=True)
    o = oracle(name).soft(camera(sp["view"], W, H), W, H, sp["area_light"], exact=True)
    for key in ("tri", "vis", "E", "Eu"):
        assert np.array_equal(r[key], o[key]), key
    rp = ref.primary(sp["view"], W, H, tile=16)
    op = oracle(name).primary(camera(sp["view"], W, H), W, H, tile=16)
    assert np.array_equal(canon(rp), canon(op))


def analytic_visibility(px, pz, H, h, occ):
    """Visible fraction of the unit square light at height H over floor point
    (px, pz) with a centred square occluder of side occ at height h: the
    occluder's shadow of the light, projected through the point onto the light
    plane, is a square of half-size occ/2 * H/h around -(px, pz) (H/h - 1)."""
    k = H / h
    cx, cz, half = px * (1 - k), pz * (1 - k), occ * k / 2
    ox = max(0.0, min(0.5, cx + half) - max(-0.5, cx - half))
    oz = max(0.0, min(0.5, cz + half) - max(-0.5, cz - half))
    return 1 - ox * oz


def test_soft_shadow_visibility_matches_closed_form():
    s, a, k, sp = load_scene("analytic")
    o = oracle("analytic")
    floor = int(np.where(np.all(a["pos"][a["tris"]][:, :, 1] == 0, axis=1))[0][0])
    rng = np.random.default_rng(17)
    offset = 1e-4
    light = np.array([0, 3, 0, 1, 0, 0, 0, 0, 1], np.float32)
    worst = 0.0
    for _ in range(300):
        px, pz = rng.uniform(-1.5, 1.5, 2)
        x = np.array([px, 0, pz], np.float32)
        V, G, st = o.light_vis(x, np.array([0, 1, 0], np.float32), floor, light, offset)
        ref = analytic_visibility(float(x[0]), float(x[2]), 3 - offset, 1 - offset, 1.0)
        worst = max(worst, abs(V - ref))
    assert worst < 2e-3  # float build tolerance of the reference's own test (tests/test_tracing.cpp:281)


def test_exact_integral_of_unoccluded_light_matches_quadrature():
    """polygonFormFactorG of the visible light (here: all of it) against a
    midpoint rule for the integral of cos(ti) cos(tl) / r^2 over the quad."""
    s, a, k, sp = load_scene("analytic")
    o = oracle("analytic")
    x = np.array([1.4, 0, 1.4], np.float32)
    light = np.array([0, 3, 0, 1, 0, 0, 0, 0, 1], np.float32)
    V, G, st = o.light_vis(x, np.array([0, 1, 0], np.float32), -1, light, 1e-4)
    assert V == 1.0
    n = 400
    u = (np.arange(n) + 0.5) / n - 0.5
    P = np.stack(np.meshgrid(u, u, indexing="ij"), -1).reshape(-1, 2)
    w = np.stack([P[:, 0] - 1.4, np.full(len(P), 3.0), P[:, 1] - 1.4], -1)
    r2 = np.sum(w * w, 1)
    r = np.sqrt(r2)
    quad = np.sum((w[:, 1] / r) * (w[:, 1] / r) / r2) / (n * n)
    assert G == pytest.approx(quad, rel=2e-4)


@pytest.mark.parametrize("name", ["room", "random_tris", "sphere"])
def test_primary_beams_tile_the_image(name):
    """The output beams of one whole-image root cover it: sum of areas plus
    dropped slivers equals the root area (4 halfW halfH)."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    cam = camera(sp["view"], 64, 64)
    p = oracle(name).primary(cam, 64, 64, tile=0)
    total = float(np.sum(p["area"].astype(np.float64))) + p["stats"]["droppedArea"]
    root = 4 * float(cam[12]) * float(cam[13])
    assert abs(total - root) / root < 1e-3


def test_oracle_exact_and_eq2_agree_when_unoccluded():
    """Without occluders the exact integral and A * G0 * V differ only by the
    light-centre approximation, and V = 1 everywhere the light is visible."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("analytic")
    cam = camera(sp["view"], 32, 32)
    o = oracle("analytic").soft(cam, 32, 32, sp["area_light"], exact=True)
    lit = (o["vis"] > 0.999)
    assert lit.sum() > 100
    rel = np.abs(o["E"][lit] - o["Eu"][lit]) / o["Eu"][lit]
    assert np.median(rel) < 0.05


@pytest.mark.parametrize("name,side,max_err,max_frac,gain", [
    ("room", 48, 2e-6, 0.0, 100), ("random_tris", 48, 1e-4, 0.02, 1), ("plant", 48, 2e-5, 0.002, 100),
    ("atrium", 40, 1e-4, 0.01, 100), ("conference", 40, 5e-5, 0.01, 100), ("blob", 24, 5e-3, 0.03, 3)])
def test_guarded_soft_shadows_approach_the_double_build(name, side, max_err, max_frac, gain):
    """The FP64-guarded mode (double hit point, the double build's apex offset,
    the apex as a float pair, the Lambert integral in double) against the
    reference's double build: the mean
    visibility error falls by `gain` or more from the float build's (up to
    ~4e-2 on the plant, half of its pixels above 1e-5), the largest error and
    the share above 1e-5 stay within the bounds here. A float apex at the
    double build's offset (no double hit point) self-shadows instead: errors
    of 1.0 on the conference hall's floor."""
    if not ref_available("d"):
        pytest.skip("reference double build not built")
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W = H = side
    cam = camera(sp["view"], W, H)
    refd = Ref("d").load(a["pos"], a["tris"])
    d = refd.soft(sp["view"], W, H, sp["area_light"], exact=True, mailbox=False, trail=True, threads=8)
    refd.free()
    o = Oracle(a, k)
    err, eerr = {}, {}
    for guard in (False, True):
        g = o.soft(cam, W, H, sp["area_light"], exact=True, guard=guard)
        assert np.array_equal(g["tri"], d["tri"])
        m = (d["vis"] >= 0) & (g["vis"] >= 0)
        err[guard] = np.abs(g["vis"].astype(np.float64) - d["vis"])[m]
        eerr[guard] = np.abs(g["E"].astype(np.float64) - d["E"])[m] / np.abs(d["E"][m]).max()
    assert err[True].max() <= max_err
    assert np.mean(err[True] > 1e-5) <= max_frac
    assert err[True].mean() * gain <= err[False].mean()
    # the Lambert integral in double: irradiance errors (relative to the
    # image's largest E) fall by 30x or more (the float images' own rounding,
    # ~6e-8, is the floor)
    assert eerr[True].mean() * 30 <= eerr[False].mean()
    assert eerr[True].max() <= 20 * max_err


LARGE = ["blob", "atrium", "conference", "building"]


def large_golden():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "large.npz"))


@pytest.mark.parametrize("name", LARGE)
def test_oracle_matches_reference_on_benchmark_scenes(name):
    """The batched workload's scenes (noisy blob, atrium), the served one
    (conference hall) and the full 2M-triangle building: soft shadows at
    40 x 30 and tiled primary visibility and point shadows at 64 x 48 (16 x 16
    tiles), bit for bit with the reference build's (tests/golden/large.npz)."""
    from beamtrace2007_amd.host import camera
    g = large_golden()
    s, a, k, sp = load_scene(name)
    o = oracle(name)
    cam = camera(sp["view"], 40, 30)
    r = o.soft(cam, 40, 30, sp["area_light"], exact=True, threads=8)
    for f in ("tri", "vis", "E", "Eu"):
        assert np.array_equal(r[f], g[f"{name}_soft_{f}"]), f
    assert np.array_equal(np.array([r["stats"][key] for key in STAT_KEYS]), g[f"{name}_soft_stats"])
    cam = camera(sp["view"], 64, 48)
    p = o.primary(cam, 64, 48, tile=16)
    c = canon(p)
    assert len(c) == int(g[f"{name}_prim16_count"]) and digest(c) == str(g[f"{name}_prim16_digest"])
    assert np.array_equal(np.array([p["stats"][key] for key in STAT_KEYS]), g[f"{name}_prim16_stats"])
    pt = o.point(cam, 64, 48, p, sp["point_light"])
    c = canon(pt)
    assert len(c) == int(g[f"{name}_point_count"]) and digest(c) == str(g[f"{name}_point_digest"])
    assert np.array_equal(np.array([pt["stats"][key] for key in STAT_KEYS]), g[f"{name}_point_stats"])


def test_strip_of_tile_rows_equals_the_whole_frame_restricted_to_it():
    """orc_primary_tiles (a strip of tile rows) gives exactly the polygons of
    those tiles in the whole frame: roots are traced independently."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("room")
    o = oracle("room")
    W, H, tile = 96, 80, 16
    cam = camera(sp["view"], W, H)
    full = o.primary(cam, W, H, tile=tile)
    parts = [o.primary(cam, W, H, tile=tile, tile_rows=(r, r + 1)) for r in range(H // tile)]
    joined = {f: np.concatenate([p[f] for p in parts]) for f in ("x", "y", "n", "tri", "area")}
    assert np.array_equal(canon(joined), canon(full))
    assert sum(p["stats"]["triTests"] for p in parts) == full["stats"]["triTests"]
