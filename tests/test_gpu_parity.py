"""GPU path (lib/libbtg.so, gfx950 kernels) against the oracle.

Bar: primary hit ids, split topology (polygon multisets) and point shadows
bit-exact; soft-shadow visible fraction within 1e-5 absolute (in practice
bit-exact: the engine sums piece areas exactly); E (exact Lambert integral)
within 1e-5 relative, since the reference itself sums it in float in an
order that depends on the traversal; Eu bit-exact.
"""
import numpy as np
import pytest

from conftest import VIS_TOL, check_soft, load_scene
from refpy import Oracle, canon

pytestmark = pytest.mark.gpu

SCENES = ["analytic", "room", "random_tris", "sphere", "plant"]
_renderers, _oracles = {}, {}


def renderer(name):
    from beamtrace2007_amd.gpu import GpuRenderer
    if name not in _renderers:
        s, a, k, sp = load_scene(name)
        _renderers[name] = GpuRenderer(s)
    return _renderers[name]


def oracle(name):
    if name not in _oracles:
        s, a, k, sp = load_scene(name)
        _oracles[name] = Oracle(a, k)
    return _oracles[name]


@pytest.mark.parametrize("name", SCENES)
def test_soft_shadows_match_oracle(require_gpu, name):
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W = H = 80
    cam = camera(sp["view"], W, H)
    print(f"soft {name} {W}x{H}", flush=True)
    g = renderer(name).soft_shadows(cam, W, H, sp["area_light"], exact=True)
    o = oracle(name).soft(cam, W, H, sp["area_light"], exact=True)
    check_soft(g, o)
    if g["info"]["killed_roots"] == 0:  # counters are engine-independent when no root is traced twice
        for key in ("kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "rootBeams", "presplitBeams",
                    "fiveSplits", "climbs"):
            assert g["info"]["stats"][key] == o["stats"][key], key


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("tile", [8, 16])
def test_primary_polygons_match_oracle(require_gpu, name, tile):
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 96, 72
    cam = camera(sp["view"], W, H)
    print(f"primary {name} tile {tile}", flush=True)
    gp, info = renderer(name).primary(cam, W, H, tile=tile)
    op = oracle(name).primary(cam, W, H, tile=tile)
    assert np.array_equal(canon(polys_to_dict(gp)), canon(op))
    if info["killed_roots"] == 0:
        for key in ("kdSteps", "leafVisits", "triTests", "splits", "beams", "fiveSplits", "climbs"):
            assert info["stats"][key] == op["stats"][key], key


@pytest.mark.parametrize("name", SCENES)
def test_point_shadows_and_masks_match_oracle(require_gpu, name):
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera, rasterize
    s, a, k, sp = load_scene(name)
    W, H = 96, 72
    cam = camera(sp["view"], W, H)
    print(f"point {name}", flush=True)
    gp, _ = renderer(name).primary(cam, W, H, tile=8)
    gs, _ = renderer(name).point_shadows(cam, W, H, gp, sp["point_light"])
    prim = polys_to_dict(gp)
    os_ = oracle(name).point(cam, W, H, prim, sp["point_light"])
    assert np.array_equal(canon(polys_to_dict(gs)), canon(os_))
    # per-pixel ids and shadow mask (polygons rasterized in canonical order)
    ids_g = rasterize(cam, W, H, prim)
    ids_o = rasterize(cam, W, H, oracle(name).primary(cam, W, H, tile=8))
    assert np.array_equal(ids_g, ids_o)
    assert np.array_equal(rasterize(cam, W, H, polys_to_dict(gs), shadow=True),
                          rasterize(cam, W, H, os_, shadow=True))


@pytest.mark.parametrize("name", ["room", "sphere", "random_tris"])
def test_primary_with_back_face_culling_matches_oracle(require_gpu, name):
    """primary(cull = true) (cuda/gpu_render.h:41): back faces never hit."""
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 96, 72
    cam = camera(sp["view"], W, H)
    gp, _ = renderer(name).primary(cam, W, H, cull=True, tile=8)
    op = oracle(name).primary(cam, W, H, cull=True, tile=8)
    assert np.array_equal(canon(polys_to_dict(gp)), canon(op))
    go, _ = renderer(name).primary(cam, W, H, cull=False, tile=8)
    if name == "random_tris":  # open, randomly oriented triangles: culling changes what is seen
        assert not np.array_equal(canon(polys_to_dict(gp)), canon(polys_to_dict(go)))


LARGE = ["atrium", "conference", "blob", "building_small"]


@pytest.mark.parametrize("name", LARGE)
def test_soft_shadows_on_benchmark_scenes_match_oracle(require_gpu, name):
    """The batched workload's scenes (blob, atrium), the served one
    (conference hall) and a smaller building, in the default float mode."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 64, 48
    cam = camera(sp["view"], W, H)
    print(f"soft {name} {W}x{H}", flush=True)
    g = renderer(name).soft_shadows(cam, W, H, sp["area_light"], exact=True)
    o = oracle(name).soft(cam, W, H, sp["area_light"], exact=True)
    check_soft(g, o)
    if g["info"]["killed_roots"] == 0:
        for key in ("kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "climbs"):
            assert g["info"]["stats"][key] == o["stats"][key], key


@pytest.mark.parametrize("name", ["atrium", "conference", "building_small"])
def test_primary_and_point_shadows_on_benchmark_scenes_match_oracle(require_gpu, name):
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 96, 72
    cam = camera(sp["view"], W, H)
    print(f"primary + point {name}", flush=True)
    gp, _ = renderer(name).primary(cam, W, H, tile=16)
    op = oracle(name).primary(cam, W, H, tile=16)
    assert np.array_equal(canon(polys_to_dict(gp)), canon(op))
    gs, _ = renderer(name).point_shadows(cam, W, H, gp, sp["point_light"])
    os_ = oracle(name).point(cam, W, H, op, sp["point_light"])
    assert np.array_equal(canon(polys_to_dict(gs)), canon(os_))


def test_plant_at_full_size_matches_oracle_on_row_bands(require_gpu):
    """The soft-shadow configuration at its full size (512 x 512): the GPU's
    whole image against the oracle on five bands of eight rows."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("plant")
    W = H = 512
    cam = camera(sp["view"], W, H)
    g = renderer("plant").soft_shadows(cam, W, H, sp["area_light"], exact=True)
    for y0 in (0, 128, 256, 384, 504):
        print(f"plant 512 rows {y0}..{y0 + 8}", flush=True)
        o = oracle("plant").soft(cam, W, H, sp["area_light"], exact=True, threads=16, rows=(y0, y0 + 8))
        band = slice(y0, y0 + 8)
        check_soft({f: g[f][band] for f in ("tri", "vis", "E", "Eu")}, {f: o[f][band] for f in ("tri", "vis", "E", "Eu")})


def test_room_at_full_size_primary_matches_oracle(require_gpu):
    """The primary-visibility configuration at its full size: 512 x 512 in
    8 x 8 tiles (4096 roots), polygons and per-pixel ids."""
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera, rasterize
    s, a, k, sp = load_scene("room")
    W = H = 512
    cam = camera(sp["view"], W, H)
    gp, info = renderer("room").primary(cam, W, H, tile=8)
    op = oracle("room").primary(cam, W, H, tile=8)
    assert np.array_equal(canon(polys_to_dict(gp)), canon(op))
    assert np.array_equal(rasterize(cam, W, H, polys_to_dict(gp)), rasterize(cam, W, H, op))
    assert info["stats"]["triTests"] == op["stats"]["triTests"]


def test_odd_image_sizes_and_partial_tiles(require_gpu):
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("room")
    for W, H in [(37, 23), (1, 1), (9, 130)]:
        cam = camera(sp["view"], W, H)
        print(f"odd size {W}x{H}", flush=True)
        check_soft(renderer("room").soft_shadows(cam, W, H, sp["area_light"]), oracle("room").soft(cam, W, H, sp["area_light"]))
        gp, _ = renderer("room").primary(cam, W, H, tile=16)
        assert np.array_equal(canon(polys_to_dict(gp)), canon(oracle("room").primary(cam, W, H, tile=16)))


def test_background_and_unlit_pixels(require_gpu):
    """A camera that sees nothing gives tri = -1 / vis = -1 / E = Eu = 0; a light
    under the floor leaves every receiver at vis = 0."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("analytic")
    away = np.array([0, 5, 5, 0, 10, 20, 0, 1, 0, 30], np.float32)
    cam = camera(away, 24, 24)
    g = renderer("analytic").soft_shadows(cam, 24, 24, sp["area_light"])
    assert np.all(g["tri"] == -1) and np.all(g["vis"] == -1) and np.all(g["E"] == 0) and np.all(g["Eu"] == 0)
    below = np.array([0, -3, 0, 1, 0, 0, 0, 0, 1], np.float32)
    cam = camera(sp["view"], 32, 32)
    g = renderer("analytic").soft_shadows(cam, 32, 32, below)
    o = oracle("analytic").soft(cam, 32, 32, below)
    check_soft(g, o)
    assert np.all(g["vis"][g["tri"] >= 0] == 0)


def test_tiny_budget_retries_on_the_gpu_and_stays_exact(require_gpu):
    """Budget 4 and a 4096-record queue force many roots over budget (the
    reference's tests/cuda case, cuda/test_cuda.cpp:264-302); they are traced
    again on the GPU with larger budgets and the results stay exact."""
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("plant")
    r = GpuRenderer(s)
    r.set_options(budget=4, capacity=4096)
    W = H = 48
    cam = camera(sp["view"], W, H)
    g = r.soft_shadows(cam, W, H, sp["area_light"])
    assert g["info"]["killed_roots"] > 100 and g["info"]["retry_passes"] >= 1
    check_soft(g, oracle("plant").soft(cam, W, H, sp["area_light"]))
    r.set_options(primary_budget=4, budget=4)
    gp, info = r.primary(cam, W, H, tile=16)
    assert info["killed_roots"] > 0
    from beamtrace2007_amd.gpu import polys_to_dict
    assert np.array_equal(canon(polys_to_dict(gp)), canon(oracle("plant").primary(cam, W, H, tile=16)))
    r.close()


def test_repeated_runs_are_bitwise_identical(require_gpu):
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("plant")
    cam = camera(sp["view"], 64, 64)
    x = renderer("plant").soft_shadows(cam, 64, 64, sp["area_light"])
    y = renderer("plant").soft_shadows(cam, 64, 64, sp["area_light"])
    for key in ("tri", "vis", "E", "Eu"):
        assert np.array_equal(x[key], y[key])


def test_invalid_arguments_are_errors_not_crashes(require_gpu):
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("room")
    cam = camera(sp["view"], 16, 16)
    r = renderer("room")
    with pytest.raises(RuntimeError):
        r.primary(cam, 16, 16, tile=0)
    with pytest.raises(RuntimeError):
        r.soft_shadows(cam, 0, 16, sp["area_light"])
    for W, H in ((40000, 16), (16, 40000)):  # sides above 32768: refused before any allocation
        with pytest.raises(RuntimeError):
            r.primary(camera(sp["view"], W, H), W, H, tile=16)
    bad = np.zeros(1, dtype=__import__("beamtrace2007_amd.gpu", fromlist=["POLY_DTYPE"]).POLY_DTYPE)
    bad["tri"] = -1
    with pytest.raises(RuntimeError):
        r.point_shadows(cam, 16, 16, bad, sp["point_light"])
    empty, _ = r.point_shadows(cam, 16, 16, bad[:0], sp["point_light"])
    assert len(empty) == 0


@pytest.mark.parametrize("name", ["room", "plant", "random_tris", "conference", "atrium"])
def test_guarded_soft_shadows_match_guarded_oracle(require_gpu, name):
    """precision = 1 (double hit point, the double build's apex offset, the
    apex as a float pair) on the GPU equals the oracle's guarded variant; its
    error against the double build is checked on the CPU (test_oracle.py)."""
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 72, 56
    cam = camera(sp["view"], W, H)
    r = renderer(name)
    saved = r.options
    r.set_options(precision=1)
    try:
        g = r.soft_shadows(cam, W, H, sp["area_light"], exact=True)
    finally:
        r.set_options(**saved)
    o = oracle(name).soft(cam, W, H, sp["area_light"], exact=True, guard=True)
    check_soft(g, o)


@pytest.mark.parametrize("name", ["room", "conference"])
def test_primary_and_point_independent_of_leaf_cap(require_gpu, name):
    """Yields (a sub-beam's record continues in the next round) and the pooled
    nearest-hit tests restarting after an all-in hit leave polygons and
    counters unchanged: leaf cap 1 (a yield after every leaf) and no cap equal
    the default (primary passes at most 4) bit for bit."""
    from beamtrace2007_amd.gpu import GpuRenderer, polys_to_dict
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene(name)
    W, H = 160, 120
    cam = camera(sp["view"], W, H)
    r = GpuRenderer(s)
    res = {}
    for cap in (None, 1, 0):
        if cap is not None:
            r.set_options(leaf_cap=cap)
        print(f"{name} leaf cap {cap}", flush=True)
        hits, info = r.primary(cam, W, H, tile=16)
        shadow, info2 = r.point_shadows(cam, W, H, hits, sp["point_light"])
        res[cap] = (canon(polys_to_dict(hits)), canon(polys_to_dict(shadow)), info["stats"], info2["stats"])
    r.close()
    for cap in (1, 0):
        assert np.array_equal(res[cap][0], res[None][0]) and np.array_equal(res[cap][1], res[None][1])
        for key in ("kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "climbs"):
            assert res[cap][2][key] == res[None][2][key], key
            assert res[cap][3][key] == res[None][3][key], key


@pytest.mark.parametrize("name", ["plant", "blob", "conference"])
def test_mailbox_changes_nothing_but_the_test_count(require_gpu, name):
    """Post Office mailboxing on the GPU (btg_options.mailbox): a sub-beam skips
    triangles it has already passed unchanged. Images, polygons and every
    counter but triTests are identical with and without it, and each skip
    replaces exactly one test (triTests + mailboxSkips is unchanged)."""
    from beamtrace2007_amd.gpu import polys_to_dict
    from beamtrace2007_amd.host import camera
    from refpy import canon
    s, a, k, sp = load_scene(name)
    r = renderer(name)
    W, H = 64, 48
    cam = camera(sp["view"], W, H)
    saved = r.options
    runs = {}
    try:
        for mb in (False, True):
            r.set_options(mailbox=mb)
            print(f"{name} mailbox {mb}", flush=True)
            soft = r.soft_shadows(cam, W, H, sp["area_light"], exact=True)
            prim, pinfo = r.primary(cam, W, H, tile=16)
            pt, qinfo = r.point_shadows(cam, W, H, prim, sp["point_light"])
            runs[mb] = (soft, prim, pinfo, pt, qinfo)
    finally:
        r.set_options(**saved)
    (s0, p0, pi0, q0, qi0), (s1, p1, pi1, q1, qi1) = runs[False], runs[True]
    for f in ("tri", "vis", "E", "Eu"):
        assert np.array_equal(s0[f], s1[f]), f
    assert np.array_equal(canon(polys_to_dict(p0)), canon(polys_to_dict(p1)))
    assert np.array_equal(canon(polys_to_dict(q0)), canon(polys_to_dict(q1)))
    skips = 0
    for i0, i1 in ((s0["info"], s1["info"]), (pi0, pi1), (qi0, qi1)):
        if i0["killed_roots"] or i1["killed_roots"]:
            continue  # roots traced twice: counters are not comparable
        a0, a1 = i0["stats"], i1["stats"]
        assert a0["mailboxSkips"] == 0
        assert a1["triTests"] + a1["mailboxSkips"] == a0["triTests"]
        for key in ("kdSteps", "leafVisits", "hits", "splits", "beams", "climbs"):
            assert a0[key] == a1[key], key
        skips += a1["mailboxSkips"]
    print(f"{name}: {skips} mailbox skips", flush=True)


FP64_CASES = [("plant", 1.0, 48), ("atrium", 4.0, 48), ("conference", 1.0, 48), ("blob", 2.0, 32), ("room", 1.0, 48)]


@pytest.mark.parametrize("name,scale,side", FP64_CASES)
def test_fp64_soft_shadows_match_the_double_build(require_gpu, name, scale, side):
    """precision = 2 (the whole beam trace in double with the double build's
    tolerances) against bt_render_d's algorithm (the reference compiled with
    Real = double, oracle/_ref), on every pixel with the same primary hit:
    * with the double build's camera (camera_d): the visible fraction and the
      exact irradiance agree to 1e-6 (the float32 rounding of the outputs;
      the precision target is 1e-5);
    * with the float camera alone: within 1e-4 (its vectors' rounding moves
      receiver points by ~1e-8 of the distance: up to ~4e-5 of visible
      fraction at sharp penumbrae)."""
    from beamtrace2007_amd.host import camera, camera_d
    from beamtrace2007_amd.workloads import scaled_light
    from refpy import Ref, ref_available
    if not ref_available("d"):
        pytest.skip("reference double build (oracle/_ref) not present")
    s, a, k, sp = load_scene(name)
    W = H = side
    cam = camera(sp["view"], W, H)
    light = scaled_light(sp["area_light"], scale)
    r = renderer(name)
    saved = r.options
    r.set_options(precision=2)
    try:
        print(f"fp64 {name} {W}x{H} light x{scale}", flush=True)
        g = r.soft_shadows(cam, W, H, light, exact=True, cam_d=camera_d(sp["view"], W, H))
        gf = r.soft_shadows(cam, W, H, light, exact=True)
    finally:
        r.set_options(**saved)
    refd = Ref("d").load(a["pos"], a["tris"])
    d = refd.soft(sp["view"], W, H, light, exact=True, mailbox=False, trail=True, threads=16)
    refd.free()
    for img, tol, what in ((g, 1e-6, "double camera"), (gf, 1e-4, "float camera")):
        m = (d["vis"] >= 0) & (img["vis"] >= 0) & (img["tri"] == d["tri"])
        assert m.sum() >= 0.9 * (d["vis"] >= 0).sum()  # the same primary hit on nearly every lit pixel
        err = np.abs(img["vis"].astype(np.float64) - d["vis"])[m]
        e_err = np.abs(img["E"].astype(np.float64) - d["E"])[m] / max(float(np.abs(d["E"][m]).max()), 1e-30)
        print(f"fp64 {name} ({what}): vis max {err.max():.2e} mean {err.mean():.2e}; E max {e_err.max():.2e}",
              flush=True)
        assert err.max() <= tol
        assert e_err.max() <= max(tol, 1e-6)
        assert np.array_equal(img["vis"] < 0, img["tri"] < 0)
    assert np.array_equal(g["Eu"] >= 0, gf["Eu"] >= 0)
