"""The request server (btg_server_*): continuous batching of soft-shadow
requests into a running persistent pass, point-shadow requests overlapping on
their own stream. Every result must equal the same request rendered alone
(bit for bit: the exact sums make per-pixel results independent of what else
is in flight; polygon lists are compared in canonical order)."""
import threading
import time

import numpy as np
import pytest

from conftest import load_scene

pytestmark = pytest.mark.gpu


# The suite's other modules keep renderers (and their queues) alive: servers
# here take 32 GiB rather than the default 128 GiB.
SERVER_HBM = 32 << 30


def scaled_light(light9, s):
    L = np.array(light9, np.float32).copy()
    L[3:] = L[3:] * np.float32(s)
    return L


def soft_alone(name, cam, W, H, light):
    from beamtrace2007_amd.gpu import GpuRenderer
    r = GpuRenderer(load_scene(name)[0])
    out = r.soft_shadows(cam, W, H, light, exact=True)
    r.close()
    return out


def point_alone(name, cam, W, H, light3):
    from beamtrace2007_amd.gpu import GpuRenderer
    r = GpuRenderer(load_scene(name)[0])
    hits, _ = r.primary(cam, W, H, cull=False, tile=16)
    shadow, _ = r.point_shadows(cam, W, H, hits, light3)
    r.close()
    return hits, shadow


def same_polys(a, b):
    return len(a) == len(b) and all(np.array_equal(a[f], b[f]) for f in ("x", "y", "n", "tri", "area"))


@pytest.fixture()
def server_setup():
    from beamtrace2007_amd.gpu import GpuRenderer, Server
    r = GpuRenderer(load_scene("room")[0])
    plant = r.add_scene(load_scene("plant")[0])
    srv = Server(r, hbm_bytes=SERVER_HBM)
    yield srv, {"room": 0, "plant": plant}
    srv.close()
    r.close()


def test_mixed_requests_equal_requests_alone(require_gpu, server_setup):
    from beamtrace2007_amd.host import camera
    srv, ids = server_setup
    plan = [("soft", "plant", 64, 48, 1.0), ("point", "room", 48, 48, None), ("soft", "room", 40, 30, 1.0),
            ("soft", "plant", 96, 96, 2.0), ("point", "room", 64, 32, None), ("soft", "room", 17, 9, 0.5),
            ("soft", "plant", 33, 65, 0.7), ("point", "room", 16, 16, None)]
    tickets = []
    for kind, name, W, H, s in plan:
        sp = load_scene(name)[3]
        cam = camera(sp["view"], W, H)
        if kind == "soft":
            light = scaled_light(sp["area_light"], s)
            tickets.append((kind, name, cam, W, H, light, srv.submit_soft(ids[name], cam, W, H, light)))
        else:
            tickets.append((kind, name, cam, W, H, sp["point_light"], srv.submit_point(cam, W, H, sp["point_light"])))
    print("submitted", len(tickets), flush=True)
    for kind, name, cam, W, H, light, t in tickets:
        res = srv.wait(t, timeout=120)
        assert res is not None and res["status"] == 0, res
        assert res["latency_s"] >= res["queue_s"] >= 0
        if kind == "soft":
            a = soft_alone(name, cam, W, H, light)
            for f in ("tri", "vis", "E", "Eu"):
                assert np.array_equal(res[f], a[f]), (name, W, H, f)
        else:
            hits, shadow = point_alone(name, cam, W, H, light)
            assert same_polys(res["hits"], hits) and same_polys(res["polys"], shadow), (name, W, H)


def test_requests_join_a_running_pass(require_gpu, server_setup):
    """A small request submitted while a large one is on the device is admitted
    into the running pass and finishes first."""
    from beamtrace2007_amd.host import camera
    srv, ids = server_setup
    sp = load_scene("plant")[3]
    big_cam = camera(sp["view"], 384, 384)
    big = srv.submit_soft(ids["plant"], big_cam, 384, 384, scaled_light(sp["area_light"], 2.0))
    time.sleep(0.05)
    small_cam = camera(sp["view"], 24, 24)
    small = srv.submit_soft(ids["plant"], small_cam, 24, 24, sp["area_light"])
    rs = srv.wait(small, timeout=120)
    done_big_early = srv.wait(big, timeout=0)
    rb = done_big_early or srv.wait(big, timeout=300)
    assert rs["status"] == 0 and rb["status"] == 0
    assert done_big_early is None, "the small request should not wait for the large one"
    a = soft_alone("plant", small_cam, 24, 24, sp["area_light"])
    assert np.array_equal(rs["vis"], a["vis"])


def test_passes_alternate_when_full(require_gpu):
    """With room for only a few requests per pass, the server fills one pass,
    hands over to the other, and reuses each once drained; results stay exact."""
    from beamtrace2007_amd.gpu import GpuRenderer, Server
    from beamtrace2007_amd.host import camera
    r = GpuRenderer(load_scene("plant")[0])
    r.set_options(capacity=1 << 20)
    # per pass: ~0.23 GiB of queues and ~2M pixels of images and root state
    srv = Server(r, hbm_bytes=int(2 * (0.25 * 2**30 + 2.2e6 * 204)), jobs_per_pass=16)
    sp = load_scene("plant")[3]
    reqs = []
    lock = threading.Lock()

    def submitter():
        for k in range(10):
            side = 300 + 37 * k
            cam = camera(sp["view"], side, side)
            t = srv.submit_soft(0, cam, side, side, sp["area_light"])
            with lock:
                reqs.append((cam, side, t))
            time.sleep(0.01)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    for cam, side, t in reqs:
        res = srv.wait(t, timeout=300)
        assert res is not None and res["status"] == 0, res
        a = r_alone = soft_alone("plant", cam, side, side, sp["area_light"])
        assert np.array_equal(res["vis"], a["vis"]) and np.array_equal(res["E"], a["E"]), side
        del r_alone
    srv.close()
    r.close()


def test_guarded_requests_equal_guarded_alone(require_gpu):
    """A server on a context in the FP64-guarded mode runs its passes guarded:
    results equal the same requests rendered alone in that mode."""
    from beamtrace2007_amd.gpu import GpuRenderer, Server
    from beamtrace2007_amd.host import camera
    r = GpuRenderer(load_scene("room")[0])
    r.set_options(precision=1)
    srv = Server(r, hbm_bytes=SERVER_HBM)
    sp = load_scene("room")[3]
    plan = [(40, 30, 1.0), (17, 9, 0.5), (64, 48, 2.0)]
    tickets = []
    for W, H, s in plan:
        cam = camera(sp["view"], W, H)
        light = scaled_light(sp["area_light"], s)
        tickets.append((cam, W, H, light, srv.submit_soft(0, cam, W, H, light)))
    for cam, W, H, light, t in tickets:
        res = srv.wait(t, timeout=120)
        assert res is not None and res["status"] == 0, res
        a = r.soft_shadows(cam, W, H, light, exact=True)
        for f in ("tri", "vis", "E", "Eu"):
            assert np.array_equal(res[f], a[f]), (W, H, f)
    srv.close()
    r.close()


def test_fp64_requests_equal_fp64_alone(require_gpu):
    """A server on a context set to FP64 soft shadows (precision 2) serves its
    soft requests in FP64 batches: requests submitted together share a batch,
    a request for a scene that is not resident fails alone, a point request
    runs beside them, and every soft result equals the same job rendered
    alone in FP64 mode from the same double camera, bit for bit (the
    per-pixel sums are fixed point); against the reference's double build it
    is within 1e-6 on every pixel with the same primary hit."""
    from beamtrace2007_amd.gpu import GpuRenderer, Server
    from beamtrace2007_amd.host import camera, camera_d
    from refpy import Ref, ref_available
    s, a, k, sp = load_scene("room")
    plan = [(40, 30, 1.0), (24, 24, 2.0), (32, 40, 0.5)]
    r = GpuRenderer(s)
    r.set_options(precision=2)
    alone = []
    for W, H, sc in plan:
        cam, cd = camera(sp["view"], W, H), camera_d(sp["view"], W, H)
        light = scaled_light(sp["area_light"], sc)
        alone.append((cam, cd, W, H, light, r.soft_shadows(cam, W, H, light, exact=True, cam_d=cd)))
    srv = Server(r, hbm_bytes=SERVER_HBM)
    tickets = [srv.submit_soft(0, cam, W, H, light, cam_d=cd) for cam, cd, W, H, light, _ in alone]
    bad = srv.submit_soft(7, alone[0][0], 40, 30, alone[0][4])
    pcam = camera(sp["view"], 32, 32)
    point = srv.submit_point(pcam, 32, 32, sp["point_light"])
    print(f"submitted {len(tickets)} FP64 soft requests, a bad one and a point request", flush=True)
    results = [srv.wait(t, timeout=300) for t in tickets]
    rb, rp = srv.wait(bad, timeout=300), srv.wait(point, timeout=300)
    srv.close()
    r.close()
    assert rb["status"] != 0 and "scene" in rb["error"]
    assert rp["status"] == 0
    hits, shadow = point_alone("room", pcam, 32, 32, sp["point_light"])
    assert same_polys(rp["hits"], hits) and same_polys(rp["polys"], shadow)
    refd = Ref("d").load(a["pos"], a["tris"]) if ref_available("d") else None
    for (cam, cd, W, H, light, g), res in zip(alone, results):
        assert res is not None and res["status"] == 0, res
        for f in ("tri", "vis", "E", "Eu"):
            assert np.array_equal(res[f], g[f]), (W, H, f)
        if refd is not None:
            d = refd.soft(sp["view"], W, H, light, exact=True, mailbox=False, trail=True, threads=16)
            m = (d["vis"] >= 0) & (res["vis"] >= 0) & (res["tri"] == d["tri"])
            err = np.abs(res["vis"].astype(np.float64) - d["vis"])[m]
            print(f"served FP64 {W}x{H}: max |vis - vis_fp64| {err.max():.2e} on {m.sum()} px", flush=True)
            assert err.max() <= 1e-6
    if refd is not None:
        refd.free()


def test_requests_over_budget_are_traced_again(require_gpu):
    """Roots that go over their live sub-beam budget inside a served pass are
    abandoned there, and their request is traced again alone on the server's
    fallback engine (GPU passes with a larger budget): at a budget of 4 the
    plant's roots go over, and every served result, soft and point, still
    equals the request rendered alone at the default budget."""
    from beamtrace2007_amd.gpu import GpuRenderer, Server
    from beamtrace2007_amd.host import camera
    sp = load_scene("plant")[3]
    plan = [(48, 40, 1.0), (32, 32, 2.0), (40, 56, 0.5)]
    r = GpuRenderer(load_scene("plant")[0])
    r.set_options(budget=4, primary_budget=4)
    cam0 = camera(sp["view"], 48, 40)
    killed = r.soft_shadows(cam0, 48, 40, sp["area_light"], exact=True)["info"]["killed_roots"]
    assert killed > 0, "a budget of 4 should abandon some of the plant's roots"
    srv = Server(r, hbm_bytes=SERVER_HBM)
    tickets = []
    for W, H, s in plan:
        cam = camera(sp["view"], W, H)
        light = scaled_light(sp["area_light"], s)
        tickets.append(("soft", cam, W, H, light, srv.submit_soft(0, cam, W, H, light)))
    pcam = camera(sp["view"], 40, 40)
    tickets.append(("point", pcam, 40, 40, sp["point_light"], srv.submit_point(pcam, 40, 40, sp["point_light"])))
    print(f"submitted {len(tickets)} requests at budget 4 ({killed} roots of the first over budget alone)",
          flush=True)
    results = []
    for kind, cam, W, H, light, t in tickets:
        res = srv.wait(t, timeout=120)
        assert res is not None and res["status"] == 0, res
        results.append((kind, cam, W, H, light, res))
    srv.close()
    r.close()
    for kind, cam, W, H, light, res in results:
        if kind == "soft":
            a = soft_alone("plant", cam, W, H, light)
            for f in ("tri", "vis", "E", "Eu"):
                assert np.array_equal(res[f], a[f]), (W, H, f)
        else:
            hits, shadow = point_alone("plant", cam, W, H, light)
            assert same_polys(res["hits"], hits) and same_polys(res["polys"], shadow), (W, H)


def test_bad_requests_fail_alone(require_gpu, server_setup):
    """A request for a scene that is not resident, or one larger than a whole
    pass, finishes with an error status and a message; requests around it
    are served normally. An invalid camera is refused at submission."""
    from beamtrace2007_amd.host import camera
    srv, ids = server_setup
    sp = load_scene("room")[3]
    good_cam = camera(sp["view"], 32, 24)
    t0 = srv.submit_soft(ids["room"], good_cam, 32, 24, sp["area_light"])
    bad = srv.submit_soft(99, good_cam, 32, 24, sp["area_light"])
    huge = srv.submit_soft(ids["room"], camera(sp["view"], 8192, 8192), 8192, 8192, sp["area_light"])
    t1 = srv.submit_soft(ids["room"], good_cam, 32, 24, sp["area_light"])
    for t in (bad, huge):
        res = srv.wait(t, timeout=120)
        assert res is not None and res["status"] != 0
    for t in (t0, t1):
        res = srv.wait(t, timeout=120)
        assert res is not None and res["status"] == 0
        a = soft_alone("room", good_cam, 32, 24, sp["area_light"])
        assert np.array_equal(res["vis"], a["vis"])
    with pytest.raises(RuntimeError):
        srv.submit_soft(ids["room"], camera(sp["view"], 32, 24), 0, 24, sp["area_light"])


def test_two_waiters_on_one_ticket(require_gpu, server_setup):
    """Two threads wait on the same ticket: exactly one receives the result,
    the other an error (the ticket is released once), and neither hangs."""
    from beamtrace2007_amd.host import camera
    srv, ids = server_setup
    sp = load_scene("room")[3]
    cam = camera(sp["view"], 48, 40)
    t = srv.submit_soft(ids["room"], cam, 48, 40, sp["area_light"])
    got, errors = [], []

    def waiter():
        try:
            got.append(srv.wait(t, timeout=120))
        except RuntimeError as e:
            errors.append(str(e))

    th = [threading.Thread(target=waiter) for _ in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=180)
    assert not any(x.is_alive() for x in th)
    assert len(got) == 1 and len(errors) == 1
    assert got[0] is not None and got[0]["status"] == 0
    assert np.array_equal(got[0]["vis"], soft_alone("room", cam, 48, 40, sp["area_light"])["vis"])


def test_footprint_of_requests(require_gpu, server_setup):
    """btg_footprint: soft requests hold their per-pixel images and hit points
    (40 B) and the state of 64 roots per 8 x 8 tile, partial tiles included;
    point requests their 16 x 16 tile roots and the primary polygon buffer
    (at least 2^20 polygons of 44 B plus root and tag words)."""
    from beamtrace2007_amd.host import camera
    srv, ids = server_setup
    sp = load_scene("room")[3]
    fp = lambda kind, W, H: srv.footprint(kind, camera(sp["view"], W, H), W, H)
    a = fp("soft", 8, 8)
    root = (a - 64 * 40) // 64
    assert a == 64 * 40 + 64 * root and 150 <= root <= 400
    assert fp("soft", 16, 8) == 2 * a
    assert fp("soft", 9, 8) == 72 * 40 + 128 * root
    assert fp("soft", 1024, 768) == 1024 * 768 * (40 + root)
    assert fp("point", 16, 16) == root + (1 << 20) * 52
    assert fp("point", 2048, 2048) == 128 * 128 * root + 2 * 2048 * 2048 * 52
