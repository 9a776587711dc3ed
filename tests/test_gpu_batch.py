"""Batched soft shadows (btg_soft_batch): jobs on several resident scenes, of
uneven sizes and light scales, traced in one persistent pass. Each job's
outputs must equal the same job rendered alone (bit for bit: the exact sums
make results independent of what else is in flight), and alone results are
checked against the oracle elsewhere (test_gpu_parity.py); one job here is
also checked against the oracle directly."""
import numpy as np
import pytest

from conftest import load_scene
from refpy import Oracle

pytestmark = pytest.mark.gpu

VIS_TOL = 1e-5


def scaled_light(light9, s):
    L = np.array(light9, np.float32).copy()
    L[3:] = L[3:] * np.float32(s)  # bt_render: L.U = cfg.lightU * Real(lightScale) (apps/bt_render.cpp:289-290)
    return L


def jobs_for(ids):
    from beamtrace2007_amd.host import camera
    spec = {name: load_scene(name)[3] for name in ids}
    plan = [("room", 40, 30, 1.0), ("plant", 64, 48, 1.0), ("random_tris", 33, 17, 0.5), ("plant", 80, 80, 2.0),
            ("room", 8, 8, 0.25), ("analytic", 24, 40, 1.5), ("plant", 1, 1, 1.0)]
    jobs = []
    for name, W, H, s in plan:
        sp = spec[name]
        jobs.append({"name": name, "scene": ids[name], "cam": camera(sp["view"], W, H), "W": W, "H": H,
                     "light": scaled_light(sp["area_light"], s)})
    return jobs


@pytest.fixture(scope="module")
def batch_renderer():
    from beamtrace2007_amd.gpu import GpuRenderer
    s0 = load_scene("room")[0]
    r = GpuRenderer(s0)
    ids = {"room": 0}
    for name in ("plant", "random_tris", "analytic"):
        ids[name] = r.add_scene(load_scene(name)[0])
    yield r, ids
    r.close()


def test_batch_jobs_equal_jobs_alone(require_gpu, batch_renderer):
    from beamtrace2007_amd.gpu import GpuRenderer
    r, ids = batch_renderer
    jobs = jobs_for(ids)
    print("batch of", len(jobs), "jobs", flush=True)
    outs, info = r.soft_batch(jobs, exact=True)
    assert info["killed_roots"] == 0
    alone = {}
    for j, o in zip(jobs, outs):
        if j["name"] not in alone:
            alone[j["name"]] = GpuRenderer(load_scene(j["name"])[0])
        a = alone[j["name"]].soft_shadows(j["cam"], j["W"], j["H"], j["light"], exact=True)
        for f in ("tri", "vis", "E", "Eu"):
            assert np.array_equal(o[f], a[f]), (j["name"], j["W"], j["H"], f)
    for a in alone.values():
        a.close()


def test_batch_job_matches_oracle(require_gpu, batch_renderer):
    r, ids = batch_renderer
    jobs = [j for j in jobs_for(ids) if j["name"] == "plant" and j["W"] == 80]
    outs, _ = r.soft_batch(jobs, exact=True)
    s, a, k, sp = load_scene("plant")
    o = Oracle(a, k).soft(jobs[0]["cam"], 80, 80, jobs[0]["light"], exact=True)
    assert np.array_equal(outs[0]["tri"], o["tri"])
    assert np.max(np.abs(outs[0]["vis"].astype(np.float64) - o["vis"])) <= VIS_TOL


def test_batch_of_one_equals_soft_shadows(require_gpu, batch_renderer):
    r, ids = batch_renderer
    j = jobs_for(ids)[0]
    outs, _ = r.soft_batch([j], exact=False)
    a = r.soft_shadows(j["cam"], j["W"], j["H"], j["light"], exact=False)
    for f in ("tri", "vis", "E", "Eu"):
        assert np.array_equal(outs[0][f], a[f]), f


def test_scene_residency_add_remove(require_gpu):
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera
    r = GpuRenderer(load_scene("analytic")[0])
    sid = r.add_scene(load_scene("plant")[0])
    assert sid >= 1 and r.scene_bytes(sid) > 0 and r.scene_bytes(0) > 0
    sp = load_scene("plant")[3]
    job = {"scene": sid, "cam": camera(sp["view"], 32, 24), "W": 32, "H": 24, "light": sp["area_light"]}
    before, _ = r.soft_batch([job])
    r.remove_scene(sid)
    assert r.scene_bytes(sid) == -1
    with pytest.raises(RuntimeError):
        r.soft_batch([job])
    with pytest.raises(RuntimeError):
        r.remove_scene(0)
    sid2 = r.add_scene(load_scene("plant")[0])
    job["scene"] = sid2
    after, _ = r.soft_batch([job])
    for f in ("tri", "vis", "E", "Eu"):
        assert np.array_equal(before[0][f], after[0][f]), f
    r.close()


def test_autotune_returns_valid_options_and_results_do_not_change(require_gpu, batch_renderer):
    r, ids = batch_renderer
    jobs = jobs_for(ids)[:4]
    before, _ = r.soft_batch(jobs, exact=True)
    opts, trials = r.autotune(jobs, exact=True, max_seconds=10.0)
    assert opts["leaf_cap"] in (2, 4, 8, 16, 32) and opts["budget"] == r.options["budget"]
    assert opts["rounds_per_check"] in (4, 8, 16, 32) and len(trials) >= 1
    saved = r.options
    r.set_options(**opts)
    after, _ = r.soft_batch(jobs, exact=True)
    r.set_options(**saved)
    for b, a in zip(before, after):
        for f in ("tri", "vis", "E", "Eu"):
            assert np.array_equal(b[f], a[f]), f


def test_graph_replays_equal_direct_launches(require_gpu, batch_renderer):
    """Rounds replayed from captured HIP graphs (the default) give the same
    results as direct launches, on every replay: three batched calls in a row
    with other work between them, then the same batch with graphs off (and
    with per-launch profiling, which launches directly); primary and
    point-shadow passes likewise."""
    from beamtrace2007_amd.host import camera
    r, ids = batch_renderer
    jobs = jobs_for(ids)[:5]
    saved = r.options
    try:
        r.set_options(graphs=True, profile=False, rounds_per_check=4)
        runs = []
        for k in range(3):
            outs, info = r.soft_batch(jobs, exact=True)
            runs.append(outs)
            r.soft_shadows(jobs[k]["cam"], jobs[k]["W"], jobs[k]["H"], jobs[k]["light"], exact=False)
        for opts in ({"graphs": False}, {"graphs": True, "profile": True}):
            r.set_options(**opts)
            direct, dinfo = r.soft_batch(jobs, exact=True)
            for outs in runs:
                for a, b in zip(outs, direct):
                    for f in ("tri", "vis", "E", "Eu"):
                        assert np.array_equal(a[f], b[f]), f
        r.set_options(profile=False)
        sp = load_scene("room")[3]
        cam = camera(sp["view"], 96, 64)
        polys = {}
        for g in (True, False):
            r.set_options(graphs=g)
            hits, _ = r.primary(cam, 96, 64, cull=False, tile=8)
            shadow, _ = r.point_shadows(cam, 96, 64, hits, sp["point_light"])
            polys[g] = (hits, shadow)
        for a, b in zip(polys[True], polys[False]):
            assert len(a) == len(b) and all(np.array_equal(a[f], b[f]) for f in ("x", "y", "n", "tri", "area"))
    finally:
        r.set_options(**saved)


def test_scenes_from_resident_files(require_gpu, tmp_path):
    """A context created from a scene file and a scene added from one render
    exactly as the same scenes uploaded from memory."""
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera
    paths = {}
    for name in ("room", "plant"):
        paths[name] = str(tmp_path / f"{name}.btgs")
        load_scene(name)[0].save(paths[name])
    a = GpuRenderer(load_scene("room")[0])
    pid = a.add_scene(load_scene("plant")[0])
    b = GpuRenderer(paths["room"])
    fid = b.add_scene(paths["plant"])
    assert b.scene_bytes(fid) == a.scene_bytes(pid) and b.scene_bytes(0) == a.scene_bytes(0)
    jobs = []
    for name, sid in (("room", 0), ("plant", pid)):
        sp = load_scene(name)[3]
        jobs.append({"scene": sid, "cam": camera(sp["view"], 48, 40), "W": 48, "H": 40, "light": sp["area_light"]})
    ra, _ = a.soft_batch(jobs, exact=True)
    jobs[1]["scene"] = fid
    rb, _ = b.soft_batch(jobs, exact=True)
    for x, y in zip(ra, rb):
        for f in ("tri", "vis", "E", "Eu"):
            assert np.array_equal(x[f], y[f]), f
    with pytest.raises(RuntimeError):
        b.add_scene(str(tmp_path / "missing.btgs"))
    a.close()
    b.close()


def test_largest_queue_capacity_gives_identical_images(require_gpu):
    """2^28 queue records (the benchmark's batch capacity on a 288 GB device:
    62 GB of queues, record offsets past 2^32 bytes) against 2^22 records (64x
    fewer roots in flight, many more rounds): identical images, bit for bit,
    on a batch whose roots x budget fill the large queues."""
    from beamtrace2007_amd.gpu import GpuRenderer, device_info
    from beamtrace2007_amd.host import camera
    if device_info()["mem_gb"] < 200:
        pytest.skip("needs a device with room for 62 GB of queues")
    s, _, _, sp = load_scene("plant")
    jobs = [{"scene": 0, "cam": camera(sp["view"], 256, 256), "W": 256, "H": 256,
             "light": scaled_light(sp["area_light"], k)} for k in (0.5, 1.0, 2.0, 4.0)]
    outs = {}
    for cap in (1 << 28, 1 << 22):
        r = GpuRenderer(s)
        r.set_options(capacity=cap)
        print("capacity", cap, flush=True)
        outs[cap], _ = r.soft_batch(jobs, exact=True)
        r.close()
    for a, b in zip(outs[1 << 28], outs[1 << 22]):
        for f in ("tri", "vis", "E", "Eu"):
            assert np.array_equal(a[f], b[f]), f
