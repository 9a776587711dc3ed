"""The benchmark configurations in their own form on the GPU, against the
oracle (which tests/test_oracle.py pins to the reference build on these very
scenes, tests/golden/large.npz):

* batched soft shadows (the 64-job workload's form): one persistent batch of
  noisy-blob and atrium jobs of uneven sides and light scales 0.25..4, every
  job against the oracle;
* the served stream's form: a request server on the conference hall with soft
  and point-shadow requests in flight together, every result against the
  oracle;
* resident scene files: a context created from a .btgs file against the
  oracle;
* point-light shadows on the full 2M-triangle building at 1024 x 1024 (16 x 16
  and 8 x 8 tiles): primary and shadow polygons of the whole frame, and the
  rasterized hit ids and shadow mask, bit for bit.

Bar: ids, polygons and masks bit-exact; visible fraction within 1e-5
absolute; exact irradiance within 1e-5 relative.
"""
import numpy as np
import pytest

from conftest import canonical_equal, check_soft, load_scene
from refpy import Oracle

pytestmark = pytest.mark.gpu

SERVER_HBM = 32 << 30
_oracles = {}


def oracle(name):
    if name not in _oracles:
        s, a, k, sp = load_scene(name)
        _oracles[name] = Oracle(a, k)
    return _oracles[name]


def test_batch_of_blob_and_atrium_jobs_matches_oracle(require_gpu):
    """The batched workload's form: scenes A (blob, 345K triangles) and B
    (atrium, 66K) resident side by side, eight jobs of sides 48..96 and light
    scales 0.25..4 in one persistent pass."""
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera
    from beamtrace2007_amd.workloads import scaled_light
    r = GpuRenderer(load_scene("blob")[0])
    ids = {"blob": 0, "atrium": r.add_scene(load_scene("atrium")[0])}
    plan = [("blob", 64, 64, 0.25), ("atrium", 96, 72, 0.25), ("blob", 48, 40, 1.0), ("atrium", 64, 48, 0.7),
            ("blob", 80, 56, 2.0), ("atrium", 48, 48, 1.5), ("blob", 56, 48, 4.0), ("atrium", 72, 56, 4.0)]
    jobs = []
    for name, W, H, sc in plan:
        sp = load_scene(name)[3]
        jobs.append({"name": name, "scene": ids[name], "cam": camera(sp["view"], W, H), "W": W, "H": H,
                     "light": scaled_light(sp["area_light"], sc)})
    print("batch of", len(jobs), "blob / atrium jobs", flush=True)
    outs, info = r.soft_batch(jobs, exact=True)
    r.close()
    print("roots traced again with a larger budget:", info["killed_roots"], flush=True)
    for j, g in zip(jobs, outs):
        print(f"oracle {j['name']} {j['W']}x{j['H']}", flush=True)
        o = oracle(j["name"]).soft(j["cam"], j["W"], j["H"], j["light"], exact=True, threads=16)
        check_soft(g, o)


def test_served_conference_requests_match_oracle(require_gpu):
    """The served stream's form: soft and point-shadow requests on the
    conference hall submitted together (soft ones join one running pass,
    point ones overlap on their own stream); every result against the
    oracle."""
    from beamtrace2007_amd.gpu import GpuRenderer, Server
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("conference")
    r = GpuRenderer(s)
    srv = Server(r, hbm_bytes=SERVER_HBM)
    plan = [("soft", 96, 96), ("point", 80, 80), ("soft", 64, 48), ("soft", 72, 72), ("point", 96, 64),
            ("soft", 48, 64), ("point", 64, 64), ("soft", 80, 60)]
    tickets = []
    for kind, W, H in plan:
        cam = camera(sp["view"], W, H)
        if kind == "soft":
            tickets.append((kind, cam, W, H, srv.submit_soft(0, cam, W, H, sp["area_light"])))
        else:
            tickets.append((kind, cam, W, H, srv.submit_point(cam, W, H, sp["point_light"])))
    print("submitted", len(tickets), "requests", flush=True)
    results = []
    for kind, cam, W, H, t in tickets:
        res = srv.wait(t, timeout=120)
        assert res is not None and res["status"] == 0, res
        results.append((kind, cam, W, H, res))
    srv.close()
    r.close()
    o = oracle("conference")
    for kind, cam, W, H, res in results:
        print(f"oracle {kind} {W}x{H}", flush=True)
        if kind == "soft":
            check_soft(res, o.soft(cam, W, H, sp["area_light"], exact=True, threads=16))
        else:
            op = o.primary(cam, W, H, tile=16)
            assert canonical_equal(res["hits"], op), (W, H)
            assert canonical_equal(res["polys"], o.point(cam, W, H, op, sp["point_light"])), (W, H)


def test_context_from_scene_file_matches_oracle(require_gpu, tmp_path):
    """A context created from a resident scene file (.btgs: the stored tree is
    uploaded, not rebuilt) renders soft shadows, primary visibility and point
    shadows as the oracle does on the scene the file was saved from."""
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera
    s, a, k, sp = load_scene("atrium")
    path = str(tmp_path / "atrium.btgs")
    s.save(path)
    r = GpuRenderer(path)
    W, H = 80, 60
    cam = camera(sp["view"], W, H)
    print("atrium from file", flush=True)
    check_soft(r.soft_shadows(cam, W, H, sp["area_light"], exact=True),
               oracle("atrium").soft(cam, W, H, sp["area_light"], exact=True, threads=16))
    gp, _ = r.primary(cam, W, H, tile=16)
    op = oracle("atrium").primary(cam, W, H, tile=16)
    assert canonical_equal(gp, op)
    gs, _ = r.point_shadows(cam, W, H, gp, sp["point_light"])
    assert canonical_equal(gs, oracle("atrium").point(cam, W, H, op, sp["point_light"]))
    r.close()


@pytest.mark.parametrize("tile", [16, 8])
def test_building_frame_at_full_size_matches_oracle(require_gpu, tile):
    """Point-light shadows at their full size: the 2M-triangle building at
    1024 x 1024, one root per tile. The whole frame's primary and shadow
    polygons equal the oracle's, and the ids and mask rasterized on the device
    (btg_images, polygons never leave the device) equal the host rasterizer's
    over the oracle's polygons."""
    from beamtrace2007_amd.gpu import GpuRenderer
    from beamtrace2007_amd.host import camera, rasterize
    s, a, k, sp = load_scene("building")
    W = H = 1024
    cam = camera(sp["view"], W, H)
    r = GpuRenderer(s)
    print(f"building {W}x{H} tile {tile}", flush=True)
    gp, ginfo = r.primary(cam, W, H, tile=tile)
    gs, _ = r.point_shadows(cam, W, H, gp, sp["point_light"])
    ids, mask, _ = r.images(cam, W, H, sp["point_light"], tile=tile)
    r.close()
    o = oracle("building")
    op = o.primary(cam, W, H, tile=tile)
    os_ = o.point(cam, W, H, op, sp["point_light"])
    print(f"oracle: {len(op['n'])} primary, {len(os_['n'])} shadow polygons", flush=True)
    assert canonical_equal(gp, op)
    assert canonical_equal(gs, os_)
    if ginfo["killed_roots"] == 0:  # counters are engine-independent when no root is traced twice
        assert ginfo["stats"]["triTests"] == op["stats"]["triTests"]
    assert np.array_equal(ids, rasterize(cam, W, H, op))
    assert np.array_equal(mask, rasterize(cam, W, H, os_, shadow=True))
