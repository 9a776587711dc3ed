"""Benchmark: exact soft shadows on one MI355X, batched vs plain, with
served-stream latency.

    python bench.py [--workload batch64|plant|room|serve|primary|building] [--steps K] [--warmup W]

batch64 (default): one step = one batch of 64 uneven exact soft-shadow jobs
on two resident scenes (a ~345K-triangle noisy blob and a ~66K-triangle
atrium, sides 128..1024, light scales 0.25..4, seed 64) traced in one
persistent wavefront pass, with the schedule chosen by the autotuner on a
scaled-down copy of the batch. `value` = batch pixels / wall time; the plain
form (one softShadows call per job) is timed on the same jobs beside it. The
same line carries a served mixed stream (`serving`: Poisson arrivals of soft
and point-shadow requests on a ~125K-triangle conference hall through the
request server, p50/p99 latency) and the reference CPU build on a bounded
sample of the batch.

plant / room: one step = one exact soft-shadow frame of one job (GPU
primary-ray pass + wavefront shadow beams, as the reference times it:
pixels / (primary_s + trace_s), scripts/make_report.py:171), with a parity
band against the oracle and the error against the double build.

primary / building: one step = one frame of primary visibility (the room at
512 x 512 from 8 x 8-pixel tile roots) or of primary visibility plus
point-light shadows (the 2M-triangle building at 1024 x 1024, 16 x 16 tile
roots), polygons traced and rasterized into hit ids (and the shadow mask) on
the device, the images downloaded; the reference CPU build renders the same
frame beside it and its images are compared bit for bit.

Scenes and trees are resident in HBM and soft-shadow images stay in device
memory (no PCIe in their timed region). The line also reports:
  roofline      the trace kernel's algorithmic bytes (one 12-B kd node per
                inner step and leaf visit, one 64-B leaf triangle per test,
                one 4-B parent link per climb) over its HIP-event time,
                against MI355X's 8 TB/s HBM3E;
  cpu_baseline  the reference's own CPU build (oracle/_ref, compiled from the
                reference sources) on a bounded sample of the same work, all
                host cores the process may use (affinity and CPU quota) and
                on one thread, its default options (mailbox on).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # name: (scene, default resolution, description)
    "plant": ("plant", 512, "exact soft shadows, seeded procedural plant (curved stem, 292 bent 3x3-grid leaves "
                            "2-6 cm, ground quad; 5378 tris), 0.5x0.5 m area light 1 m above, light-scale 1"),
    "room": ("room", 512, "exact soft shadows, seeded 10x8x3 m room (888 tris), 1x1 m ceiling light"),
    "primary": ("room", 512, "primary visibility, seeded 10x8x3 m room (888 tris), pinhole camera at the room "
                             "centre, 8x8-pixel tile roots, hit ids rasterized on the device"),
    "building": ("building", 1024, "primary visibility + point-light shadows, seeded 7-floor office building (12x8 "
                                   "rooms per floor, door and window cut-outs, desks and chairs; 2.03M tris), "
                                   "16x16-pixel tile roots, hit ids and shadow mask rasterized on the device"),
    "serve": (None, 0, "served mixed stream on the seeded conference hall (20x12 m, 8 tables, 120 chairs; 124800 "
                       "tris): Poisson arrivals (seed 5), 70% exact soft shadows / 30% point shadows, square side "
                       "uniform 256..1024"),
    "batch64": (None, 0, "exact soft shadows, 64 jobs drawn with seed 64: scene A (noisy level-7 icosphere blob with "
                         "limbs, 345106 tris) or B (two-storey atrium with arches, columns and cloth, 65840 tris) by "
                         "coin, square side uniform 128..1024, light scale log-uniform 0.25..4"),
}
METRIC = "exact soft-shadow pixels/s, batched vs plain (1 GPU); p50/p99 request latency"
# Measured HBM bytes per trace-kernel launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
# scaled by a 256 MiB copy: the *_hbm_traffic.txt summaries under profiles/) and
# the trace kernel's issue figures (rocprofv3 SQ counters: *_valu_busy.txt,
# *_sq_counters.txt), valid only for the library build and schedule they were
# measured with: keyed by the SHA-256 prefix of lib/libbtg.so and the schedule,
# so that any kernel change or another schedule reports null, never a stale figure.
TRAFFIC = {("0536bc727452f3d8", "plant", 512): 479315442}
TRAFFIC_BATCH = {("0536bc727452f3d8", 1 << 28, 1024, 8, 8, False): 5057140028,
                 ("0536bc727452f3d8", 1 << 28, 1024, 8, 4, False): 3980057483}
ISSUE_BATCH = {("0536bc727452f3d8", 1 << 28, 1024, 8, 8, False): {"valu_busy": 0.843, "lane_utilisation": 0.259},
               ("0536bc727452f3d8", 1 << 28, 1024, 8, 4, False): {"valu_busy": 0.845, "lane_utilisation": 0.263}}


def library_hash() -> str:
    import hashlib
    path = os.environ.get("BTG_LIBRARY") or os.path.join(ROOT, "beamtrace2007_amd", "lib", "libbtg.so")
    try:
        with open(path, "rb") as f:
            return hashlib.sha256(f.read()).hexdigest()[:16]
    except OSError:
        return ""


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: 2 for batch64, 10 otherwise)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default: 1 for batch64, 2 otherwise)")
    ap.add_argument("--workload", default="batch64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-serve", action="store_true", help="batch64: skip the served-stream measurement")
    ap.add_argument("--no-plain", action="store_true", help="batch64: skip the plain form (one call per job)")
    ap.add_argument("--no-tune", action="store_true",
                    help="batch64: skip the autotuner and keep the options given (for profiling the batched steps)")
    ap.add_argument("--res", type=int, default=0)
    ap.add_argument("--cpu-stride", type=int, default=1, help="plant / room: the CPU baseline traces every k-th row")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip per-launch HIP events")
    ap.add_argument("--budget", type=int, default=0, help="live sub-beams per root (0: library default)")
    ap.add_argument("--capacity", type=int, default=0,
                    help="sub-beam records per queue (0: batch64 sizes it from device memory, others use the "
                         "library default)")
    ap.add_argument("--leaf-cap", type=int, default=-1, help="leaves per sub-beam per trace launch (-1: default)")
    ap.add_argument("--mailbox", type=int, default=-1, choices=(-1, 0, 1),
                    help="GPU mailboxing of repeated triangle tests (-1: library default, off)")
    ap.add_argument("--qps", default="10,0",
                    help="serve: request arrival rates, comma-separated (0: every request at once, the saturation "
                         "capacity)")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests in the list")
    ap.add_argument("--fp64-requests", type=int, default=24,
                    help="serve: FP64 soft requests served in batches against the plain form (0: skip)")
    ap.add_argument("--no-plain-serve", dest="plain_serve", action="store_false",
                    help="serve: skip the plain form on the same request list")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 10 device-to-device copies of MIB MiB (the known-size copy that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and print their size")
    a = ap.parse_args()
    if a.steps is None:
        a.steps = 2 if a.workload == "batch64" else 10
    if a.warmup is None:
        a.warmup = 1 if a.workload == "batch64" else 2
    return a


def copy_calibration(mib: int) -> int:
    """A plain copy kernel of known size, for calibrating the HBM counters:
    profiled with rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE), each copy
    dispatch should read (write) exactly `mib` MiB."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(10):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"copy_bytes": n * 4, "copies": 10, "ok": ok}), flush=True)
    return 0 if ok else 1


def main() -> int:
    args = parse()
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the tracer shards by independent roots (pixels, tiles, jobs) and needs no collectives", file=sys.stderr)
        return 2
    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "serve":
        res = bench_serve(args)
        if res is None:
            return 1
        print(json.dumps(res), flush=True)
        return 0
    if args.workload in ("primary", "building"):
        return bench_visibility(args)
    from beamtrace2007_amd import gpu
    from beamtrace2007_amd.host import Scene, camera, camera_d

    scene_name, res0, desc = WORKLOADS[args.workload]
    res = args.res or res0
    W = H = res
    t0 = time.perf_counter()
    s = Scene.generate(scene_name)
    s.build_kd(0.4)
    build_s = time.perf_counter() - t0
    sp = s.spec()
    cam = camera(sp["view"], W, H)
    light = sp["area_light"]
    dev = gpu.device_info()
    if dev is None:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    r = gpu.GpuRenderer(s)
    kw = {"profile": not args.no_profile}
    if args.budget > 0:
        kw["budget"] = args.budget
    if args.capacity > 0:
        kw["capacity"] = args.capacity
    if args.leaf_cap >= 0:
        kw["leaf_cap"] = args.leaf_cap
    if args.mailbox >= 0:
        kw["mailbox"] = bool(args.mailbox)
    r.set_options(**kw)
    opts = r.options
    for _ in range(args.warmup):
        r.soft_shadows(cam, W, H, light, exact=True, download=False)
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        infos.append(r.soft_shadows(cam, W, H, light, exact=True, download=False)["info"])
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    px = W * H
    value = px * args.steps / wall

    st = infos[-1]["stats"]
    b_trace = 12 * (st["kdSteps"] + st["leafVisits"]) + 64 * st["triTests"] + 4 * st["climbs"]
    b_job = 40 * px + b_trace
    trace_s = sum(i["kernel_s"]["trace"] for i in infos)
    trace_launches = sum(i["launches"]["trace"] for i in infos)
    roofline = None
    if trace_s > 0:
        achieved = b_trace * args.steps / trace_s / 1e9
        defaults = args.budget <= 0 and args.capacity <= 0 and args.leaf_cap < 0
        traffic = TRAFFIC.get((library_hash(), args.workload, res)) if defaults else None
        roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                    "kernel": "traceKernel", "launches_per_step": trace_launches / args.steps,
                    "avg_launch_ms": round(trace_s / max(1, trace_launches) * 1e3, 4),
                    "alg_bytes_per_step": int(b_trace)}
    kernel_ms = {k: round(sum(i["kernel_s"][k] for i in infos) / args.steps * 1e3, 3) for k in gpu.KERNELS}

    # one downloaded frame for the parity band and error against the double build
    t1 = time.perf_counter()
    full = r.soft_shadows(cam, W, H, light, exact=True, download=True)
    pcie_s = time.perf_counter() - t1
    r.set_options(precision=1)
    guarded = r.soft_shadows(cam, W, H, light, exact=True, download=True)
    # FP64 mode from the double build's camera (one frame, images downloaded)
    r.set_options(precision=2)
    t1 = time.perf_counter()
    fp64 = r.soft_shadows(cam, W, H, light, exact=True, download=True, cam_d=camera_d(sp["view"], W, H))
    fp64_s = time.perf_counter() - t1
    r.close()
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "pixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 exact integral)", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"{args.workload} {W}x{H}: {desc}", "mode": "plain (one job per call)",
                   "tris": s.num_tris, "kd_nodes": len(s.kd_arrays()["nodes"]), "exact": True,
                   "budget": opts["budget"], "capacity": opts["capacity"], "leaf_cap": opts["leaf_cap"]},
        "device": device_label(dev), "arch": dev["arch"],
        "beams_per_s": round(st["beams"] * args.steps / wall, 1),
        "rounds": infos[-1]["rounds"], "killed_roots": infos[-1]["killed_roots"],
        "kernel_ms_per_step": kernel_ms, "alg_bytes_per_step": int(b_job),
        "pcie_inclusive_px_per_s": round(px / pcie_s, 1),  # one frame with tri/vis/E/Eu copied to pageable host memory
        "roofline": roofline, "host_kd_build_s": round(build_s, 3),
        "fp64_mode": {"pixels_per_s": round(px / fp64_s, 1), "seconds": round(fp64_s, 3),
                      "mode": "precision 2 from the double build's camera, one frame, images downloaded"},
    }
    if any(infos[-1]["phase_cycles"].values()):  # instrumented builds only
        out["trace_phase_cycles"] = infos[-1]["phase_cycles"]
    out.update(cpu_and_parity(args, s, sp, cam, W, H, light, full, guarded, fp64))
    print(json.dumps(out), flush=True)
    return 0


def trace_bytes(st: dict) -> int:
    """Algorithmic bytes of the trace kernel for a pass's counters: a 12-B kd
    node per inner step and leaf visit, a 64-B leaf triangle per test, a 4-B
    parent link per climbed level."""
    return 12 * (st["kdSteps"] + st["leafVisits"]) + 64 * st["triTests"] + 4 * st["climbs"]


def bench_visibility(args) -> int:
    """Primary visibility (room, 8 x 8 tile roots) or primary visibility plus
    point-light shadows (the 2M-triangle building at 1024 x 1024, 16 x 16 tile
    roots): one step = one frame's polygons traced on the device and
    rasterized there into per-pixel hit ids (and the shadow mask), the images
    then downloaded (5 B per pixel). The reference CPU build renders the same
    frame with the same algorithm (restart trail, no mailbox: its primary and
    point-shadow paths are single-threaded, render.cpp:150-158, 209-232) as
    the baseline, and its ids and mask are compared with the GPU's."""
    from beamtrace2007_amd import gpu
    from beamtrace2007_amd.host import Scene, camera

    scene_name, res0, desc = WORKLOADS[args.workload]
    W = H = args.res or res0
    tile = 8 if args.workload == "primary" else 16
    dev = gpu.device_info()
    if dev is None:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    t0 = time.perf_counter()
    s = Scene.generate(scene_name)
    s.build_kd(0.4)
    build_s = time.perf_counter() - t0
    sp = s.spec()
    cam = camera(sp["view"], W, H)
    light3 = sp["point_light"] if args.workload == "building" else None
    r = gpu.GpuRenderer(s)
    kw = {}
    if args.leaf_cap >= 0:
        kw["leaf_cap"] = args.leaf_cap
    if args.mailbox >= 0:
        kw["mailbox"] = bool(args.mailbox)
    r.set_options(**kw)
    # The timed frames run as a caller runs them (round groups replayed from
    # HIP graphs); per-launch HIP events would add host work to every launch of
    # these latency-bound passes, so the kernel times come from extra frames
    # profiled after the timed ones.
    for _ in range(args.warmup):
        r.images(cam, W, H, light3, tile=tile)
    infos = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ids, mask, info = r.images(cam, W, H, light3, tile=tile)
        infos.append(info)
    wall = time.perf_counter() - t0
    px = W * H
    print(f"bench.py: {args.workload} {W}x{H}: {args.steps} frames in {wall:.3f} s", file=sys.stderr, flush=True)
    passes = ("primary", "point") if light3 is not None else ("primary",)
    prof = []
    if not args.no_profile:
        r.set_options(profile=True)
        for _ in range(max(1, min(args.steps, 3))):
            prof.append(r.images(cam, W, H, light3, tile=tile)[2])
        r.set_options(profile=False)
    r.close()
    b_trace = sum(trace_bytes(infos[-1][k]["stats"]) for k in passes)
    roofline, kernel_ms = None, None
    if prof:
        n_prof = len(prof)
        trace_s = sum(i[k]["kernel_s"]["trace"] for i in prof for k in passes)
        launches = sum(i[k]["launches"]["trace"] for i in prof for k in passes)
        if trace_s > 0:
            achieved = b_trace * n_prof / trace_s / 1e9
            roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "kernel": "traceKernel",
                        "launches_per_step": launches / n_prof,
                        "avg_launch_ms": round(trace_s / max(1, launches) * 1e3, 4),
                        "alg_bytes_per_step": int(b_trace),
                        "timed_on": f"{n_prof} profiled frames after the timed ones (per-launch HIP events)"}
        kernel_ms = {k: {n: round(sum(i[k]["kernel_s"][n] for i in prof) / n_prof * 1e3, 3) for n in gpu.KERNELS}
                     for k in passes}
    per_pass = {k: {"seconds": round(sum(i[k]["total_s"] for i in infos) / args.steps, 5),
                    "rounds": infos[-1][k]["rounds"], "killed_roots": infos[-1][k]["killed_roots"],
                    "beams": infos[-1][k]["stats"]["beams"],
                    "kernel_ms_profiled": kernel_ms[k] if kernel_ms else None} for k in passes}
    beams = sum(infos[-1][k]["stats"]["beams"] for k in passes)
    out = {
        "metric": "primary visibility" + (" + point-shadow" if light3 is not None else "") + " pixels/s (1 GPU)",
        "value": round(px * args.steps / wall, 1), "unit": "pixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
        "config": {"workload": f"{args.workload} {W}x{H}: {desc}", "tile": tile, "tris": s.num_tris,
                   "kd_nodes": len(s.kd_arrays()["nodes"])},
        "device": device_label(dev), "arch": dev["arch"], "beams_per_s": round(beams * args.steps / wall, 1),
        "passes": per_pass, "raster_ms": round(sum(i["raster_s"] for i in infos) / args.steps * 1e3, 3),
        "roofline": roofline, "host_kd_build_s": round(build_s, 3), "cpu_baseline": None, "parity": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from refpy import Ref, ref_available
        if ref_available("f"):
            a = s.arrays()
            ref = Ref("f").load(a["pos"], a["tris"])
            t0 = time.perf_counter()
            prim = ref.primary(sp["view"], W, H, mailbox=False, trail=True, tile=tile)
            rid = ref.raster(sp["view"], W, H, prim)
            rmask = None
            if light3 is not None:
                sh = ref.point(sp["view"], W, H, prim, light3, mailbox=False, trail=True)
                rmask = ref.raster(sp["view"], W, H, sh, shadow=True)
            secs = time.perf_counter() - t0
            ref.free()
            _, cores = host_cores()
            out["cpu_baseline"] = {"value": round(px / secs, 1), "unit": "pixels/s", "cores": 1, "kind": "reference",
                                   "host": cores,
                                   "sample": f"the whole {W}x{H} frame on one thread ({secs:.2f} s: primary beams"
                                             + (" + point-shadow beams" if light3 is not None else "")
                                             + " + rasterization, restart trail, no mailbox)"}
            out["parity"] = {"against": "the reference CPU build, same frame",
                             "ids_bitexact": bool(np.array_equal(ids, rid)),
                             "mask_bitexact": None if rmask is None else bool(np.array_equal(mask, rmask))}
    print(json.dumps(out), flush=True)
    return 0


def device_label(dev: dict) -> str:
    """The device's marketing name, or (where the runtime reports none) its
    compute units, memory and architecture."""
    return dev["name"] or f"{dev['compute_units']}-CU {dev['mem_gb']:.0f} GiB {dev['arch'].split(':')[0]}"


def batch_capacity(mem_gb: float) -> int:
    """Queue records for both forms of the 64-job batch: the largest power of
    two up to 2^28 whose three sub-beam queues and shade queue (232 B per
    record) fit in 22% of device memory (2^28, 62 GB, on a 288 GB MI355X; the
    library's own default takes 11%, so that several contexts fit). More roots
    in flight per round, fewer rounds: the batch ran 3.2% faster at 2^28 than
    at 2^27."""
    fit = mem_gb * 2**30 * 0.22 / 232
    cap = 1 << 20
    while cap < (1 << 28) and cap * 2 <= fit:
        cap *= 2
    return cap


def bench_batch(args) -> int:
    """The batched form: 64 uneven jobs on two resident scenes in one
    persistent pass per step, against the plain form (one softShadows call per
    job, one context per scene) on the same jobs."""
    from beamtrace2007_amd import gpu
    from beamtrace2007_amd.host import Scene, camera
    from beamtrace2007_amd.workloads import batch_draw, scaled_light

    dev = gpu.device_info()
    if dev is None:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    t0 = time.perf_counter()
    scenes = {}
    for name in ("blob", "atrium"):
        sc = Scene.generate(name)
        sc.build_kd(0.4)
        scenes[name] = sc
    build_s = time.perf_counter() - t0
    draw = batch_draw(64, seed=64)
    r = gpu.GpuRenderer(scenes["blob"])
    ids = {"blob": 0, "atrium": r.add_scene(scenes["atrium"])}
    kw = {"profile": not args.no_profile,
          "capacity": args.capacity if args.capacity > 0 else batch_capacity(dev["mem_gb"])}
    if args.budget > 0:
        kw["budget"] = args.budget
    if args.leaf_cap >= 0:
        kw["leaf_cap"] = args.leaf_cap
    if args.mailbox >= 0:
        kw["mailbox"] = bool(args.mailbox)
    r.set_options(**kw)
    jobs = []
    for d in draw:
        sp = scenes[d["scene"]].spec()
        jobs.append({"name": d["scene"], "scene": ids[d["scene"]], "cam": camera(sp["view"], d["side"], d["side"]),
                     "W": d["side"], "H": d["side"], "light": scaled_light(sp["area_light"], d["light_scale"])})
    px = sum(j["W"] * j["H"] for j in jobs)
    # warm-up: one small job per scene in both forms
    small = [dict(j, W=64, H=64, cam=camera(scenes[j["name"]].spec()["view"], 64, 64)) for j in
             [next(j for j in jobs if j["name"] == n) for n in ("blob", "atrium")]]
    # the plain form runs as a caller runs it: no per-launch HIP events, so its
    # round groups replay from HIP graphs (the library default)
    plain = {n: gpu.GpuRenderer(scenes[n]) for n in scenes}
    for p_ in plain.values():
        p_.set_options(**dict(kw, profile=False))
    for _ in range(max(1, args.warmup)):
        r.soft_batch(small, exact=True, download=False)
        for j in small:
            plain[j["name"]].soft_shadows(j["cam"], j["W"], j["H"], j["light"], exact=True, download=False)
    # the schedule for the batched form comes from the autotuner, timed on the
    # same 64 jobs at a quarter of their side (outside the timed region)
    sample = [dict(j, W=max(32, j["W"] // 4), H=max(32, j["H"] // 4),
                   cam=camera(scenes[j["name"]].spec()["view"], max(32, j["W"] // 4), max(32, j["H"] // 4)))
              for j in jobs]
    t1 = time.perf_counter()
    if args.no_tune:
        o = r.options
        tuned, trials = {k: o[k] for k in ("budget", "rounds_per_check", "leaf_cap", "mailbox")}, []
    else:
        tuned, trials = r.autotune(sample, exact=True, max_seconds=15.0)
    tune_s = time.perf_counter() - t1
    print(f"bench.py: autotuned in {tune_s:.1f} s: {tuned}", file=sys.stderr, flush=True)
    r.set_options(**tuned)
    # plain form: one call per job, the same queue capacity and knobs before tuning
    plain_s = None
    if not args.no_plain:
        t0 = time.perf_counter()
        for j in jobs:
            plain[j["name"]].soft_shadows(j["cam"], j["W"], j["H"], j["light"], exact=True, download=False)
        plain_s = time.perf_counter() - t0
        print(f"bench.py: plain form {plain_s:.2f} s", file=sys.stderr, flush=True)
    for p_ in plain.values():
        p_.close()
    # batched form: the timed steps
    infos = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        infos.append(r.soft_batch(jobs, exact=True, download=False)[1])
        print(f"bench.py: batched step {k + 1}: {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    capacity = r.options["capacity"]
    mailbox = r.options["mailbox"]
    precision = None if args.no_cpu else batch_precision(r, scenes, jobs, px)
    r.close()
    st = infos[-1]["stats"]
    b_trace = 12 * (st["kdSteps"] + st["leafVisits"]) + 64 * st["triTests"] + 4 * st["climbs"]
    trace_s = sum(i["kernel_s"]["trace"] for i in infos)
    trace_launches = sum(i["launches"]["trace"] for i in infos)
    roofline = None
    if trace_s > 0:
        achieved = b_trace * args.steps / trace_s / 1e9
        traffic = TRAFFIC_BATCH.get((library_hash(), capacity, tuned["budget"], tuned["rounds_per_check"],
                                     tuned["leaf_cap"], mailbox))
        roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "kernel": "traceKernel",
                    "launches_per_step": trace_launches / args.steps,
                    "avg_launch_ms": round(trace_s / max(1, trace_launches) * 1e3, 4), "alg_bytes_per_step": int(b_trace)}
    value = px * args.steps / wall
    sched = (library_hash(), capacity, tuned["budget"], tuned["rounds_per_check"], tuned["leaf_cap"], mailbox)
    issue = None if sched not in ISSUE_BATCH else dict(
        ISSUE_BATCH[sched], kernel="traceKernel",
        meaning="fraction of SIMD cycles issuing a VALU instruction, and of lanes active per VALU instruction "
                "(rocprofv3 PMC, this schedule)")
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "pixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 exact integral)",
        "data": "synthetic (seeded procedural scenes, seeded job draw)",
        "config": {"workload": f"batch64: {WORKLOADS['batch64'][2]}", "mode": "batched (64 jobs per call)",
                   "jobs": len(jobs), "pixels_per_step": px, "tris": {n: scenes[n].num_tris for n in scenes},
                   "draw_seed": 64, "scene_seed": 2007},
        "device": device_label(dev), "arch": dev["arch"],
        "schedule": dict(tuned, capacity=capacity, mailbox=mailbox), "library_sha256_16": library_hash(),
        "plain": None if plain_s is None else {"value": round(px / plain_s, 1), "unit": "pixels/s",
                                               "seconds": round(plain_s, 3), "form": "one softShadows call per job"},
        "batched_over_plain": None if plain_s is None else round(plain_s / (wall / args.steps), 3),
        "beams_per_s": round(st["beams"] * args.steps / wall, 1),
        "rounds": infos[-1]["rounds"], "killed_roots": infos[-1]["killed_roots"],
        "retry_passes": infos[-1]["retry_passes"],
        "kernel_ms_per_step": {k: round(sum(i["kernel_s"][k] for i in infos) / args.steps * 1e3, 3)
                               for k in gpu.KERNELS},
        "roofline": roofline, "issue_bound": issue, "host_kd_build_s": round(build_s, 3),
        "cpu_baseline": None if args.no_cpu else cpu_batch_sample(scenes, jobs, draw),
        "precision": precision,
        "autotune": None if args.no_tune else {"chosen": tuned, "seconds": round(tune_s, 2), "trials": trials,
                                               "sample": "the 64 jobs at a quarter of their side"},
        "draw": [[d["scene"][0], d["side"], round(d["light_scale"], 4)] for d in draw],
    }
    if not args.no_serve:
        out["serving"] = bench_serve(args)
    print(json.dumps(out), flush=True)
    return 0


def batch_precision(r, scenes, jobs, px, side: int = 128) -> dict | None:
    """The FP64-guarded mode's batched throughput on the full batch, and every
    mode's error against the reference's double build on a sample: the first
    two jobs of each scene (their scene, view and light) re-rendered at
    side x side and compared on whole frames (modes 0 and 1 from the float
    camera, mode 2 from the double build's camera, all three batched)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refpy import Ref, ref_available
    from beamtrace2007_amd.host import camera, camera_d

    saved = r.options
    r.set_options(precision=1)
    t0 = time.perf_counter()
    r.soft_batch(jobs, exact=True, download=False)
    guarded_s = time.perf_counter() - t0
    r.set_options(**saved)
    print(f"bench.py: guarded batch {guarded_s:.2f} s", file=sys.stderr, flush=True)
    out = {"guarded": {"value": round(px / guarded_s, 1), "unit": "pixels/s", "seconds": round(guarded_s, 3),
                       "mode": "precision 1: double hit point, the double build's apex offset "
                               "carried as a float pair, FP64 Lambert integral"}}
    if not ref_available("d"):
        return out
    names = sorted({j["name"] for j in jobs})
    sample = []
    for n in names:
        for j in [x for x in jobs if x["name"] == n][:2]:
            view = scenes[n].spec()["view"]
            sample.append(dict(j, W=side, H=side, cam=camera(view, side, side), cam_d=camera_d(view, side, side)))
    res, secs = {}, {}
    for mode in (0, 1, 2):
        r.set_options(precision=mode)
        t0 = time.perf_counter()
        res[mode] = r.soft_batch(sample, exact=True, download=True)[0]
        secs[mode] = time.perf_counter() - t0
    r.set_options(**saved)
    threads, _ = host_cores()
    refs, dimg = {}, []
    t0 = time.perf_counter()
    for j in sample:
        name = j["name"]
        if name not in refs:
            a = scenes[name].arrays()
            refs[name] = Ref("d").load(a["pos"], a["tris"])
        dimg.append(refs[name].soft(scenes[name].spec()["view"], side, side, j["light"], exact=True, mailbox=False,
                                    trail=True, threads=threads))
    d_s = time.perf_counter() - t0
    for rf in refs.values():
        rf.free()
    cat = {k: np.concatenate([im[k].ravel() for im in dimg]) for k in ("tri", "vis", "E")}
    stated = {}
    for mode, key in ((0, "0_float_build"), (1, "1_guarded"), (2, "2_fp64")):
        stated[key] = err_vs_fp64({k: np.concatenate([o[k].ravel() for o in res[mode]]) for k in ("tri", "vis", "E")},
                                  cat)
        stated[key]["sample_px_per_s"] = round(len(sample) * side * side / secs[mode], 1)
    stated["against"] = (f"bt_render_d's algorithm (the reference compiled with Real = double) on whole frames: "
                         f"the first two jobs of each scene ({', '.join(names)}) at {side}x{side}, their lights and "
                         f"views; {d_s:.1f} s on {threads} threads")
    out["precision"] = stated
    return out


def host_cores() -> tuple[int, dict]:
    """Threads for the CPU baseline: every core this process may run on (its
    affinity mask, further limited by a cgroup CPU quota where one is set),
    with what the machine reports beside it."""
    n_os = os.cpu_count() or 1
    try:
        n_aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n_aff = n_os
    quota = None
    for path in ("/sys/fs/cgroup/cpu.max", "/sys/fs/cgroup/cpu/cpu.cfs_quota_us"):
        try:
            with open(path) as f:
                parts = f.read().split()
            if path.endswith("cpu.max") and parts and parts[0] != "max":
                quota = float(parts[0]) / float(parts[1])
            elif path.endswith("quota_us") and parts and int(parts[0]) > 0:
                with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as g:
                    quota = int(parts[0]) / int(g.read().split()[0])
            break
        except (OSError, ValueError, IndexError, ZeroDivisionError):
            continue
    n = n_aff if quota is None else max(1, min(n_aff, int(quota + 0.999)))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, {"os_cpu_count": n_os, "affinity": n_aff, "cgroup_quota": quota, "model": model}


def cpu_batch_sample(scenes, jobs, draw, budget_s: float = 20.0, single_s: float = 8.0) -> dict | None:
    """The reference CPU build on a bounded sample of the batch, on every
    host core and on one: each job traces every k-th row of its image (k set
    so that a call has several rows per thread, rows spread over the image's
    height), jobs in draw order until about budget_s seconds of tracing; the
    one-thread figure traces a subset of the same rows (every 4th of the
    sampled rows of the first jobs) for about single_s seconds. Rows are
    handed to the reference's own row workers (src/render/render.cpp:386-421
    hands out rows the same way)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refpy import Ref, ref_available

    if not ref_available("f"):
        return None
    threads, cores = host_cores()
    refs = {}

    def ref_of(name):
        if name not in refs:
            a = scenes[name].arrays()
            refs[name] = Ref("f").load(a["pos"], a["tris"])
        return refs[name]

    def run(nthreads, rows_per_job, budget):
        px = rows = 0
        secs = 0.0
        njobs = 0
        for j in jobs:
            sp = scenes[j["name"]].spec()
            stride = max(1, j["H"] // rows_per_job)
            y0 = stride // 2
            res = ref_of(j["name"]).soft_rows(sp["view"], j["W"], j["H"], j["light"], y0, j["H"], exact=True,
                                              mailbox=True, trail=False, threads=nthreads, stride=stride)
            secs += float(res["seconds"][0] + res["seconds"][1])
            n = res["vis"].shape[0]
            rows += n
            px += n * j["W"]
            njobs += 1
            if secs >= budget:
                break
        return px, rows, secs, njobs

    px, rows, secs, njobs = run(threads, max(8, 4 * threads), budget_s)
    px1, rows1, secs1, njobs1 = run(1, 4, single_s)
    for rf in refs.values():
        rf.free()
    allc, one = px / secs, px1 / secs1
    print(f"bench.py: CPU reference {threads} threads: {rows} rows of {njobs} jobs, {secs:.1f} s, {allc:.0f} px/s; "
          f"1 thread: {rows1} rows of {njobs1} jobs, {secs1:.1f} s, {one:.0f} px/s ({cores['model']})",
          file=sys.stderr, flush=True)
    return {"value": round(allc, 1), "unit": "pixels/s", "cores": threads, "kind": "reference",
            "single_thread": {"value": round(one, 1), "rows": rows1, "jobs": njobs1, "seconds": round(secs1, 2)},
            "speedup_over_one_thread": round(allc / one, 2), "host": cores,
            "sample": f"about {max(8, 4 * threads)} evenly spaced rows of each of the first {njobs} jobs of the batch "
                      f"({rows} rows, {px} px, {secs:.2f} s on {threads} threads: primary rays + shadow beams, "
                      "mailbox on, frame stack, the reference's row workers)"}


def bench_serve(args) -> dict | None:
    """The served mixed stream on the conference hall through the request
    server, and the plain form on the same request list.

    One seeded request list (args.requests requests: 70% exact soft shadows,
    30% point shadows, sides uniform 256..1024) is served at each rate of
    args.qps: a rate > 0 is a Poisson arrival stream at that rate (latency
    submit -> results in host memory, p50/p99), 0 submits every request at
    once (saturation: the server's capacity in pixels/s). The plain form then
    renders the same list one call at a time on the same context (soft:
    softShadows with the images downloaded; point: primary + point shadows
    with the polygon lists downloaded, as a served point request returns
    them)."""
    import threading

    from beamtrace2007_amd import gpu
    from beamtrace2007_amd.host import Scene, camera
    from beamtrace2007_amd.workloads import served_draw

    if gpu.device_info() is None:
        print("bench.py: no HIP device", file=sys.stderr)
        return None
    sc = Scene.generate("conference")
    sc.build_kd(0.4)
    sp = sc.spec()
    r = gpu.GpuRenderer(sc)
    rates = [float(q) for q in str(args.qps).split(",") if q.strip()]
    base = served_draw(args.requests, 1.0, seed=5)  # kinds and sides do not depend on the rate
    px_of = [d["side"] ** 2 for d in base]
    total_px = sum(px_of)

    def pct(v, q):
        return round(float(np.percentile(np.array(v), q)) * 1e3, 2) if len(v) else None

    kept = []  # served soft results kept for their error against the double build

    def serve(rate):
        srv = gpu.Server(r)
        for t in (srv.submit_soft(0, camera(sp["view"], 64, 64), 64, 64, sp["area_light"]),
                  srv.submit_point(camera(sp["view"], 64, 64), 64, 64, sp["point_light"])):
            srv.wait(t)  # warm-up outside the measurement
        arrivals = [d["t"] / rate for d in base] if rate > 0 else [0.0] * len(base)
        tickets = [None] * len(base)
        t0 = time.perf_counter()

        def submitter():
            for i, d in enumerate(base):
                delay = t0 + arrivals[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                cam = camera(sp["view"], d["side"], d["side"])
                if d["kind"] == "soft":
                    tickets[i] = srv.submit_soft(0, cam, d["side"], d["side"], sp["area_light"])
                else:
                    tickets[i] = srv.submit_point(cam, d["side"], d["side"], sp["point_light"])

        th = threading.Thread(target=submitter)
        th.start()
        lat = {"soft": [], "point": []}
        queue = []
        last = time.perf_counter()
        for i, d in enumerate(base):
            while tickets[i] is None:
                time.sleep(0.0005)
            res = srv.wait(tickets[i])
            if res["status"] != 0:
                print(f"bench.py: request {i} failed: {res.get('error')}", file=sys.stderr)
                th.join()
                srv.close()
                return None
            lat[d["kind"]].append(res["latency_s"])
            queue.append(res["queue_s"])
            if d["kind"] == "soft" and len(kept) < 4:
                kept.append((d["side"], res))
            if time.perf_counter() - last > 10:
                print(f"bench.py: rate {rate:g}: served {i + 1}/{len(base)}", file=sys.stderr, flush=True)
                last = time.perf_counter()
        th.join()
        span = time.perf_counter() - t0
        srv.close()
        allv = lat["soft"] + lat["point"]
        print(f"bench.py: rate {rate:g}: {len(base)} requests in {span:.1f} s, {total_px / span:.0f} px/s, p50 "
              f"{pct(allv, 50)} ms, p99 {pct(allv, 99)} ms", file=sys.stderr, flush=True)
        return {"qps": rate if rate > 0 else "saturation", "seconds": round(span, 3),
                "pixels_per_s": round(total_px / span, 1),
                "offered_pixels_per_s": round(total_px / arrivals[-1], 1) if rate > 0 else None,
                "latency_ms": {"p50": pct(allv, 50), "p99": pct(allv, 99), "soft_p50": pct(lat["soft"], 50),
                               "soft_p99": pct(lat["soft"], 99), "point_p50": pct(lat["point"], 50),
                               "point_p99": pct(lat["point"], 99)},
                "queue_ms": {"p50": pct(queue, 50), "p99": pct(queue, 99)}}

    points = []
    for rate in rates:
        p = serve(rate)
        if p is None:
            r.close()
            return None
        points.append(p)
    plain = None
    if args.plain_serve:
        for d in base[:2]:  # warm-up
            cam = camera(sp["view"], 64, 64)
            r.soft_shadows(cam, 64, 64, sp["area_light"], exact=True)
        lat = []
        t0 = time.perf_counter()
        for d in base:
            t1 = time.perf_counter()
            cam = camera(sp["view"], d["side"], d["side"])
            if d["kind"] == "soft":
                r.soft_shadows(cam, d["side"], d["side"], sp["area_light"], exact=True, download=True)
            else:
                hits, _ = r.primary(cam, d["side"], d["side"], tile=16)
                r.point_shadows(cam, d["side"], d["side"], hits, sp["point_light"])
            lat.append(time.perf_counter() - t1)
        span = time.perf_counter() - t0
        print(f"bench.py: plain form on the same requests {span:.1f} s, {total_px / span:.0f} px/s",
              file=sys.stderr, flush=True)
        plain = {"seconds": round(span, 3), "pixels_per_s": round(total_px / span, 1),
                 "call_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                 "form": "one call per request, in arrival order (soft: softShadows, images downloaded; point: "
                         "primary + pointShadows, polygon lists downloaded)"}
    r.close()
    sat = next((p for p in points if p["qps"] == "saturation"), None)
    served_err = None if args.no_cpu else served_precision(sc, sp, kept)
    fp64 = serve_fp64(args, sc, sp) if args.fp64_requests > 0 else None
    return {"requests": len(base), "arrival_seed": 5, "pixels": total_px, "precision": served_err,
            "mix": {"soft": sum(d["kind"] == "soft" for d in base), "point": sum(d["kind"] == "point" for d in base)},
            "points": points, "plain": plain,
            "served_over_plain": round(sat["pixels_per_s"] / plain["pixels_per_s"], 3) if sat and plain else None,
            "fp64": fp64, "config": WORKLOADS["serve"][2]}


def serve_fp64(args, sc, sp) -> dict | None:
    """FP64 soft shadows (precision 2, each request from its double camera)
    served in batches on the conference hall, against the plain form (one
    call per request) on the same seeded list: args.fp64_requests requests,
    sides uniform 64..128 (seed 7), all submitted at once; the error of the
    served results against the reference's double build on two of them."""
    from beamtrace2007_amd import gpu
    from beamtrace2007_amd.host import camera, camera_d

    g = np.random.Generator(np.random.PCG64(7))
    sides = [int(x) for x in g.integers(64, 129, size=args.fp64_requests)]
    total_px = sum(x * x for x in sides)
    r = gpu.GpuRenderer(sc)
    r.set_options(precision=2)
    r.soft_shadows(camera(sp["view"], 32, 32), 32, 32, sp["area_light"], exact=True,
                   cam_d=camera_d(sp["view"], 32, 32))  # warm-up
    srv = gpu.Server(r)
    t0 = time.perf_counter()
    tickets = [srv.submit_soft(0, camera(sp["view"], x, x), x, x, sp["area_light"], cam_d=camera_d(sp["view"], x, x))
               for x in sides]
    res, lat = [], []
    for t in tickets:
        q = srv.wait(t)
        if q["status"] != 0:
            print(f"bench.py: FP64 request failed: {q.get('error')}", file=sys.stderr)
            srv.close()
            r.close()
            return None
        res.append(q)
        lat.append(q["latency_s"])
    served_s = time.perf_counter() - t0
    srv.close()
    t0 = time.perf_counter()
    for x in sides:
        r.soft_shadows(camera(sp["view"], x, x), x, x, sp["area_light"], exact=True, cam_d=camera_d(sp["view"], x, x))
    plain_s = time.perf_counter() - t0
    r.close()
    print(f"bench.py: FP64 served {args.fp64_requests} requests in {served_s:.2f} s, plain {plain_s:.2f} s",
          file=sys.stderr, flush=True)
    err = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from refpy import Ref, ref_available
        if ref_available("d"):
            threads, _ = host_cores()
            a = sc.arrays()
            refd = Ref("d").load(a["pos"], a["tris"])
            imgs, refs = [], []
            for x, q in list(zip(sides, res))[:2]:
                d = refd.soft(sp["view"], x, x, sp["area_light"], exact=True, mailbox=False, trail=True, threads=threads)
                imgs.append({k: q[k].ravel() for k in ("tri", "vis", "E")})
                refs.append({k: d[k].ravel() for k in ("tri", "vis", "E")})
            refd.free()
            cat = lambda L: {k: np.concatenate([z[k] for z in L]) for k in ("tri", "vis", "E")}  # noqa: E731
            err = err_vs_fp64(cat(imgs), cat(refs))
            err["sample"] = f"the first two served FP64 requests (sides {sides[0]}, {sides[1]}), whole frames"
    return {"requests": len(sides), "pixels": total_px, "sides": "uniform 64..128, seed 7",
            "served": {"seconds": round(served_s, 3), "pixels_per_s": round(total_px / served_s, 1),
                       "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                                      "p99": round(float(np.percentile(lat, 99)) * 1e3, 1)},
                       "form": "all submitted at once to a server on a precision-2 context (FP64 batches)"},
            "plain": {"seconds": round(plain_s, 3), "pixels_per_s": round(total_px / plain_s, 1),
                      "form": "one FP64 softShadows call per request"},
            "served_over_plain": round(plain_s / served_s, 3), "precision": err}


def err_vs_fp64(img: dict, d: dict) -> dict:
    """Visible fraction and exact irradiance of `img` against the double
    build's image `d`, on the pixels where both have the same primary hit."""
    m = (d["vis"] >= 0) & (img["vis"] >= 0) & (img["tri"] == d["tri"])
    err = np.abs(img["vis"].astype(np.float64) - d["vis"])[m]
    e_err = np.abs(img["E"].astype(np.float64) - d["E"])[m] / max(float(np.abs(d["E"][m]).max()) if m.any() else 0.0,
                                                                   1e-30)
    if not err.size:
        return {"pixels": 0}
    return {"max": float(err.max()), "mean": float(err.mean()), "frac_above_1e-5": float(np.mean(err > 1e-5)),
            "pixels": int(err.size), "same_hit_frac": round(float(m.sum()) / max(int((d["vis"] >= 0).sum()), 1), 6),
            "E_max_rel": float(e_err.max())}


def served_precision(sc, sp, kept, rows: int = 8) -> dict | None:
    """Served soft results (the server runs the float build's mode) against
    the reference's double build: `rows` middle rows of each kept result."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refpy import Ref, ref_available
    if not ref_available("d") or not kept:
        return None
    threads, _ = host_cores()
    a = sc.arrays()
    refd = Ref("d").load(a["pos"], a["tris"])
    imgs, refs = [], []
    for side, res in kept:
        y0 = side // 2 - rows // 2
        rr = refd.soft_rows(sp["view"], side, side, sp["area_light"], y0, y0 + rows, exact=True, mailbox=False,
                            trail=True, threads=threads)
        imgs.append({k: res[k][y0:y0 + rows].ravel() for k in ("tri", "vis", "E")})
        refs.append({k: rr[k].ravel() for k in ("tri", "vis", "E")})
    refd.free()
    cat = lambda L: {k: np.concatenate([x[k] for x in L]) for k in ("tri", "vis", "E")}  # noqa: E731
    out = err_vs_fp64(cat(imgs), cat(refs))
    out["sample"] = (f"{rows} middle rows of the first {len(kept)} served soft requests (sides "
                     f"{', '.join(str(k[0]) for k in kept)}), precision 0, against bt_render_d's algorithm")
    return out


def cpu_and_parity(args, s, sp, cam, W, H, light, full, guarded=None, fp64=None) -> dict:
    """The CPU baseline, the parity band (test oracles, never the product)
    and every mode's error against the double build on the whole frame."""
    res = {"cpu_baseline": None, "parity": None}
    if args.no_cpu:
        return res
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refpy import Oracle, Ref, ref_available

    stride = max(1, args.cpu_stride)
    y0, y1 = 0, H
    rows = (H + stride - 1) // stride
    threads, cores = host_cores()
    a = s.arrays()
    if ref_available("f"):
        ref = Ref("f").load(a["pos"], a["tris"])
        cpu = ref.soft_rows(sp["view"], W, H, light, y0, y1, exact=True, mailbox=True, trail=False, threads=threads,
                            stride=stride)
        secs = float(cpu["seconds"][0] + cpu["seconds"][1])
        res["cpu_baseline"] = {"value": round(rows * W / secs, 1), "unit": "pixels/s", "cores": threads,
                               "kind": "reference", "host": cores,
                               "sample": f"every {stride}th row of the same {W}x{H} frame ({rows * W} px, "
                                         f"{secs:.2f} s: primary rays + shadow beams, mailbox on, frame stack)"}
        ref.free()
    # parity band: a few rows through the plant (the oracle is a scalar port, kept short)
    py0 = H // 2 + H // 8
    py1 = min(H, py0 + max(1, H // 32))
    orc = Oracle(a, s.kd_arrays())
    o = orc.soft(cam, W, H, light, exact=True, threads=threads, rows=(py0, py1))
    gv, ov = full["vis"][py0:py1], o["vis"][py0:py1]
    par = {"rows": [py0, py1], "tri_bitexact": bool(np.array_equal(full["tri"][py0:py1], o["tri"][py0:py1])),
           "vis_bitexact_frac": float(np.mean(gv == ov)),
           "max_abs_vis_diff": float(np.max(np.abs(gv.astype(np.float64) - ov)))}
    res["parity"] = par
    if ref_available("d"):
        refd = Ref("d").load(a["pos"], a["tris"])
        t0 = time.perf_counter()
        d = refd.soft(sp["view"], W, H, light, exact=True, mailbox=False, trail=True, threads=threads)
        d_s = time.perf_counter() - t0
        refd.free()
        modes = {"0_float_build": full}
        if guarded is not None:
            modes["1_guarded"] = guarded
        if fp64 is not None:
            modes["2_fp64"] = fp64
        res["precision"] = {k: err_vs_fp64(v, d) for k, v in modes.items()}
        res["precision"]["against"] = (f"bt_render_d's algorithm (the reference compiled with Real = double) on the "
                                       f"whole {W}x{H} frame, {d_s:.1f} s on {threads} threads; modes 0 and 1 shoot "
                                       f"from the float camera, mode 2 from the double build's camera")
    return res


if __name__ == "__main__":
    sys.exit(main())
