"""Python mirror of the reference's GpuRenderer (cuda/gpu_render.h:26-73) over
the C ABI of lib/libbtg.so (include/btg.h).

    r = GpuRenderer(scene)                 # scene: beamtrace2007_amd.host.Scene with its kd-tree
    res = r.soft_shadows(cam, light, exact=True)
    hits = r.primary(cam, cull=False, tile=16)
    shadow = r.point_shadows(cam, hits, light_pos)

Every call runs the gfx950 kernels; there is no CPU path. If the library or a
device is missing, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .host import LIB_DIR, Scene

_lib = None


class btg_device_info(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("computeUnits", C.c_int),
                ("wavefrontSize", C.c_int), ("memGB", C.c_double)]


class btg_scene_desc(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("numVerts", C.c_int64), ("tri3", C.c_void_p), ("triN", C.c_void_p),
                ("numTris", C.c_int64), ("nodes", C.c_void_p), ("numNodes", C.c_int64), ("triIndices", C.c_void_p),
                ("numRefs", C.c_int64), ("parent", C.c_void_p), ("nodeBox", C.c_void_p),
                ("kdBounds", C.c_float * 6), ("sceneBounds", C.c_float * 6)]


class btg_options(C.Structure):
    _fields_ = [("budget", C.c_int), ("primaryBudget", C.c_int), ("capacity", C.c_int64),
                ("roundsPerCheck", C.c_int), ("profileKernels", C.c_int), ("leafCap", C.c_int),
                ("precision", C.c_int), ("useGraphs", C.c_int), ("mailbox", C.c_int)]


class btg_camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3), ("down", C.c_float * 3),
                ("halfW", C.c_float), ("halfH", C.c_float), ("W", C.c_int), ("H", C.c_int)]


class btg_camera_d(C.Structure):
    _fields_ = [("eye", C.c_double * 3), ("fwd", C.c_double * 3), ("right", C.c_double * 3),
                ("down", C.c_double * 3), ("halfW", C.c_double), ("halfH", C.c_double)]


class btg_area_light(C.Structure):
    _fields_ = [("c", C.c_float * 3), ("U", C.c_float * 3), ("V", C.c_float * 3)]


class btg_autotune_trial(C.Structure):
    _fields_ = [("leafCap", C.c_int), ("budget", C.c_int), ("roundsPerCheck", C.c_int), ("mailbox", C.c_int),
                ("seconds", C.c_double)]


class btg_server_config(C.Structure):
    _fields_ = [("hbmBytes", C.c_int64), ("roundsPerStep", C.c_int), ("jobsPerPass", C.c_int)]


class btg_request(C.Structure):
    _fields_ = [("kind", C.c_int), ("scene", C.c_int), ("cam", btg_camera), ("light", btg_area_light),
                ("pointLight", C.c_float * 3), ("exact", C.c_int), ("tri", C.c_void_p), ("vis", C.c_void_p),
                ("E", C.c_void_p), ("Eu", C.c_void_p), ("camD", C.POINTER(btg_camera_d))]


class btg_request_status(C.Structure):
    _fields_ = [("done", C.c_int), ("status", C.c_int), ("queueSeconds", C.c_double),
                ("latencySeconds", C.c_double), ("hits", C.c_void_p), ("polys", C.c_void_p)]


class btg_soft_job(C.Structure):
    _fields_ = [("scene", C.c_int), ("cam", btg_camera), ("light", btg_area_light), ("tri", C.c_void_p),
                ("vis", C.c_void_p), ("E", C.c_void_p), ("Eu", C.c_void_p), ("camD", C.POINTER(btg_camera_d))]


STAT_FIELDS = ["kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "rootBeams", "presplitBeams",
               "mailboxSkips", "fiveSplits", "visibleTris", "rays", "climbs"]


class btg_trace_stats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in STAT_FIELDS] + [("droppedArea", C.c_double)]


class btg_run_info(C.Structure):
    _fields_ = [("primarySeconds", C.c_double), ("traceSeconds", C.c_double), ("totalSeconds", C.c_double),
                ("rounds", C.c_int), ("maxQueue", C.c_int), ("killedRoots", C.c_int), ("retryPasses", C.c_int),
                ("stats", btg_trace_stats), ("rayStats", btg_trace_stats), ("kernelSeconds", C.c_double * 5),
                ("kernelLaunches", C.c_int * 5), ("phaseCycles", C.c_uint64 * 8)]


POLY_DTYPE = np.dtype([("x", "<f4", (4,)), ("y", "<f4", (4,)), ("n", "<i4"), ("tri", "<i4"), ("area", "<f4")])
assert POLY_DTYPE.itemsize == 44

# Every entry point include/btg.h declares (checked by the tests).
EXPORTS = ["btg_device_query", "btg_create", "btg_destroy", "btg_last_error", "btg_get_options", "btg_set_options",
           "btg_soft", "btg_scene_add", "btg_scene_remove", "btg_scene_bytes", "btg_soft_batch", "btg_autotune", "btg_server_create", "btg_server_submit", "btg_server_wait",
           "btg_server_error", "btg_server_destroy", "btg_polys_count",
           "btg_polys_data", "btg_polys_free", "btg_primary", "btg_point", "btg_raster", "btg_raster_device_noalloc",
           "btg_images",
           "btg_create_file", "btg_scene_add_file", "btg_footprint"]


def lib() -> C.CDLL:
    """Loads lib/libbtg.so; raises if it was not built."""
    global _lib
    if _lib is None:
        path = os.environ.get("BTG_LIBRARY") or os.path.join(LIB_DIR, "libbtg.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python -m beamtrace2007_amd.build` first")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.btg_device_query.argtypes = [C.POINTER(btg_device_info)]
        L.btg_create.argtypes = [C.POINTER(vp), C.POINTER(btg_scene_desc)]
        L.btg_destroy.argtypes = [vp]
        L.btg_last_error.argtypes = [vp]
        L.btg_last_error.restype = C.c_char_p
        L.btg_get_options.argtypes = [vp, C.POINTER(btg_options)]
        L.btg_set_options.argtypes = [vp, C.POINTER(btg_options)]
        L.btg_soft.argtypes = [vp, C.POINTER(btg_camera), C.POINTER(btg_area_light), C.c_int, vp, vp, vp, vp,
                               C.POINTER(btg_run_info)]
        L.btg_scene_add.argtypes = [vp, C.POINTER(btg_scene_desc), C.POINTER(C.c_int)]
        L.btg_scene_remove.argtypes = [vp, C.c_int]
        L.btg_scene_bytes.argtypes = [vp, C.c_int]
        L.btg_scene_bytes.restype = C.c_int64
        L.btg_soft_batch.argtypes = [vp, C.POINTER(btg_soft_job), C.c_int, C.c_int, C.POINTER(btg_run_info)]
        L.btg_autotune.argtypes = [vp, C.POINTER(btg_soft_job), C.c_int, C.c_int, C.c_double, C.POINTER(btg_options),
                                   C.POINTER(btg_autotune_trial), C.c_int, C.POINTER(C.c_int)]
        L.btg_server_create.argtypes = [vp, C.POINTER(btg_server_config), C.POINTER(vp)]
        L.btg_server_submit.argtypes = [vp, C.POINTER(btg_request), C.POINTER(C.c_uint64)]
        L.btg_server_wait.argtypes = [vp, C.c_uint64, C.c_double, C.POINTER(btg_request_status)]
        L.btg_server_error.argtypes = [vp]
        L.btg_server_error.restype = C.c_char_p
        L.btg_server_destroy.argtypes = [vp]
        L.btg_polys_count.argtypes = [vp]
        L.btg_polys_count.restype = C.c_int64
        L.btg_polys_data.argtypes = [vp]
        L.btg_polys_data.restype = vp
        L.btg_polys_free.argtypes = [vp]
        L.btg_primary.argtypes = [vp, C.POINTER(btg_camera), C.c_int, C.c_int, C.c_int, C.POINTER(vp),
                                  C.POINTER(btg_run_info)]
        L.btg_point.argtypes = [vp, C.POINTER(btg_camera), vp, C.c_int64, C.POINTER(C.c_float), C.c_int,
                                C.POINTER(vp), C.POINTER(btg_run_info)]
        L.btg_raster.argtypes = [vp, C.POINTER(btg_camera), vp, C.c_int64, C.c_int, vp, vp]
        L.btg_raster_device_noalloc.argtypes = [vp, C.POINTER(btg_camera), vp, C.c_int64, C.c_int, vp, vp]
        L.btg_create_file.argtypes = [C.POINTER(vp), C.c_char_p]
        L.btg_scene_add_file.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
        L.btg_footprint.argtypes = [vp, C.POINTER(btg_request), C.POINTER(C.c_int64)]
        L.btg_images.argtypes = [vp, C.POINTER(btg_camera), C.c_int, C.c_int, vp, vp, vp, C.POINTER(btg_run_info),
                                 C.POINTER(btg_run_info), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def device_info() -> dict | None:
    """gpuAvailable (cuda/gpu_render.h:24): None without a usable device."""
    info = btg_device_info()
    if not lib().btg_device_query(C.byref(info)):
        return None
    return {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.computeUnits,
            "wavefront": info.wavefrontSize, "mem_gb": info.memGB}


def make_camera(cam14, W: int, H: int) -> btg_camera:
    c = np.asarray(cam14, dtype=np.float32)
    bc = btg_camera()
    for i, name in enumerate(["eye", "fwd", "right", "down"]):
        getattr(bc, name)[:] = [float(v) for v in c[3 * i:3 * i + 3]]
    bc.halfW, bc.halfH, bc.W, bc.H = float(c[12]), float(c[13]), int(W), int(H)
    return bc


def info_dict(ri: btg_run_info) -> dict:
    st = {f: int(getattr(ri.stats, f)) for f in STAT_FIELDS}
    st["droppedArea"] = float(ri.stats.droppedArea)
    rs = {f: int(getattr(ri.rayStats, f)) for f in STAT_FIELDS}
    return {"primary_s": ri.primarySeconds, "trace_s": ri.traceSeconds, "total_s": ri.totalSeconds,
            "rounds": ri.rounds, "max_queue": ri.maxQueue, "killed_roots": ri.killedRoots,
            "retry_passes": ri.retryPasses, "stats": st, "ray_stats": rs,
            "kernel_s": dict(zip(KERNELS, [float(v) for v in ri.kernelSeconds])),
            "launches": dict(zip(KERNELS, [int(v) for v in ri.kernelLaunches])),
            "phase_cycles": dict(zip(PHASES, [int(v) for v in ri.phaseCycles]))}


KERNELS = ["admit", "trace", "split", "primary_rays", "other"]
# trace-loop regions timed by BTG_PHASE_CYCLES builds (btg_run_info.phaseCycles)
PHASES = ["refill", "descend", "climb_walk", "climb_box", "emit", "test", "append", "advance_rest"]


def polys_to_dict(arr: np.ndarray) -> dict:
    return {"x": arr["x"].copy(), "y": arr["y"].copy(), "n": arr["n"].copy(), "tri": arr["tri"].copy(),
            "area": arr["area"].copy()}


def dict_to_polys(p: dict) -> np.ndarray:
    n = len(p["n"])
    a = np.zeros(n, POLY_DTYPE)
    a["x"], a["y"], a["n"], a["tri"], a["area"] = p["x"], p["y"], p["n"], p["tri"], p["area"]
    return a


def _scene_desc(scene: Scene):
    a = scene.arrays()
    k = scene.kd_arrays()
    d = btg_scene_desc()
    d.pos, d.numVerts = a["pos"].ctypes.data, len(a["pos"])
    d.tri3, d.triN, d.numTris = a["tris"].ctypes.data, a["triN"].ctypes.data, len(a["tris"])
    d.nodes, d.numNodes = k["nodes"].ctypes.data, len(k["nodes"])
    d.triIndices, d.numRefs = k["triIndices"].ctypes.data, len(k["triIndices"])
    d.parent, d.nodeBox = k["parent"].ctypes.data, k["nodeBox"].ctypes.data
    d.kdBounds[:] = [float(v) for v in k["bounds"]]
    d.sceneBounds[:] = [float(v) for v in a["bounds"]]
    return d, (a, k)


def _light(light9) -> btg_area_light:
    lt = btg_area_light()
    l9 = np.asarray(light9, np.float32)
    lt.c[:], lt.U[:], lt.V[:] = [float(v) for v in l9[0:3]], [float(v) for v in l9[3:6]], [float(v) for v in l9[6:9]]
    return lt


class GpuRenderer:
    """The reference's GpuRenderer over one device context. The scene must
    have its kd-tree built (Scene.build_kd)."""

    def __init__(self, scene):
        """`scene`: a Scene with its kd-tree built, or the path of a resident
        scene file (Scene.save), uploaded without rebuilding the tree."""
        L = lib()
        h = C.c_void_p()
        if isinstance(scene, (str, os.PathLike)):
            if L.btg_create_file(C.byref(h), os.fsencode(scene)) != 0:
                raise RuntimeError("btg_create_file: " + L.btg_last_error(None).decode())
        else:
            d, keep = _scene_desc(scene)
            if L.btg_create(C.byref(h), C.byref(d)) != 0:
                raise RuntimeError("btg_create: " + L.btg_last_error(None).decode())
        self._h = h

    def add_scene(self, scene) -> int:
        """Makes another scene resident (a Scene with its kd-tree built, or a
        resident scene file's path); returns its id for soft_batch jobs."""
        sid = C.c_int(-1)
        if isinstance(scene, (str, os.PathLike)):
            rc = lib().btg_scene_add_file(self._h, os.fsencode(scene), C.byref(sid))
            if rc:
                self._err("btg_scene_add_file", rc)
            return sid.value
        d, keep = _scene_desc(scene)
        rc = lib().btg_scene_add(self._h, C.byref(d), C.byref(sid))
        if rc:
            self._err("btg_scene_add", rc)
        return sid.value

    def remove_scene(self, scene_id: int) -> None:
        rc = lib().btg_scene_remove(self._h, int(scene_id))
        if rc:
            self._err("btg_scene_remove", rc)

    def scene_bytes(self, scene_id: int) -> int:
        return int(lib().btg_scene_bytes(self._h, int(scene_id)))

    def _jobs(self, jobs, outs=None):
        """The jobs as btg_soft_job records (a job's optional "cam_d": its camera in 14 doubles, for
        precision 2); the double cameras are kept alive on the array."""
        arr = (btg_soft_job * len(jobs))()
        arr._camd = []
        for i, j in enumerate(jobs):
            arr[i].scene = int(j.get("scene", 0))
            arr[i].cam = make_camera(j["cam"], int(j["W"]), int(j["H"]))
            arr[i].light = _light(j["light"])
            for f in ("tri", "vis", "E", "Eu"):
                setattr(arr[i], f, outs[i][f].ctypes.data if outs else None)
            if j.get("cam_d") is not None:
                c = np.ascontiguousarray(j["cam_d"], np.float64)
                if c.shape != (14,):
                    raise ValueError("cam_d must hold 14 doubles (eye, fwd, right, down, halfW, halfH)")
                cd = btg_camera_d.from_buffer_copy(c.tobytes())
                arr._camd.append(cd)
                arr[i].camD = C.pointer(cd)
        return arr

    def autotune(self, jobs, exact: bool = True, max_seconds: float = 20.0):
        """Tunes leaf cap, budget and check interval on the sample batch `jobs`; returns (options dict, trials).
        The renderer's options are not changed (apply with set_options)."""
        arr = self._jobs(jobs)
        best = btg_options()
        trials = (btg_autotune_trial * 64)()
        n = C.c_int(0)
        rc = lib().btg_autotune(self._h, arr, len(jobs), int(exact), float(max_seconds), C.byref(best), trials, 64,
                                C.byref(n))
        if rc:
            self._err("btg_autotune", rc)
        opts = {"budget": best.budget, "rounds_per_check": best.roundsPerCheck, "leaf_cap": best.leafCap,
                "mailbox": bool(best.mailbox)}
        tr = [{"leaf_cap": t.leafCap, "budget": t.budget, "rounds_per_check": t.roundsPerCheck, "mailbox": bool(t.mailbox),
               "seconds": t.seconds} for t in trials[:min(n.value, 64)]]
        return opts, tr

    def soft_batch(self, jobs, exact: bool = True, download: bool = True):
        """Soft shadows for many jobs in one persistent pass. jobs: dicts with scene (id), cam (14 floats),
        W, H, light (9 floats). Returns (list of per-job dicts with tri/vis/E/Eu, run info)."""
        n = len(jobs)
        outs = []
        for j in jobs:
            W, H = int(j["W"]), int(j["H"])
            outs.append({"tri": np.empty((H, W), np.int32), "vis": np.empty((H, W), np.float32),
                         "E": np.empty((H, W), np.float32), "Eu": np.empty((H, W), np.float32)} if download else {})
        arr = self._jobs(jobs, outs if download else None)
        ri = btg_run_info()
        rc = lib().btg_soft_batch(self._h, arr, n, int(exact), C.byref(ri))
        if rc:
            self._err("btg_soft_batch", rc)
        return outs, info_dict(ri)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().btg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, what: str, rc: int):
        raise RuntimeError(f"{what} failed ({rc}): " + lib().btg_last_error(self._h).decode())

    @property
    def options(self) -> dict:
        o = btg_options()
        lib().btg_get_options(self._h, C.byref(o))
        return {"budget": o.budget, "primary_budget": o.primaryBudget, "capacity": o.capacity,
                "rounds_per_check": o.roundsPerCheck, "profile": bool(o.profileKernels), "leaf_cap": o.leafCap,
                "precision": o.precision, "graphs": bool(o.useGraphs), "mailbox": bool(o.mailbox)}

    def set_options(self, **kw) -> None:
        cur = self.options
        cur.update(kw)
        o = btg_options(cur["budget"], cur["primary_budget"], cur["capacity"], cur["rounds_per_check"],
                        int(cur["profile"]), int(cur["leaf_cap"]), int(cur["precision"]), int(cur["graphs"]),
                        int(cur["mailbox"]))
        rc = lib().btg_set_options(self._h, C.byref(o))
        if rc:
            self._err("btg_set_options", rc)

    def soft_shadows(self, cam14, W: int, H: int, light9, exact: bool = True, download: bool = True,
                     cam_d=None) -> dict:
        """softShadows (cuda/gpu_render.h:36): tri, vis, E, Eu images (H, W) and run info.
        download=False leaves the images in device memory (no PCIe transfer). cam_d: the camera in 14
        doubles (host.camera_d), which precision 2 shoots its primary rays from (a one-job batch)."""
        if cam_d is not None:
            outs, info = self.soft_batch([{"scene": 0, "cam": cam14, "W": W, "H": H, "light": light9,
                                           "cam_d": cam_d}], exact=exact, download=download)
            return dict(outs[0], info=info)
        n = W * H if download else 0
        tri, vis = np.empty(n, np.int32), np.empty(n, np.float32)
        E, Eu = np.empty(n, np.float32), np.empty(n, np.float32)
        lt = _light(light9)
        ri = btg_run_info()
        ptr = (lambda a: a.ctypes.data) if download else (lambda a: None)
        rc = lib().btg_soft(self._h, C.byref(make_camera(cam14, W, H)), C.byref(lt), int(exact), ptr(tri), ptr(vis),
                            ptr(E), ptr(Eu), C.byref(ri))
        if rc:
            self._err("btg_soft", rc)
        if not download:
            return {"info": info_dict(ri)}
        return {"tri": tri.reshape(H, W), "vis": vis.reshape(H, W), "E": E.reshape(H, W), "Eu": Eu.reshape(H, W),
                "info": info_dict(ri)}

    def _take(self, ph: C.c_void_p) -> np.ndarray:
        L = lib()
        n = L.btg_polys_count(ph)
        out = np.zeros(n, POLY_DTYPE)
        if n:
            C.memmove(out.ctypes.data, L.btg_polys_data(ph), n * POLY_DTYPE.itemsize)
        L.btg_polys_free(ph)
        return out

    def primary(self, cam14, W: int, H: int, cull: bool = False, tile: int = 16, canonical: bool = True):
        """primary (cuda/gpu_render.h:41): hit polygons as a structured array, and run info."""
        ph, ri = C.c_void_p(), btg_run_info()
        rc = lib().btg_primary(self._h, C.byref(make_camera(cam14, W, H)), int(cull), int(tile), int(canonical),
                               C.byref(ph), C.byref(ri))
        if rc:
            self._err("btg_primary", rc)
        return self._take(ph), info_dict(ri)

    def point_shadows(self, cam14, W: int, H: int, prim: np.ndarray, light3, canonical: bool = True):
        """pointShadows (cuda/gpu_render.h:45): shadowed polygons for primary hit polygons."""
        prim = np.ascontiguousarray(prim, dtype=POLY_DTYPE)
        lt = (C.c_float * 3)(*[float(v) for v in light3])
        ph, ri = C.c_void_p(), btg_run_info()
        rc = lib().btg_point(self._h, C.byref(make_camera(cam14, W, H)), prim.ctypes.data if len(prim) else None,
                             len(prim), lt, int(canonical), C.byref(ph), C.byref(ri))
        if rc:
            self._err("btg_point", rc)
        return self._take(ph), info_dict(ri)

    def raster(self, cam14, W: int, H: int, polys: np.ndarray, shadow: bool = False) -> np.ndarray:
        """rasterize (src/render/render.cpp:49-95) on the device: an int32 ID image
        (shadow = False) or a uint8 mask (shadow = True), H x W."""
        polys = np.ascontiguousarray(polys, dtype=POLY_DTYPE)
        out = np.empty((H, W), np.uint8 if shadow else np.int32)
        rc = lib().btg_raster(self._h, C.byref(make_camera(cam14, W, H)), polys.ctypes.data if len(polys) else None,
                              len(polys), int(shadow), None if shadow else out.ctypes.data,
                              out.ctypes.data if shadow else None)
        if rc:
            self._err("btg_raster", rc)
        return out

    def raster_device_noalloc(self, cam14, W: int, H: int, polys_dev: int, count: int, out_dev: int,
                              shadow: bool = False) -> None:
        """btg_raster_device_noalloc: rasterizes `count` polygons at device address
        polys_dev into the device image at out_dev (int32 ids, or uint8 mask when
        shadow), allocating nothing (at most BTG_RASTER_NOALLOC_MAX_POLYS)."""
        rc = lib().btg_raster_device_noalloc(self._h, C.byref(make_camera(cam14, W, H)), polys_dev or None, int(count),
                                             int(shadow), None if shadow else out_dev, out_dev if shadow else None)
        if rc:
            self._err("btg_raster_device_noalloc", rc)

    def images(self, cam14, W: int, H: int, light3=None, cull: bool = False, tile: int = 16):
        """Primary IDs and (light3 given) the point-shadow mask of one frame, both
        rasterized on the device: (ids, mask or None, info)."""
        ids = np.empty((H, W), np.int32)
        mask = np.empty((H, W), np.uint8) if light3 is not None else None
        lt = (C.c_float * 3)(*[float(v) for v in light3]) if light3 is not None else None
        a, b, rs = btg_run_info(), btg_run_info(), C.c_double()
        rc = lib().btg_images(self._h, C.byref(make_camera(cam14, W, H)), int(cull), int(tile), lt, ids.ctypes.data,
                              mask.ctypes.data if mask is not None else None, C.byref(a), C.byref(b), C.byref(rs))
        if rc:
            self._err("btg_images", rc)
        return ids, mask, {"primary": info_dict(a), "point": info_dict(b), "raster_s": rs.value}


def _take_polys(ph) -> np.ndarray:
    L = lib()
    n = L.btg_polys_count(ph)
    out = np.zeros(n, POLY_DTYPE)
    if n:
        C.memmove(out.ctypes.data, L.btg_polys_data(ph), n * POLY_DTYPE.itemsize)
    L.btg_polys_free(ph)
    return out


class Server:
    """Request server over a renderer's context (btg_server_*): soft-shadow
    requests join a running persistent pass (continuous batching), point-shadow
    requests overlap on their own stream. The renderer must not be used while
    the server exists."""

    def __init__(self, renderer: GpuRenderer, hbm_bytes: int = 0, rounds_per_step: int = 0, jobs_per_pass: int = 0):
        self._r = renderer
        cfg = btg_server_config(int(hbm_bytes), int(rounds_per_step), int(jobs_per_pass))
        h = C.c_void_p()
        rc = lib().btg_server_create(renderer._h, C.byref(cfg), C.byref(h))
        if rc:
            renderer._err("btg_server_create", rc)
        self._h = h
        self._outs = {}

    def submit_soft(self, scene: int, cam14, W: int, H: int, light9, cam_d=None) -> int:
        """A soft-shadow request; cam_d (14 doubles, host.camera_d) is read by
        FP64 servers (a context at precision 2) and copied at submission."""
        o = {"tri": np.empty((H, W), np.int32), "vis": np.empty((H, W), np.float32),
             "E": np.empty((H, W), np.float32), "Eu": np.empty((H, W), np.float32)}
        q = btg_request()
        q.kind, q.scene, q.exact = 0, int(scene), 1
        q.cam = make_camera(cam14, W, H)
        q.light = _light(light9)
        if cam_d is not None:
            c = np.ascontiguousarray(cam_d, np.float64)
            if c.shape != (14,):
                raise ValueError("cam_d must hold 14 doubles (eye, fwd, right, down, halfW, halfH)")
            cd = btg_camera_d.from_buffer_copy(c.tobytes())
            q.camD = C.pointer(cd)
        for f in ("tri", "vis", "E", "Eu"):
            setattr(q, f, o[f].ctypes.data)
        return self._submit(q, o)

    def footprint(self, kind: str, cam14, W: int, H: int) -> int:
        """btg_footprint: device bytes a "soft" or "point" request holds while in flight."""
        q = btg_request()
        q.kind, q.scene, q.exact = (0 if kind == "soft" else 1), 0, 1
        q.cam = make_camera(cam14, W, H)
        b = C.c_int64(0)
        rc = lib().btg_footprint(self._r._h, C.byref(q), C.byref(b))
        if rc:
            raise RuntimeError(f"btg_footprint failed ({rc})")
        return int(b.value)

    def submit_point(self, cam14, W: int, H: int, light3) -> int:
        q = btg_request()
        q.kind, q.scene = 1, 0
        q.cam = make_camera(cam14, W, H)
        q.pointLight[:] = [float(v) for v in light3]
        return self._submit(q, {})

    def _submit(self, q, outs) -> int:
        t = C.c_uint64(0)
        rc = lib().btg_server_submit(self._h, C.byref(q), C.byref(t))
        if rc:
            raise RuntimeError(f"btg_server_submit failed ({rc}): " + lib().btg_server_error(self._h).decode())
        self._outs[t.value] = outs
        return t.value

    def wait(self, ticket: int, timeout: float = -1.0) -> dict | None:
        st = btg_request_status()
        rc = lib().btg_server_wait(self._h, int(ticket), float(timeout), C.byref(st))
        if rc:
            raise RuntimeError(f"btg_server_wait failed ({rc}): " + lib().btg_server_error(self._h).decode())
        if not st.done:
            return None
        res = dict(self._outs.pop(ticket))
        res.update({"status": st.status, "queue_s": st.queueSeconds, "latency_s": st.latencySeconds})
        if st.hits:
            res["hits"] = _take_polys(C.c_void_p(st.hits))
        if st.polys:
            res["polys"] = _take_polys(C.c_void_p(st.polys))
        if st.status:
            res["error"] = lib().btg_server_error(self._h).decode()
        return res

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().btg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
