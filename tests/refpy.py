"""ctypes bindings for the test oracles (never imported by the product).

* `Ref`    — the reference's own CPU build, compiled from its sources into
             oracle/_ref/libbtref_{f,d}.so by oracle/Makefile (oracle/ref_shim.cpp);
* `Oracle` — this project's C++ restatement, oracle/build/libbt_oracle.so.

Both take scenes as raw arrays and return numpy arrays. Polygon lists are
dicts {x (n,4), y (n,4), n, tri, area}.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "libbt_oracle.so")

STAT_NAMES = ["kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "rootBeams", "presplitBeams",
              "mailboxSkips", "fiveSplits", "visibleTris", "rays", "climbs", "droppedArea"]

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p
i64 = C.c_int64


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)


def ref_available(prec: str = "f") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libbtref_{prec}.so"))


def stats_dict(a) -> dict:
    return {k: (float(v) if k == "droppedArea" else int(v)) for k, v in zip(STAT_NAMES, a)}


class Ref:
    """The reference CPU build (float: prec='f', double: prec='d')."""

    _libs: dict = {}

    def __init__(self, prec: str = "f"):
        if prec not in Ref._libs:
            L = C.CDLL(os.path.join(REF_DIR, f"libbtref_{prec}.so"))
            L.ref_scene_new.restype = vp
            L.ref_scene_new.argtypes = [fp, i64, u32p, i64]
            L.ref_scene_free.argtypes = [vp]
            L.ref_scene_ntris.restype = i64
            L.ref_scene_ntris.argtypes = [vp]
            L.ref_scene_arrays.argtypes = [vp, u32p, dp, dp, dp]
            L.ref_kd_new.restype = vp
            L.ref_kd_new.argtypes = [vp, C.c_double, C.c_int]
            L.ref_kd_free.argtypes = [vp]
            L.ref_kd_counts.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
            L.ref_kd_arrays.argtypes = [vp, dp, u32p, u32p, u32p, u32p, dp, dp]
            L.ref_camera.argtypes = [dp, C.c_int, C.c_int, dp]
            L.ref_soft.argtypes = [vp, vp, dp, C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int, i32p, fp, fp,
                                   fp, dp, dp]
            L.ref_soft_rows.argtypes = [vp, vp, dp, C.c_int, C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, fp, dp, dp, i32p, fp]
            L.ref_light_vis.restype = C.c_double
            L.ref_light_vis.argtypes = [vp, vp, dp, dp, C.c_int, dp, C.c_double, C.c_int, C.c_int, dp, dp]
            L.ref_primary.restype = vp
            L.ref_primary.argtypes = [vp, vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
            L.ref_point.restype = vp
            L.ref_point.argtypes = [vp, vp, dp, C.c_int, C.c_int, dp, dp, i32p, i32p, dp, i64, dp, C.c_int, C.c_int, dp]
            L.ref_polys_count.restype = i64
            L.ref_polys_count.argtypes = [vp]
            L.ref_polys_get.argtypes = [vp, dp, dp, i32p, i32p, dp]
            L.ref_polys_free.argtypes = [vp]
            L.ref_raster.argtypes = [dp, C.c_int, C.c_int, dp, dp, i32p, i32p, i64, C.c_int, i32p,
                                     C.POINTER(C.c_uint8)]
            L.ref_primary_rays.argtypes = [vp, vp, dp, C.c_int, C.c_int, i32p, dp, dp, dp]
            Ref._libs[prec] = L
        self.L = Ref._libs[prec]
        self.prec = prec
        self.scene = None
        self.kd = None

    def load(self, pos: np.ndarray, tris: np.ndarray, cost_intersect: float = 0.4, single_leaf: bool = False):
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        tris = np.ascontiguousarray(tris, dtype=np.uint32)
        self.free()
        self.scene = self.L.ref_scene_new(pos.ctypes.data_as(fp), len(pos), tris.ctypes.data_as(u32p), len(tris))
        self.kd = self.L.ref_kd_new(self.scene, cost_intersect, int(single_leaf))
        return self

    def free(self):
        if self.kd:
            self.L.ref_kd_free(self.kd)
            self.kd = None
        if self.scene:
            self.L.ref_scene_free(self.scene)
            self.scene = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def scene_arrays(self) -> dict:
        nt = self.L.ref_scene_ntris(self.scene)
        tris = np.zeros((nt, 3), np.uint32)
        triN = np.zeros((nt, 3))
        area = np.zeros(nt)
        b = np.zeros(6)
        self.L.ref_scene_arrays(self.scene, tris.ctypes.data_as(u32p), triN.ctypes.data_as(dp),
                                area.ctypes.data_as(dp), b.ctypes.data_as(dp))
        return {"tris": tris, "triN": triN, "triArea": area, "bounds": b}

    def kd_arrays(self) -> dict:
        nn, ni = i64(), i64()
        self.L.ref_kd_counts(self.kd, C.byref(nn), C.byref(ni))
        N, R = nn.value, ni.value
        split, data, off = np.zeros(N), np.zeros(N, np.uint32), np.zeros(N, np.uint32)
        idx, par = np.zeros(R, np.uint32), np.zeros(N, np.uint32)
        box, b = np.zeros((N, 6)), np.zeros(6)
        self.L.ref_kd_arrays(self.kd, split.ctypes.data_as(dp), data.ctypes.data_as(u32p), off.ctypes.data_as(u32p),
                             idx.ctypes.data_as(u32p), par.ctypes.data_as(u32p), box.ctypes.data_as(dp),
                             b.ctypes.data_as(dp))
        return {"split": split, "data": data, "offset": off, "triIndices": idx, "parent": par, "nodeBox": box,
                "bounds": b}

    def camera(self, view10, W, H) -> np.ndarray:
        out = np.zeros(14)
        self.L.ref_camera(_d(view10), W, H, out.ctypes.data_as(dp))
        return out

    def soft(self, view10, W, H, light9, exact=True, mailbox=False, trail=True, threads=8):
        n = W * H
        tri, vis, E, Eu = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        st, sec = np.zeros(14), np.zeros(2)
        self.L.ref_soft(self.scene, self.kd, _d(view10), W, H, _d(light9), int(exact), int(mailbox), int(trail),
                        threads, tri.ctypes.data_as(i32p), vis.ctypes.data_as(fp), E.ctypes.data_as(fp),
                        Eu.ctypes.data_as(fp), st.ctypes.data_as(dp), sec.ctypes.data_as(dp))
        return {"tri": tri.reshape(H, W), "vis": vis.reshape(H, W), "E": E.reshape(H, W), "Eu": Eu.reshape(H, W),
                "stats": stats_dict(st), "seconds": sec}

    def soft_rows(self, view10, W, H, light9, y0, y1, exact=True, mailbox=True, trail=False, threads=8, stride=1):
        rows = (y1 - y0 + stride - 1) // stride
        vis, st, sec = np.zeros(rows * W, np.float32), np.zeros(14), np.zeros(2)
        tri = np.zeros(rows * W, np.int32)
        E = np.zeros(rows * W, np.float32)
        self.L.ref_soft_rows(self.scene, self.kd, _d(view10), W, H, _d(light9), int(exact), int(mailbox), int(trail),
                             threads, y0, y1, stride, vis.ctypes.data_as(fp), st.ctypes.data_as(dp),
                             sec.ctypes.data_as(dp), tri.ctypes.data_as(i32p), E.ctypes.data_as(fp))
        out = {"vis": vis.reshape(rows, W), "tri": tri.reshape(rows, W), "stats": stats_dict(st), "seconds": sec}
        if exact:
            out["E"] = E.reshape(rows, W)
        return out

    def light_vis(self, x, ns, tri, light9, offset, mailbox=False, trail=True):
        G, st = np.zeros(1), np.zeros(14)
        V = self.L.ref_light_vis(self.scene, self.kd, _d(x), _d(ns), int(tri), _d(light9), float(offset),
                                 int(mailbox), int(trail), G.ctypes.data_as(dp), st.ctypes.data_as(dp))
        return V, float(G[0]), stats_dict(st)

    def _polys(self, h) -> dict:
        n = self.L.ref_polys_count(h)
        x, y = np.zeros((n, 4)), np.zeros((n, 4))
        nn, tri, area = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        self.L.ref_polys_get(h, x.ctypes.data_as(dp), y.ctypes.data_as(dp), nn.ctypes.data_as(i32p),
                             tri.ctypes.data_as(i32p), area.ctypes.data_as(dp))
        self.L.ref_polys_free(h)
        return {"x": x, "y": y, "n": nn, "tri": tri, "area": area}

    def primary(self, view10, W, H, cull=False, mailbox=False, trail=True, tile=0):
        st = np.zeros(14)
        h = self.L.ref_primary(self.scene, self.kd, _d(view10), W, H, int(cull), int(mailbox), int(trail), tile,
                               st.ctypes.data_as(dp))
        p = self._polys(h)
        p["stats"] = stats_dict(st)
        return p

    def point(self, view10, W, H, prim: dict, light3, mailbox=False, trail=True):
        st = np.zeros(14)
        n = len(prim["n"])
        h = self.L.ref_point(self.scene, self.kd, _d(view10), W, H, _d(prim["x"]), _d(prim["y"]),
                             np.ascontiguousarray(prim["n"], np.int32).ctypes.data_as(i32p),
                             np.ascontiguousarray(prim["tri"], np.int32).ctypes.data_as(i32p), _d(prim["area"]), n,
                             _d(light3), int(mailbox), int(trail), st.ctypes.data_as(dp))
        p = self._polys(h)
        p["stats"] = stats_dict(st)
        return p

    def raster(self, view10, W, H, polys: dict, shadow=False, base=None):
        n = len(polys["n"])
        tri_o = None if shadow else (np.full(W * H, -1, np.int32) if base is None else base.astype(np.int32).ravel().copy())
        flag_o = (np.zeros(W * H, np.uint8) if base is None else base.astype(np.uint8).ravel().copy()) if shadow else None
        self.L.ref_raster(_d(view10), W, H, _d(polys["x"]), _d(polys["y"]),
                          np.ascontiguousarray(polys["n"], np.int32).ctypes.data_as(i32p),
                          np.ascontiguousarray(polys["tri"], np.int32).ctypes.data_as(i32p), n, int(shadow),
                          tri_o.ctypes.data_as(i32p) if tri_o is not None else None,
                          flag_o.ctypes.data_as(C.POINTER(C.c_uint8)) if flag_o is not None else None)
        return (flag_o if shadow else tri_o).reshape(H, W)

    def primary_rays(self, view10, W, H):
        n = W * H
        tri, X, N, st = np.zeros(n, np.int32), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(14)
        self.L.ref_primary_rays(self.scene, self.kd, _d(view10), W, H, tri.ctypes.data_as(i32p),
                                X.ctypes.data_as(dp), N.ctypes.data_as(dp), st.ctypes.data_as(dp))
        return {"tri": tri.reshape(H, W), "X": X.reshape(H, W, 3), "N": N.reshape(H, W, 3), "stats": stats_dict(st)}


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(fp)


class Oracle:
    """This project's C++ restatement of the reference CPU tracer (trail, no mailbox)."""

    _lib = None

    def __init__(self, scene_arrays: dict, kd_arrays: dict):
        if Oracle._lib is None:
            L = C.CDLL(ORACLE_LIB)
            L.orc_scene_new.restype = vp
            L.orc_scene_new.argtypes = [fp, i64, u32p, fp, i64, vp, i64, u32p, i64, u32p, fp, fp, fp]
            L.orc_scene_free.argtypes = [vp]
            L.orc_soft.argtypes = [vp, fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, C.c_int, C.c_int, i32p, fp, fp, fp,
                                   dp, C.c_int]
            L.orc_primary.restype = vp
            L.orc_primary.argtypes = [vp, fp, C.c_int, C.c_int, C.c_int, C.c_int, dp]
            L.orc_primary_tiles.restype = vp
            L.orc_primary_tiles.argtypes = [vp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp]
            L.orc_point.restype = vp
            L.orc_point.argtypes = [vp, fp, C.c_int, C.c_int, fp, fp, i32p, i32p, fp, i64, fp, dp]
            L.orc_polys_count.restype = i64
            L.orc_polys_count.argtypes = [vp]
            L.orc_polys_get.argtypes = [vp, fp, fp, i32p, i32p, fp]
            L.orc_polys_free.argtypes = [vp]
            L.orc_primary_rays.argtypes = [vp, fp, C.c_int, C.c_int, i32p, fp, fp, dp]
            L.orc_light_vis.restype = C.c_float
            L.orc_light_vis.argtypes = [vp, fp, fp, C.c_int, fp, C.c_float, fp, dp]
            Oracle._lib = L
        self.L = Oracle._lib
        s, k = scene_arrays, kd_arrays
        self._keep = [np.ascontiguousarray(s["pos"], np.float32), np.ascontiguousarray(s["tris"], np.uint32),
                      np.ascontiguousarray(s["triN"], np.float32), np.ascontiguousarray(k["nodes"]),
                      np.ascontiguousarray(k["triIndices"], np.uint32), np.ascontiguousarray(k["parent"], np.uint32),
                      np.ascontiguousarray(k["nodeBox"], np.float32), np.ascontiguousarray(k["bounds"], np.float32),
                      np.ascontiguousarray(s["bounds"], np.float32)]
        pos, tris, triN, nodes, idx, par, box, kb, sb = self._keep
        self.h = self.L.orc_scene_new(pos.ctypes.data_as(fp), len(pos), tris.ctypes.data_as(u32p),
                                      triN.ctypes.data_as(fp), len(tris), nodes.ctypes.data, len(nodes),
                                      idx.ctypes.data_as(u32p), len(idx), par.ctypes.data_as(u32p),
                                      box.ctypes.data_as(fp), kb.ctypes.data_as(fp), sb.ctypes.data_as(fp))

    def __del__(self):
        try:
            self.L.orc_scene_free(self.h)
        except Exception:
            pass

    def soft(self, cam14, W, H, light9, exact=True, threads=8, rows=None, guard=False):
        n = W * H
        tri, vis = np.full(n, -1, np.int32), np.full(n, -1, np.float32)
        E, Eu, st = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(14)
        y0, y1 = rows if rows is not None else (0, 0)
        self.L.orc_soft(self.h, _f(cam14), W, H, _f(light9), int(exact), threads, y0, y1, tri.ctypes.data_as(i32p),
                        vis.ctypes.data_as(fp), E.ctypes.data_as(fp), Eu.ctypes.data_as(fp), st.ctypes.data_as(dp),
                        int(guard))
        return {"tri": tri.reshape(H, W), "vis": vis.reshape(H, W), "E": E.reshape(H, W), "Eu": Eu.reshape(H, W),
                "stats": stats_dict(st)}

    def _polys(self, h) -> dict:
        n = self.L.orc_polys_count(h)
        x, y = np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32)
        nn, tri, area = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.float32)
        self.L.orc_polys_get(h, x.ctypes.data_as(fp), y.ctypes.data_as(fp), nn.ctypes.data_as(i32p),
                             tri.ctypes.data_as(i32p), area.ctypes.data_as(fp))
        self.L.orc_polys_free(h)
        return {"x": x, "y": y, "n": nn, "tri": tri, "area": area}

    def primary(self, cam14, W, H, cull=False, tile=0, tile_rows=None):
        """tile_rows=(ty0, ty1): only the tiles of those tile rows (tile > 0)."""
        st = np.zeros(14)
        if tile_rows is None:
            h = self.L.orc_primary(self.h, _f(cam14), W, H, int(cull), tile, st.ctypes.data_as(dp))
        else:
            h = self.L.orc_primary_tiles(self.h, _f(cam14), W, H, int(cull), tile, int(tile_rows[0]),
                                         int(tile_rows[1]), st.ctypes.data_as(dp))
        p = self._polys(h)
        p["stats"] = stats_dict(st)
        return p

    def point(self, cam14, W, H, prim: dict, light3):
        st = np.zeros(14)
        n = len(prim["n"])
        h = self.L.orc_point(self.h, _f(cam14), W, H, _f(prim["x"]), _f(prim["y"]),
                             np.ascontiguousarray(prim["n"], np.int32).ctypes.data_as(i32p),
                             np.ascontiguousarray(prim["tri"], np.int32).ctypes.data_as(i32p), _f(prim["area"]), n,
                             _f(light3), st.ctypes.data_as(dp))
        p = self._polys(h)
        p["stats"] = stats_dict(st)
        return p

    def primary_rays(self, cam14, W, H):
        n = W * H
        tri, X, N, st = np.zeros(n, np.int32), np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(14)
        self.L.orc_primary_rays(self.h, _f(cam14), W, H, tri.ctypes.data_as(i32p), X.ctypes.data_as(fp),
                                N.ctypes.data_as(fp), st.ctypes.data_as(dp))
        return {"tri": tri.reshape(H, W), "X": X.reshape(H, W, 3), "N": N.reshape(H, W, 3), "stats": stats_dict(st)}


def canon(p: dict) -> np.ndarray:
    """Polygons as a canonically sorted structured array (tri, n, corners, area):
    the order-free multiset compared bit for bit (float32 values)."""
    n = len(p["n"])
    dt = np.dtype([("tri", "<i4"), ("n", "<i4"), ("xy", "<f4", (8,)), ("area", "<f4")])
    a = np.zeros(n, dt)
    a["tri"] = p["tri"]
    a["n"] = p["n"]
    xy = np.zeros((n, 8), np.float32)
    xy[:, 0::2] = np.asarray(p["x"], np.float32)
    xy[:, 1::2] = np.asarray(p["y"], np.float32)
    a["xy"] = xy
    a["area"] = np.asarray(p["area"], np.float32)
    keys = [a["area"]] + [xy[:, k] for k in range(7, -1, -1)] + [a["n"], a["tri"]]
    return a[np.lexsort(keys)]


def _oracle_light_vis(self, x3, ns3, tri, light9, offset):
    G, st = np.zeros(1, np.float32), np.zeros(14)
    V = self.L.orc_light_vis(self.h, _f(x3), _f(ns3), int(tri), _f(light9), float(offset),
                             G.ctypes.data_as(fp), st.ctypes.data_as(dp))
    return float(V), float(G[0]), stats_dict(st)


Oracle.light_vis = _oracle_light_vis
