"""Writes the golden fixtures of tests/golden/*.npz from the reference's own
CPU build (oracle/_ref, compiled from the reference sources by oracle/Makefile).

Run in the development container (where the reference sources exist):
    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds the reference float build's outputs on this project's
synthetic scenes (inputs: the triangle arrays the host generator produces,
handed to the reference's Scene::finalize and KdTree::build):
  * hashes of the finalized scene and of the kd-tree arrays;
  * the camera (Camera::make) of the scene's view at 64 x 64;
  * soft shadows at 48 x 48, exact integration, restart trail, no mailbox
    (tri, vis, E, Eu and the counters), and the double build's vis;
  * primary visibility at 64 x 64 with 8 x 8 tiles and as one root
    (canonical polygon lists as SHA-256 digests, counts and counters);
  * point shadows of the tiled primary hits (digest, count, counters);
  * the rasterized hit ids and shadow mask of the tiled run.
Only data is stored; nothing of the reference's source.

`python tests/golden/make_golden.py large` writes tests/golden/large.npz for
the benchmark scenes (blob, atrium, conference hall, the full 2M-triangle
building): scene and kd-tree digests, soft shadows at 40 x 30 (tri, vis, E,
Eu, counters), and tiled primary visibility (16 x 16 tiles) and point
shadows at 64 x 48 (digests, counts, counters).
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from beamtrace2007_amd.host import Scene  # noqa: E402
from refpy import Ref, canon  # noqa: E402

SCENES = ["analytic", "room", "random_tris", "sphere", "plant"]
SOFT_RES = 48
PRIM_RES = 64
STAT_KEYS = ["kdSteps", "leafVisits", "triTests", "hits", "splits", "beams", "rootBeams", "presplitBeams",
             "fiveSplits", "climbs"]


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def kd_digest(split, data, offset, idx, parent, box) -> str:
    return digest(np.asarray(split, np.float32), np.asarray(data, np.uint32), np.asarray(offset, np.uint32),
                  np.asarray(idx, np.uint32), np.asarray(parent, np.uint32), np.asarray(box, np.float32))


def stats_vec(st: dict) -> np.ndarray:
    return np.array([st[k] for k in STAT_KEYS], np.int64)


def main() -> None:
    for name in SCENES:
        s = Scene.generate(name)
        a = s.arrays()
        sp = s.spec()
        ref = Ref("f").load(a["pos"], a["tris"])
        ra = ref.scene_arrays()
        rk = ref.kd_arrays()
        out = {}
        out["num_tris"] = np.int64(len(ra["tris"]))
        out["scene_digest"] = np.array(digest(ra["tris"], np.asarray(ra["triN"], np.float32),
                                              np.asarray(ra["triArea"], np.float32)))
        out["kd_digest"] = np.array(kd_digest(rk["split"], rk["data"], rk["offset"], rk["triIndices"], rk["parent"],
                                              rk["nodeBox"]))
        out["kd_nodes"] = np.int64(len(rk["split"]))
        view, light9, point3 = sp["view"], sp["area_light"], sp["point_light"]
        out["view"], out["light9"], out["point3"] = view, light9, point3
        out["cam64"] = ref.camera(view, PRIM_RES, PRIM_RES).astype(np.float32)
        out["cam48"] = ref.camera(view, SOFT_RES, SOFT_RES).astype(np.float32)
        soft = ref.soft(view, SOFT_RES, SOFT_RES, light9, exact=True, mailbox=False, trail=True, threads=8)
        out["soft_tri"], out["soft_vis"], out["soft_E"], out["soft_Eu"] = soft["tri"], soft["vis"], soft["E"], soft["Eu"]
        out["soft_stats"] = stats_vec(soft["stats"])
        out["soft_dropped"] = np.float64(soft["stats"]["droppedArea"])
        refd = Ref("d").load(a["pos"], a["tris"])
        softd = refd.soft(view, SOFT_RES, SOFT_RES, light9, exact=True, mailbox=False, trail=True, threads=8)
        out["soft_vis_fp64"] = softd["vis"].astype(np.float64)
        out["soft_E_fp64"] = softd["E"].astype(np.float64)
        refd.free()
        for tile in (0, 8):
            p = ref.primary(view, PRIM_RES, PRIM_RES, tile=tile)
            c = canon({k: p[k] for k in ("x", "y", "n", "tri", "area")})
            out[f"prim{tile}_digest"] = np.array(digest(c))
            out[f"prim{tile}_count"] = np.int64(len(c))
            out[f"prim{tile}_stats"] = stats_vec(p["stats"])
            if tile == 8:
                prim8 = p
                if len(c) <= 4000:
                    out["prim8_polys"] = c
        pt = ref.point(view, PRIM_RES, PRIM_RES, prim8, point3)
        cp = canon({k: pt[k] for k in ("x", "y", "n", "tri", "area")})
        out["point_digest"] = np.array(digest(cp))
        out["point_count"] = np.int64(len(cp))
        out["point_stats"] = stats_vec(pt["stats"])
        if len(cp) <= 4000:
            out["point_polys"] = cp
        # rasterized in canonical order (ties on shared edges are order-dependent)
        ids = ref.raster(view, PRIM_RES, PRIM_RES, {k: c[k] if k in ("n", "tri") else None for k in ("n", "tri")}
                         | {"x": c["xy"][:, 0::2], "y": c["xy"][:, 1::2]})
        mask = ref.raster(view, PRIM_RES, PRIM_RES, {"x": cp["xy"][:, 0::2], "y": cp["xy"][:, 1::2], "n": cp["n"],
                                                     "tri": cp["tri"]}, shadow=True)
        out["raster_ids"], out["raster_mask"] = ids.astype(np.int32), mask.astype(np.uint8)
        ref.free()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {int(out['num_tris'])} tris, {os.path.getsize(path)} bytes", flush=True)


LARGE = ["blob", "atrium", "conference", "building"]
LARGE_SOFT = (40, 30)
LARGE_PRIM = (64, 48)


def large() -> None:
    from beamtrace2007_amd.host import camera
    out = {}
    for name in LARGE:
        s = Scene.generate(name)
        a = s.arrays()
        sp = s.spec()
        ref = Ref("f").load(a["pos"], a["tris"])
        ra = ref.scene_arrays()
        rk = ref.kd_arrays()
        out[f"{name}_num_tris"] = np.int64(len(ra["tris"]))
        out[f"{name}_scene_digest"] = np.array(digest(ra["tris"], np.asarray(ra["triN"], np.float32),
                                                      np.asarray(ra["triArea"], np.float32)))
        out[f"{name}_kd_digest"] = np.array(kd_digest(rk["split"], rk["data"], rk["offset"], rk["triIndices"],
                                                      rk["parent"], rk["nodeBox"]))
        out[f"{name}_kd_nodes"] = np.int64(len(rk["split"]))
        view, light9, point3 = sp["view"], sp["area_light"], sp["point_light"]
        W, H = LARGE_SOFT
        soft = ref.soft(view, W, H, light9, exact=True, mailbox=False, trail=True, threads=8)
        for f in ("tri", "vis", "E", "Eu"):
            out[f"{name}_soft_{f}"] = soft[f]
        out[f"{name}_soft_stats"] = stats_vec(soft["stats"])
        W, H = LARGE_PRIM
        p = ref.primary(view, W, H, tile=16)
        c = canon({k: p[k] for k in ("x", "y", "n", "tri", "area")})
        out[f"{name}_prim16_digest"] = np.array(digest(c))
        out[f"{name}_prim16_count"] = np.int64(len(c))
        out[f"{name}_prim16_stats"] = stats_vec(p["stats"])
        pt = ref.point(view, W, H, p, point3)
        cp = canon({k: pt[k] for k in ("x", "y", "n", "tri", "area")})
        out[f"{name}_point_digest"] = np.array(digest(cp))
        out[f"{name}_point_count"] = np.int64(len(cp))
        out[f"{name}_point_stats"] = stats_vec(pt["stats"])
        ref.free()
        print(f"{name}: {int(out[f'{name}_num_tris'])} tris, {int(out[f'{name}_kd_nodes'])} kd nodes, "
              f"{len(c)} primary and {len(cp)} shadow polygons", flush=True)
    path = os.path.join(HERE, "large.npz")
    np.savez_compressed(path, **out)
    print(f"large: {os.path.getsize(path)} bytes", flush=True)


if __name__ == "__main__":
    large() if sys.argv[1:] == ["large"] else main()
