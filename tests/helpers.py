"""Shared fixtures-as-functions for the tests: seeded scenes, host octrees,
oracle scenes, and the comparison of a GPU frame with an oracle frame."""
import functools

import numpy as np

import qubatron_amd as q
from oracle import pyoracle as po


@functools.lru_cache(maxsize=8)
def scene(config, seed=0, npoints=0):
    sc = q.Scene(config, seed=seed, npoints=npoints)
    pnt, nrm, col = sc.points()
    tree = q.octree_from_points(pnt, sc.basecube, sc.depth)
    nodes = tree.nodes()
    return sc, pnt, nrm, col, nodes


def oracle_scene(config, seed=0, npoints=0, maxlevel=None):
    sc, pnt, nrm, col, nodes = scene(config, seed, npoints)
    return po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth if maxlevel is None else maxlevel)


def compare_frames(gpu, orc, rgb_tol=1.0 / 255.0):
    """Bit-exact hit/leaf/shadow/flags/isp, RGBA within rgb_tol (float) and
    1 unit (RGBA8). Returns a dict of mismatch counts."""
    res = {}
    for k in ("hit", "leaf", "shadow", "flags", "dyn"):
        if k in gpu:
            res[k] = int(np.count_nonzero(gpu[k] != orc[k]))
    if "isp" in gpu:
        res["isp"] = int(np.count_nonzero(np.any(gpu["isp"] != orc["isp"], axis=-1)))
    if "rgbaf" in gpu:
        d = np.abs(gpu["rgbaf"] - orc["rgbaf"])
        res["rgbaf"] = int(np.count_nonzero(np.any(d > rgb_tol, axis=-1)))
    if "rgba8" in gpu:
        a = gpu["rgba8"].view(np.uint8).reshape(gpu["rgba8"].shape + (4,)).astype(np.int16)
        b = orc["rgba8"].view(np.uint8).reshape(orc["rgba8"].shape + (4,)).astype(np.int16)
        res["rgba8"] = int(np.count_nonzero(np.any(np.abs(a - b) > 1, axis=-1)))
    if "exp" in gpu:
        res["exp"] = int(np.count_nonzero(gpu["exp"] != orc["exp"]))
    return res
