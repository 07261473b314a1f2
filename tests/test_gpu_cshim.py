"""The reference's renderer connector replaced by the C shim a maintainer
adds to the reference's build (tests/cshim/octree_glc_qbt.c), driven by a compiled C caller
(tests/cshim/glc_demo.c) that replays the reference's static-scene calls:
load .pnt/.col/.nrm flat files (model.c:53-111), insert the points in order
(modelutil.c:203-217), upload colours, normals and the octree
(modelutil.c:227-255), octree_glc_update, then the punch-hole edit pattern
(modelutil.c:405-500: single 48-byte node uploads of zeroed slots, then
uploads of every node a re-insertion touched) and octree_glc_update again.

The flat files are written by the qmc step (qbt_qmc_quantize, qmc.c:41-165,
then qbt_flat_write, qmc.c:291-327) from a seeded scene. Both frames are
compared with the oracle on the same arrays: hit, leaf and shadow bits
bit-exact, RGBA8 within 1 unit. The edited node array is replayed
independently with the host octree API and must be identical.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import qubatron_amd as q
from qubatron_amd import _ffi
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cshim", "glc_demo")


def _qmc_flat(tmp_path, pnt, col, nrm, size, levels):
    p, c, n = (np.ascontiguousarray(a, np.float32).copy() for a in (pnt, col, nrm))
    rng = np.zeros(p.shape[0] * 3 + 3, np.int32)
    nr = C.c_size_t(0)
    k = _ffi.host().qbt_qmc_quantize(_ffi.fptr(p), _ffi.fptr(c), _ffi.fptr(n), p.shape[0], size, levels,
                                     _ffi.iptr(rng), C.byref(nr))
    assert k > 0
    p, c, n = p[:k], c[:k], n[:k]
    paths = [str(tmp_path / f"scene.{e}") for e in ("pnt", "col", "nrm")]
    assert _ffi.host().qbt_flat_write(*(x.encode() for x in paths), _ffi.fptr(p), _ffi.fptr(c), _ffi.fptr(n),
                                      k) == 0
    return paths


def test_flat_files_through_the_c_shim(tmp_path):
    if not os.path.exists(DEMO):
        pytest.fail(f"{DEMO} is not built: run __graft_entry__.build()")
    sc = q.Scene(1, seed=3)  # 1M points: closed surfaces at depth 10
    pnt, nrm, col = sc.points()
    depth = sc.depth  # 10: leaf edge 1800 / 1024
    paths = _qmc_flat(tmp_path, pnt, col, nrm, 1800.0, depth - 1)
    # what model_load_flat reads: raw little-endian float32 triples
    fp, fc, fn = (np.fromfile(x, dtype="<f4").reshape(-1, 3) for x in paths)
    cam, ang, lighta, quality, shoot = (900.0, 900.0, 900.0), (np.pi, 0.1), 0.7, 6, 1
    centre, radius = np.float32([420.0, 1294.0, 1799.0]), np.float32(80.0)  # on the wall in view
    out = str(tmp_path / "frames.bin")
    args = [DEMO, *paths, out, "640", "480", *(repr(float(x)) for x in cam), repr(float(ang[0])),
            repr(float(ang[1])), repr(lighta), str(quality), str(depth), "1800", str(shoot),
            *(repr(float(x)) for x in centre), repr(float(radius))]
    print("[case] glc_demo", " ".join(args[5:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip(), flush=True)

    raw = np.fromfile(out, dtype=np.uint8)
    w, h, nnodes, nrem = raw[:16].view(np.int32)
    n = int(w) * int(h)
    off = 16
    frames = []
    for _ in range(2):
        f = {}
        for k, dt, sz in (("hit", np.int32, 4), ("leaf", np.int32, 4), ("shadow", np.uint8, 1),
                          ("rgba8", np.uint32, 4)):
            f[k] = raw[off:off + n * sz].view(dt).reshape(h, w)
            off += n * sz
        frames.append(f)
    edited = raw[off:off + int(nnodes) * 48].view(np.int32).reshape(-1, 12)
    assert off + int(nnodes) * 48 == raw.size
    assert (w, h) == (213, 160) and nrem > 100

    # the same edit on the host octree API, independently of the C caller
    t = q.octree_from_points(fp, sc.basecube, depth)
    before = t.nodes()
    d = fp - centre
    near = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) < radius * radius
    moved = []
    for i in np.flatnonzero(near):
        m, o = t.remove_point(fp[i])
        if o >= 0:
            moved.append((m, fp[i] + fn[m] * np.float32(5.0)))
    assert len(moved) == nrem
    for m, p in moved:
        t.insert_point(m, p.astype(np.float32), octindarr=np.zeros(24, np.int32))
    assert np.array_equal(t.nodes(), edited)

    # the light and dimensions octree_glc_update derives (octree_glc.c:263-270)
    libm = C.CDLL("libm.so.6")
    libm.sinf.argtypes, libm.sinf.restype = [C.c_float], C.c_float
    s = float(libm.sinf(lighta))
    light = np.array([420.0, np.float32(200.0 - s * 20.0), np.float32(680.0 - s * 200.0)], np.float32)
    W, Hh = np.float32(640.0 / 3.0), np.float32(480.0 / 3.0)
    for i, nodes in enumerate((before, edited)):
        o = po.OracleScene(nodes, fc, fn, sc.basecube, depth).render(cam, ang, W, Hh, light, shoot=shoot)
        mm = H.compare_frames(frames[i], o)
        assert all(v == 0 for v in mm.values()), f"frame {i + 1}: {mm}"
        assert (o["hit"] >= 0).mean() > 0.03
    assert np.count_nonzero(frames[0]["leaf"] != frames[1]["leaf"]) > 50  # the edit is in view
