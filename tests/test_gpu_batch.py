"""Batched rendering on the GPU: many views of uneven sizes in one call (one
set of persistent launches), its hipGraph capture and replays, the stack
placement and shadow-ray reordering variants, and views with different
basecubes / maxlevels in one call. Every frame must equal the oracle's, or
the single-view render that the oracle already pins, bit for bit.
"""
import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

LIGHTS4 = np.array([[880.25, 1700, 915.75], [300, 1500, 1400], [1500, 1600, 300], [1450, 950, 1500]], np.float32)
SIZES = [(64, 48), (33, 17), (100, 75), (1, 1), (128, 96), (8, 8), (72, 40)]
FIELDS = ("hit", "leaf", "shadow", "flags", "isp", "rgba8", "dyn")
MASK = q.OUT_HIT | q.OUT_LEAF | q.OUT_SHADOW | q.OUT_FLAGS | q.OUT_ISP | q.OUT_RGBA8 | q.OUT_DYN


@pytest.fixture
def scene_ctx(gpu_ctx):
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=300000)
    gpu_ctx.upload_octree(nodes)
    gpu_ctx.upload_model(col, nrm)
    gpu_ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    yield gpu_ctx, sc, po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)
    gpu_ctx.set_tuning(path=1, layout=0, stack=0, reorder=0, graph=1, strip=0, merge=1)


def _views(sc, seed, sizes=SIZES, maxlevel=None):
    pos, ang = sc.poses(len(sizes), seed=seed)
    L = sc.depth if maxlevel is None else maxlevel
    return [q.view(pos[i], ang[i], w, h, sc.basecube, L, shoot=i % 2) for i, (w, h) in enumerate(sizes)], pos, ang


def _frames(ctx, n):
    return [ctx.read_frame(i) for i in range(n)]


def _same(a, b, what):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs ({np.count_nonzero(a[k] != b[k])} elements)"


@pytest.mark.parametrize("stack,reorder,strip,merge", [(0, 1, 0, 2), (0, 0, 0, 2), (0, 0, 0, 0), (1, 1, 0, 2),
                                                      (1, 0, 0, 2), (1, 0, 0, 0), (0, 0, 2, 2), (0, 0, 5, 1)])
def test_batch_equals_oracle(scene_ctx, stack, reorder, strip, merge):
    ctx, sc, osc = scene_ctx
    ctx.set_tuning(path=1, stack=stack, reorder=reorder, graph=0, strip=strip, merge=merge)
    views, pos, ang = _views(sc, 21)
    print(f"[case] batch of {len(views)} views stack={stack} reorder={reorder} merge={merge}", flush=True)
    ctx.render(views, LIGHTS4, MASK)
    st = ctx.stats()
    assert st["primary_rays"] == sum(w * h for w, h in SIZES)
    if not merge or reorder:  # the merged launch runs only for exact renders without reordering
        assert st["shadow_deferred"] == 0
    for i, (w, h) in enumerate(SIZES):
        o = osc.render(pos[i], ang[i], w, h, LIGHTS4, shoot=i % 2)
        _same(ctx.read_frame(i), o, f"view {i} {w}x{h}")


@pytest.mark.parametrize("merge", [1, 2])
def test_graph_replays_match_direct(scene_ctx, merge):
    """Capture once, replay for other poses of the same shape, with direct
    (uncaptured) renders in between; every replay equals the direct render.
    merge 2: the captured sequence is the merged camera-and-shadow launch,
    whose hand-over slots are reset by a kernel of the graph on every replay."""
    ctx, sc, _ = scene_ctx
    ctx.set_tuning(merge=merge)
    for rep in range(4):
        views, _, _ = _views(sc, 100 + rep)
        ctx.set_tuning(graph=0)
        print(f"[case] direct render {rep} merge={merge}", flush=True)
        ctx.render(views, LIGHTS4, MASK)
        ref = _frames(ctx, len(views))
        ctx.set_tuning(graph=1)
        print(f"[case] graph render {rep} merge={merge}", flush=True)
        ctx.render(views, LIGHTS4, MASK)
        got = _frames(ctx, len(views))
        for i in range(len(views)):
            _same(got[i], ref[i], f"replay {rep} view {i}")


def test_graph_recaptures_on_shape_change(scene_ctx):
    ctx, sc, osc = scene_ctx
    ctx.set_tuning(graph=1)
    for sizes, lights in (([(40, 30), (20, 10)], LIGHTS4[:2]), ([(40, 30), (20, 10)], LIGHTS4),
                          ([(24, 24)], LIGHTS4[1:3]), ([(40, 30), (20, 10)], LIGHTS4[:2])):
        views, pos, ang = _views(sc, 7, sizes)
        print(f"[case] graph shape {sizes} lights {len(lights)}", flush=True)
        ctx.render(views, lights, MASK)
        for i, (w, h) in enumerate(sizes):
            _same(ctx.read_frame(i), osc.render(pos[i], ang[i], w, h, lights, shoot=i % 2), f"{sizes} view {i}")


def test_mixed_maxlevel_views_split_into_batches(scene_ctx):
    ctx, sc, _ = scene_ctx
    pos, ang = sc.poses(4, seed=3)
    views = [q.view(pos[i], ang[i], 48, 32, sc.basecube, sc.depth - (i % 2)) for i in range(4)]
    print("[case] alternating maxlevel", flush=True)
    ctx.render(views, LIGHTS4[:2], MASK)
    for i in range(4):
        osc = H.oracle_scene(1, npoints=300000, maxlevel=sc.depth - (i % 2))
        _same(ctx.read_frame(i), osc.render(pos[i], ang[i], 48, 32, LIGHTS4[:2]), f"view {i}")


def test_fused_path_batch(scene_ctx):
    ctx, sc, osc = scene_ctx
    ctx.set_tuning(path=0)
    views, pos, ang = _views(sc, 31, SIZES[:4])
    print("[case] fused path, 4 views", flush=True)
    ctx.render(views, LIGHTS4, MASK)
    for i, (w, h) in enumerate(SIZES[:4]):
        _same(ctx.read_frame(i), osc.render(pos[i], ang[i], w, h, LIGHTS4, shoot=i % 2), f"view {i}")


def test_autotune_picks_a_valid_tuning(scene_ctx):
    ctx, sc, osc = scene_ctx
    views, pos, ang = _views(sc, 41, SIZES[:3])
    print("[case] autotune on 3 views", flush=True)
    best, table = ctx.autotune(views, LIGHTS4[:2], MASK, reps=2)
    assert len(table) >= 8
    assert best["path"] == 1 and best["block_threads"] in (64, 128, 256)
    assert best["stack"] in (0, 1) and best["reorder"] in (0, 1) and best["merge"] in (0, 1, 2)
    assert min(ms for _, ms in table) > 0
    ctx.render(views, LIGHTS4[:2], MASK)
    for i, (w, h) in enumerate(SIZES[:3]):
        _same(ctx.read_frame(i), osc.render(pos[i], ang[i], w, h, LIGHTS4[:2], shoot=i % 2), f"view {i}")
