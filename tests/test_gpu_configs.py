"""Parity at the benchmark workloads: the scenes of configs 2, 3 and 4 at
their full point counts, depths, resolutions, light counts and batch shapes,
rendered through the C ABI and compared with the CPU oracle on the same node
array (the pixel program of octree_fsh.c:399-464).

Bar (as tests/test_gpu_parity.py): hit model index, leaf node index, shadow
bits, pixel flags and the entry point bit-exact, colour within 1/255 (float)
and 1 RGBA8 unit, expansion counts equal in counting renders. Whole frames
where the oracle finishes in seconds, row bands elsewhere (stated per test).
The octree is built on the GPU from the scene's points (qbt_build_octree)
and downloaded for the oracle; the builder itself is checked here against
the host builder (octree.c:95-147 numbering) at these sizes too.

The learned shadow surrogate at the config-5 workload is in
tests/test_gpu_surrogate_full.py.
"""
import os
import time

import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
COUNT = q.OUT_ALL | q.RENDER_COUNT
EXACT = q.OUT_HIT | q.OUT_LEAF | q.OUT_SHADOW | q.OUT_RGBA8 | q.OUT_RGBAF | q.OUT_ISP | q.OUT_FLAGS
BATCH_RES = ((320, 240), (640, 480), (1280, 720), (1920, 1080), (3840, 2160))


class Loaded:
    """A config's scene generated, uploaded, built on the GPU and downloaded
    for the oracle."""

    def __init__(self, config):
        t0 = time.time()
        self.sc = q.Scene(config)
        self.pnt, self.nrm, self.col = self.sc.points(nthreads=NT)
        self.ctx = q.Context(0)
        self.ctx.upload_points(self.pnt)
        self.ctx.upload_model(self.col, self.nrm)
        self.ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
        self.nnodes = self.ctx.build_octree(self.sc.basecube, self.sc.depth, self.sc.npoints)
        self.nodes = self.ctx.download_octree(self.nnodes)
        self.osc = po.OracleScene(self.nodes, self.col, self.nrm, self.sc.basecube, self.sc.depth)
        print(f"[scene] config {config}: {self.sc.npoints} points, {self.nnodes} nodes, "
              f"{len(self.sc.lights)} lights, ready in {time.time() - t0:.1f}s", flush=True)

    def view(self, pos, ang, w, h):
        return q.view(pos, ang, w, h, self.sc.basecube, self.sc.depth)

    def close(self):
        self.ctx.close()
        self.pnt = self.nrm = self.col = self.nodes = self.osc = None


_current = {}


def _get(config):
    """the scene of `config`, loaded once; the previous config's is released
    first (one 90M-point scene in memory at a time)"""
    if config not in _current:
        for s in _current.values():
            s.close()
        _current.clear()
        _current[config] = Loaded(config)
    return _current[config]


@pytest.fixture(scope="module", autouse=True)
def _release_scenes():
    yield
    for s in _current.values():
        s.close()
    _current.clear()


def _compare(g, o, what, rows=None, rgb_tol=1.0 / 255.0):
    if rows is not None:
        r0, r1 = rows
        g = {k: v[r0:r1] for k, v in g.items()}
        o = {k: v[r0:r1] for k, v in o.items()}
    mm = H.compare_frames(g, o, rgb_tol)
    assert all(v == 0 for v in mm.values()), f"{what}: mismatches {mm}"


def _bands(h, n=3, rows=16):
    """n bands of `rows` rows: bottom, middle, top"""
    starts = [0, (h - rows) // 2, h - rows][:n]
    return [(s, s + rows) for s in starts]


# ------------------------------------------------------------- config 2 ----

@pytest.fixture
def cfg2():
    return _get(2)


def test_config2_builder_matches_host(cfg2):
    """The GPU builder's node array equals the host builder's (insertion in
    model-index order, octree.c:95-147) node for node at 10M points."""
    t0 = time.time()
    host = q.octree_from_points(cfg2.pnt, cfg2.sc.basecube, cfg2.sc.depth).nodes()
    print(f"[case] config 2 host build: {host.shape[0]} nodes in {time.time() - t0:.1f}s", flush=True)
    assert host.shape == cfg2.nodes.shape
    assert np.array_equal(host, cfg2.nodes)


def test_config2_full_frames(cfg2):
    """Config 2 (10M points, depth 12, 1920x1080, one light): three poses of
    the seeded 64-pose path, whole frames, counting renders."""
    pos, ang = cfg2.sc.poses(64, seed=1)
    for i in range(3):
        print(f"[case] config 2 pose {i} 1920x1080 whole frame", flush=True)
        cfg2.ctx.render(cfg2.view(pos[i], ang[i], 1920, 1080), cfg2.sc.lights, COUNT)
        g = cfg2.ctx.read_frame(0)
        st = cfg2.ctx.stats()
        o = cfg2.osc.render(pos[i], ang[i], 1920, 1080, cfg2.sc.lights, nthreads=NT)
        _compare(g, o, f"config 2 pose {i}")
        assert st["expansions"] == int(o["exp"].sum())
        assert st["hits"] == int(np.count_nonzero(o["flags"] & po.PX_SHADED))
        assert (o["hit"] >= 0).mean() > 0.05  # depth-12 leaves leave gaps in 10M-point surfaces


def test_config2_all_64_poses_bands(cfg2):
    """All 64 poses of the path at 1920x1080: bottom, middle and top bands of
    16 rows of every frame against the oracle."""
    pos, ang = cfg2.sc.poses(64, seed=1)
    for i in range(64):
        cfg2.ctx.render(cfg2.view(pos[i], ang[i], 1920, 1080), cfg2.sc.lights, EXACT)
        g = cfg2.ctx.read_frame(0)
        for r0, r1 in _bands(1080):
            o = cfg2.osc.render(pos[i], ang[i], 1920, 1080, cfg2.sc.lights, rows=(r0, r1), nthreads=NT)
            _compare(g, o, f"config 2 pose {i} rows {r0}-{r1}", rows=(r0, r1))
        if i % 16 == 15:
            print(f"[case] config 2 poses 0..{i} banded: equal", flush=True)


# ------------------------------------------------------------- config 3 ----

@pytest.fixture
def cfg3():
    return _get(3)


def test_config3_builder_prefix_matches_host(cfg3):
    """Inserting points in index order only appends nodes and fills empty
    slots, so the host tree of the first 4M of the 90M points is a prefix of
    the GPU tree of all of them: same slot 8 for its nodes, and every child
    slot it has set holds the same node in the full tree."""
    k = 4000000
    pre = q.octree_from_points(cfg3.pnt[:k], cfg3.sc.basecube, cfg3.sc.depth).nodes()
    full = cfg3.nodes[:pre.shape[0]]
    assert np.array_equal(full[:, 8], pre[:, 8])
    set_ = pre[:, :8] != 0
    assert np.array_equal(full[:, :8][set_], pre[:, :8][set_])
    assert cfg3.nnodes > 150_000_000


def test_config3_1080p_full_frames(cfg3):
    """Config 3 (90M points, depth 12, 1920x1080, two lights): three poses of
    the timed camera path, whole frames, counting renders."""
    pos, ang = cfg3.sc.poses(3, seed=1)
    for i in range(3):
        print(f"[case] config 3 pose {i} 1920x1080 whole frame", flush=True)
        cfg3.ctx.render(cfg3.view(pos[i], ang[i], 1920, 1080), cfg3.sc.lights, COUNT)
        g = cfg3.ctx.read_frame(0)
        st = cfg3.ctx.stats()
        o = cfg3.osc.render(pos[i], ang[i], 1920, 1080, cfg3.sc.lights, nthreads=NT)
        _compare(g, o, f"config 3 pose {i}")
        assert st["expansions"] == int(o["exp"].sum())
        assert st["shadow_rays"] == 2 * int(np.count_nonzero(o["flags"] & po.PX_SHADED))


def test_config3_2160p_full_frame(cfg3):
    """Config 3 at 3840x2160: one pose of the 2160p path, whole frame."""
    pos, ang = cfg3.sc.poses(2, seed=2161)
    print("[case] config 3 3840x2160 whole frame", flush=True)
    cfg3.ctx.render(cfg3.view(pos[1], ang[1], 3840, 2160), cfg3.sc.lights, COUNT)
    g = cfg3.ctx.read_frame(0)
    o = cfg3.osc.render(pos[1], ang[1], 3840, 2160, cfg3.sc.lights, nthreads=NT)
    _compare(g, o, "config 3 2160p")
    assert cfg3.ctx.stats()["expansions"] == int(o["exp"].sum())


# ------------------------------------------------------------- config 4 ----

@pytest.fixture
def cfg4():
    return _get(4)


def _batch(cfg4, n=64, seed=5):
    rng = np.random.default_rng(seed)
    sizes = [BATCH_RES[i] for i in rng.integers(0, len(BATCH_RES), n)]
    pos, ang = cfg4.sc.poses(n, seed=seed)
    views = [cfg4.view(pos[i], ang[i], w, h) for i, (w, h) in enumerate(sizes)]
    return sizes, pos, ang, views


def _checked_views(sizes):
    """two views of every resolution class present"""
    out = []
    for r in BATCH_RES:
        out += [i for i, s in enumerate(sizes) if s == r][:2]
    return out


@pytest.mark.parametrize("tuning", [dict(reorder=0, graph=1), dict(reorder=1, graph=1)],
                         ids=["coherent", "octant-reordered"])
def test_config4_batch_of_64_views(cfg4, tuning):
    """Config 4 (90M points, four lights): 64 views of resolutions drawn from
    320x240 ... 3840x2160 rendered as ONE batch (one launch sequence,
    captured into a hipGraph and replayed), with the shadow rays in the
    coherent order and reordered by direction octant. Two views of every
    resolution class: whole frames up to 1280x720, three 16-row bands
    above. Replaying the captured graph gives the same frames."""
    ctx = cfg4.ctx
    ctx.set_tuning(**tuning)
    try:
        sizes, pos, ang, views = _batch(cfg4)
        lights = cfg4.sc.lights
        assert len(lights) == 4
        ctx.render(views, lights, EXACT)  # capture
        ctx.render(views, lights, EXACT)  # replay
        st = ctx.stats()
        assert st["primary_rays"] == sum(w * h for w, h in sizes)
        for i in _checked_views(sizes):
            w, h = sizes[i]
            g = ctx.read_frame(i)
            bands = [None] if w * h <= 1280 * 720 else _bands(h)
            for rows in bands:
                print(f"[case] config 4 {tuning} view {i} {w}x{h} rows {rows}", flush=True)
                o = cfg4.osc.render(pos[i], ang[i], w, h, lights, rows=rows, nthreads=NT)
                _compare(g, o, f"config 4 view {i}", rows=rows)
    finally:
        ctx.set_tuning(reorder=0, graph=1)


def test_config4_autotuned_batch(cfg4):
    """The autotuner (block size, rays per refill, stack placement, octant
    order) picks a tuning on 8 of the views; the batch rendered with it equals
    the default tuning's batch on the checked views, bit for bit."""
    ctx = cfg4.ctx
    sizes, pos, ang, views = _batch(cfg4)
    lights = cfg4.sc.lights
    ctx.set_tuning(block_threads=0, refill=0, stack=0, reorder=0, graph=1)
    ctx.render(views, lights, EXACT)
    ref = {i: ctx.read_frame(i) for i in _checked_views(sizes)}
    best, table = ctx.autotune(views[:8], lights, EXACT, reps=1)
    print(f"[case] config 4 autotuned: {best} of {len(table)} candidates", flush=True)
    try:
        ctx.set_tuning(graph=1)
        ctx.render(views, lights, EXACT)
        for i, r in ref.items():
            g = ctx.read_frame(i)
            for k in r:
                assert np.array_equal(g[k], r[k]), f"view {i} {k}"
    finally:
        ctx.set_tuning(block_threads=0, refill=0, stack=0, reorder=0, graph=1)
