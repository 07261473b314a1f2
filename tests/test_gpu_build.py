"""GPU octree construction (qbt_build_octree) against the host builder, which
follows octree_insert_point (src/qubatron/octree.c:95-147) point by point:
the node arrays must be identical (same indices, same slot 8)."""
import numpy as np
import pytest

import qubatron_amd as q
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _gpu_build(ctx, pts, basecube, levels):
    ctx.upload_points(pts)
    n = ctx.build_octree(basecube, levels, pts.shape[0])
    return ctx.download_octree(n)


def _host_build(pts, basecube, levels):
    return q.octree_from_points(pts, basecube, levels).nodes()


@pytest.mark.parametrize("config,npoints", [(1, 0), (2, 2000000), (3, 3000000)])
def test_build_matches_host(gpu_ctx, config, npoints):
    sc = q.Scene(config, seed=0, npoints=npoints)
    pnt, _, _ = sc.points()
    print(f"[case] build config {config} n={pnt.shape[0]}", flush=True)
    g = _gpu_build(gpu_ctx, pnt, sc.basecube, sc.depth)
    h = _host_build(pnt, sc.basecube, sc.depth)
    assert g.shape == h.shape
    assert np.array_equal(g, h)


def test_build_special_points(gpu_ctx):
    """Duplicates (first inserter keeps the leaf), p == edge (truncation wraps
    to the low half, octree.c:104-109), small negatives (truncate toward 0),
    and points in reverse spatial order."""
    bc = np.array([0, 1800, 1800, 1800], np.float32)
    rng = np.random.default_rng(0)
    base = rng.uniform(0, 1800, (5000, 3)).astype(np.float32)
    pts = np.concatenate([base, base[:1000], np.array([[1800, 1800, 1800], [0, 0, 0], [-0.1, 5, 5],
                                                        [1799.9, -0.2, 1800]], np.float32), base[::-1]])
    for levels in (0, 1, 3, 12, 16, 20, 21):  # 21: 63-bit keys, the top octet at bits 60-62
        print(f"[case] special levels={levels}", flush=True)
        try:
            h = _host_build(pts, bc, levels)
        except q.QbtError as e:  # the small negatives reach an octet of -1 deep down
            assert e.code == q._ffi.QBT_ERANGE
            with pytest.raises(q.QbtError) as eg:
                _gpu_build(gpu_ctx, pts, bc, levels)
            assert eg.value.code == q._ffi.QBT_ERANGE
            continue
        g = _gpu_build(gpu_ctx, pts, bc, levels)
        assert np.array_equal(g, h), f"levels {levels}"


def test_build_single_and_empty(gpu_ctx):
    bc = np.array([0, 1800, 1800, 1800], np.float32)
    one = np.array([[10, 690, 10]], np.float32)
    g = _gpu_build(gpu_ctx, one, bc, 12)
    assert g.shape == (13, 12)
    assert np.array_equal(g, _host_build(one, bc, 12))
    gpu_ctx.upload_points(one)
    assert gpu_ctx.build_octree(bc, 12, 0) == 1


def test_build_rejects_bad_point(gpu_ctx):
    """x <= -size with odd y and z cells gives octet -1 in the reference."""
    bc = np.array([0, 1800, 1800, 1800], np.float32)
    pts = np.array([[-1000.0, 1000.0, 1000.0]], np.float32)
    with pytest.raises(q.QbtError):
        _host_build(pts, bc, 4)
    gpu_ctx.upload_points(pts)
    with pytest.raises(q.QbtError) as e:
        gpu_ctx.build_octree(bc, 4, 1)
    assert e.value.code == q._ffi.QBT_ERANGE


def test_build_then_render(gpu_ctx):
    """A GPU-built tree renders identically to the host-built one."""
    sc, pnt, nrm, col, nodes = H.scene(1)
    gpu_ctx.upload_points(pnt)
    gpu_ctx.upload_model(col, nrm)
    gpu_ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    n = gpu_ctx.build_octree(sc.basecube, sc.depth, pnt.shape[0])
    assert n == nodes.shape[0]
    v = q.view((900.0, 900.0, 900.0), (np.pi, 0.0), 320, 240, sc.basecube, sc.depth)
    gpu_ctx.render(v, sc.lights, q.OUT_ALL)
    g = gpu_ctx.read_frame(0)
    o = H.oracle_scene(1).render((900.0, 900.0, 900.0), (np.pi, 0.0), 320, 240, sc.lights)
    mm = H.compare_frames(g, o)
    assert all(v == 0 for k, v in mm.items() if k != "exp"), mm


def test_dynamic_rebuilds_reuse_scratch(gpu_ctx):
    """Per-frame rebuilds of the dynamic tree: the builder carves its scratch
    from blocks it keeps between builds, so builds of different sizes in a
    row (larger, smaller, larger again) must each equal the host build."""
    rng = np.random.default_rng(17)
    bc = np.array([0.0, 1800.0, 1800.0, 1800.0], np.float32)
    for n, levels in ((400000, 12), (30000, 9), (400000, 12), (120000, 11)):
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        pts = (np.array([900.0, 700.0, 900.0]) + u * np.array([120.0, 300.0, 120.0])).astype(np.float32)
        print(f"[case] dynamic rebuild n={n} levels={levels}", flush=True)
        gpu_ctx.upload_points(pts, dynamic=True)
        nn = gpu_ctx.build_octree(bc, levels, n, dynamic=True)
        g = gpu_ctx.download_octree(nn, dynamic=True)
        assert np.array_equal(g, _host_build(pts, bc, levels))
    gpu_ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)


def test_kept_build_scratch_yields_to_the_cap():
    """A dynamic-tree build keeps its scratch memory for the next per-frame
    rebuild; under max_device_bytes, a later allocation that would exceed the
    cap releases it instead of failing. The same build + render sequence under
    a cap 16 MiB below the uncapped run's final footprint gives the same frame
    and ends below the cap."""
    sc = q.Scene(2, seed=0, npoints=2000000)
    pnt, nrm, col = sc.points()
    rng = np.random.default_rng(4)
    m = 3000000
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    dp = (np.array([900.0, 700.0, 900.0]) + u * np.array([120.0, 300.0, 120.0])).astype(np.float32)
    dc = rng.uniform(0, 1, (m, 3)).astype(np.float32)
    v = q.view((900.0, 900.0, 300.0), (np.pi, 0.0), 1920, 1080, sc.basecube, sc.depth)

    def run(cap):
        with q.Context(0, max_device_bytes=cap) as c:
            c.upload_points(pnt)
            c.upload_model(col, nrm)
            c.build_octree(sc.basecube, sc.depth, pnt.shape[0])
            c.upload_points(dp, dynamic=True)
            c.upload_model(dc, u.astype(np.float32), dynamic=True)
            c.build_octree(sc.basecube, sc.depth, m, dynamic=True)
            held = c.device_bytes()
            c.render(v, sc.lights, q.OUT_DEFAULT | q.OUT_FLAGS)
            return held, c.device_bytes(), c.read_frame(0)

    print("[case] uncapped", flush=True)
    held0, end0, f0 = run(0)
    cap = end0 - (16 << 20)
    print(f"[case] cap {cap} (uncapped: {held0} after the builds, {end0} after the render)", flush=True)
    held1, end1, f1 = run(cap)
    assert held1 == held0  # the dynamic build kept its scratch under the cap too
    assert end1 <= cap
    for k in f0:
        assert np.array_equal(f0[k], f1[k]), k
    assert np.count_nonzero(f0["hit"] >= 0) > 100000
