"""Parity of the gfx950 traversal/shading (libqbt.so, called through its C ABI)
with the CPU oracle on the same seeded inputs.

Bar: hit model index, leaf node index, dynamic model index, shadow bits, pixel
flags and the primary entry point (x, y, z, t) bit-exact; colour within 1/255
(float) and 1 unit of RGBA8; expansion counts equal.
"""
import ctypes

import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

ALL = q.OUT_ALL | q.RENDER_COUNT


@pytest.fixture(params=[dict(path=1, layout=0), dict(path=1, layout=1), dict(path=0, layout=1),
                        dict(path=1, layout=0, stack=1, reorder=0, strip=3, merge=2), dict(path=1, layout=0, merge=2)],
                ids=["wavefront-compact", "wavefront-direct", "fused", "wavefront-gstack-merged", "wavefront-merged"])
def ctx(gpu_ctx, request):
    """Every render path: wavefront stages on the compact node layout
    (default tuning: these small frames take separate camera and shadow
    launches), wavefront on the direct layout, the fused kernel, and the
    merged camera-and-shadow launch, on LDS stacks and on global-memory
    stacks with tiles in 3-tile strips."""
    gpu_ctx.set_tuning(**request.param)
    yield gpu_ctx
    gpu_ctx.set_tuning(path=1, layout=0, stack=0, reorder=0, graph=1, strip=0, merge=1)


def test_shared_child_falls_back_to_direct(ctx):
    """A node array in which two slots point at the same node is not a tree:
    the compact relayout refuses it and the direct layout renders it, still
    equal to the oracle (which follows the slots as the shader does)."""
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=50000)
    nodes = nodes.copy()
    root = nodes[0]
    kids = [o for o in range(8) if root[o] > 0]
    empty = [o for o in range(8) if root[o] == 0]
    if empty:
        nodes[0, empty[0]] = nodes[0, kids[0]]  # a second root slot at the same subtree
    else:
        nodes[0, kids[1]] = nodes[0, kids[0]]
    _upload_scene(ctx, nodes, col, nrm)
    osc = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 900.0), (2.3, 0.2), 96, 64, sc.lights, sc.basecube, sc.depth)
    assert ctx.stats()["compact"] == 0
    _assert_same(g, o, "shared child")


def _render_both(ctx, osc, camfp, angle, W, H_, lights, basecube, maxlevel, shoot=0):
    v = q.view(camfp, angle, W, H_, basecube, maxlevel, shoot)
    print(f"[case] render cam={tuple(np.round(camfp, 2))} ang={tuple(np.round(angle, 3))} {W}x{H_} L={maxlevel}",
          flush=True)
    ctx.render(v, lights, ALL)
    g = ctx.read_frame(0)
    o = osc.render(camfp, angle, W, H_, lights, shoot=shoot)
    return g, o


def _assert_same(g, o, what):
    mm = H.compare_frames(g, o)
    assert all(v == 0 for v in mm.values()), f"{what}: mismatches {mm}"


def _upload_scene(ctx, nodes, col, nrm, dyn_nodes=None, dyn_col=None, dyn_nrm=None):
    ctx.upload_octree(nodes)
    ctx.upload_model(col, nrm)
    if dyn_nodes is None:
        dyn_nodes = np.zeros((1, 12), np.int32)
    ctx.upload_octree(dyn_nodes, dynamic=True)
    if dyn_col is not None:
        ctx.upload_model(dyn_col, dyn_nrm, dynamic=True)


def test_config1_reference_view(ctx):
    """Config 1: 1M points, depth 10, 320x240, camera at the centre looking +z."""
    sc, pnt, nrm, col, nodes = H.scene(1)
    _upload_scene(ctx, nodes, col, nrm)
    osc = H.oracle_scene(1)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 900.0), (np.pi, 0.0), 320, 240, sc.lights, sc.basecube,
                        sc.depth)
    _assert_same(g, o, "config1")
    st = ctx.stats()
    assert st["primary_rays"] == 320 * 240
    assert st["hits"] == int(np.count_nonzero(o["flags"] & po.PX_SHADED))
    assert st["expansions"] == int(o["exp"].sum())
    assert st["grid_exact"] == 1
    assert (o["hit"] >= 0).mean() > 0.1  # the view sees surfaces


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_config1_random_poses(ctx, seed):
    sc, pnt, nrm, col, nodes = H.scene(1)
    _upload_scene(ctx, nodes, col, nrm)
    osc = H.oracle_scene(1)
    pos, ang = sc.poses(2, seed=seed)
    for i in range(2):
        g, o = _render_both(ctx, osc, pos[i], ang[i], 160, 120, sc.lights, sc.basecube, sc.depth,
                            shoot=i)
        _assert_same(g, o, f"pose {seed}/{i}")


def test_multi_light_and_views(ctx):
    """Two views in one call, four lights (multi-light extension)."""
    sc, pnt, nrm, col, nodes = H.scene(1)
    _upload_scene(ctx, nodes, col, nrm)
    osc = H.oracle_scene(1)
    lights = np.array([[900, 1700, 900], [300, 1500, 1400], [1500, 1600, 300], [1450, 900, 1500]], np.float32)
    pos, ang = sc.poses(2, seed=11)
    views = [q.view(pos[i], ang[i], 96 + 32 * i, 64 + 16 * i, sc.basecube, sc.depth) for i in range(2)]
    print("[case] two views, four lights", flush=True)
    ctx.render(views, lights, ALL)
    for i in range(2):
        g = ctx.read_frame(i)
        o = osc.render(pos[i], ang[i], 96 + 32 * i, 64 + 16 * i, lights)
        _assert_same(g, o, f"view {i}")


def test_non_grid_basecube(ctx):
    """A basecube whose subdivision is not exact in fp32: the running-sum cube
    path replays the reference's additions."""
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=200000)
    bc = np.array([0.3, 1000.7, 1001.1, 1000.7], np.float32)
    lo = np.array([bc[0], bc[1] - bc[3], bc[2] - bc[3]], np.float32)
    p2 = (pnt * np.float32(bc[3] / 1800.0)).astype(np.float32) + lo
    p2 = np.clip(p2, lo + 0.01, lo + bc[3] - 0.01).astype(np.float32)
    tree = q.octree_from_points(p2, bc, 9)
    nodes2 = tree.nodes()
    _upload_scene(ctx, nodes2, col, nrm)
    osc = po.OracleScene(nodes2, col, nrm, bc, 9)
    cam = np.array([500.0, 500.0, 500.0], np.float32)
    g, o = _render_both(ctx, osc, cam, (2.0, 0.3), 128, 96, [[500.0, 950.0, 500.0]], bc, 9)
    assert ctx.stats()["grid_exact"] == 0
    _assert_same(g, o, "non-grid")


def test_dual_tree(ctx):
    """Static scene + a populated dynamic tree traversed in lockstep
    (octree_fsh.c:223-241, 313-317, 355-356)."""
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=300000)
    rng = np.random.default_rng(5)
    # a dense "figure" of dynamic points in front of the camera + a few inside static walls
    n_d = 60000
    dp = np.stack([rng.uniform(820, 980, n_d), rng.uniform(700, 1100, n_d), rng.uniform(1150, 1250, n_d)], 1)
    dp = dp.astype(np.float32)
    dn = np.tile(np.array([[0, 0, -1]], np.float32), (n_d, 1))
    dc = rng.uniform(0, 1, (n_d, 3)).astype(np.float32)
    dtree = q.octree_from_points(dp, sc.basecube, sc.depth)
    dnodes = dtree.nodes()
    _upload_scene(ctx, nodes, col, nrm, dnodes, dc, dn)
    osc = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth, dyn_nodes=dnodes, dyn_col=dc, dyn_nrm=dn)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 700.0), (np.pi, 0.1), 200, 150, sc.lights, sc.basecube,
                        sc.depth)
    assert ctx.stats()["dual"] == 1
    _assert_same(g, o, "dual")
    assert np.count_nonzero(o["dyn"] > 0) > 100  # dynamic points are visible


def test_camera_outside_and_discard(ctx):
    """Camera outside the cube (rays entering through faces, discards at the
    root, OCTTEST camera of qubatron.c:140) on the 5-point OCTTEST scene
    (modelutil.c:91-110) at depth 12."""
    pts = np.array([[10, 690, 10], [10, 340, 10], [10, 340, 690], [10, 10, 10], [690, 10, 690]], np.float32)
    nrm = np.tile(np.array([[0, 0, -1]], np.float32), (5, 1))
    col = np.ones((5, 3), np.float32)
    bc = np.array([0.0, 1800.0, 1800.0, 1800.0], np.float32)
    tree = q.octree_from_points(pts, bc, 12)
    nodes = tree.nodes()
    _upload_scene(ctx, nodes, col, nrm)
    osc = po.OracleScene(nodes, col, nrm, bc, 12)
    for cam, ang in [((900.0, 900.0, 3000.0), (0.0, 0.0)), ((-500.0, 2500.0, 900.0), (1.2, -0.5)),
                     ((10.2, 690.2, 2000.0), (0.0, 0.0))]:
        g, o = _render_both(ctx, osc, cam, ang, 128, 128, [[420.0, 200.0, 680.0]], bc, 12)
        _assert_same(g, o, f"outside {cam}")


@pytest.mark.parametrize("order", ["discard-second", "discard-first", "discard-middle-of-three"])
def test_shadow_discard_by_any_light(ctx, order):
    """A shadow ray that the root rejects discards the fragment
    (octree_fsh.c:195,198): shading stops at the first such light, later
    lights are not traced (octree_fsh.c:431-434 run per light in order).
    Scene: a layer of leaves under the cube's top face, seen from above, so
    every entry point has y = 1800 exactly; a light at y = 1800 outside the
    cube sends rays parallel to that face, which no face test accepts. The
    discarding light sits first, second, or between two others; the frame
    and the expansion counts must not depend on which rays end first."""
    levels = 8
    u = 1800.0 / 2 ** levels
    g = (np.arange(40, 216) + 0.5) * u
    xx, zz = np.meshgrid(g, g)
    pts = np.stack([xx.ravel(), np.full(xx.size, 1800.0 - 0.5 * u), zz.ravel()], 1).astype(np.float32)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (pts.shape[0], 1))
    col = np.random.default_rng(3).uniform(0.2, 1.0, pts.shape).astype(np.float32)
    bc = np.array([0.0, 1800.0, 1800.0, 1800.0], np.float32)
    nodes = q.octree_from_points(pts, bc, levels).nodes()
    _upload_scene(ctx, nodes, col, nrm)
    osc = po.OracleScene(nodes, col, nrm, bc, levels)
    l_in, l_in2, l_disc = [900.0, 1000.0, 900.0], [400.0, 1300.0, 1200.0], [-300.0, 1800.0, 900.0]
    lights = {"discard-second": [l_in, l_disc], "discard-first": [l_disc, l_in],
              "discard-middle-of-three": [l_in, l_disc, l_in2]}[order]
    for rep in range(3):
        gg, o = _render_both(ctx, osc, (900.0, 2700.0, 1500.0), (0.0, -1.1), 96, 80, lights, bc, levels)
        _assert_same(gg, o, f"{order} #{rep}")
        assert ctx.stats()["expansions"] == int(o["exp"].sum())
    leaf, disc = (o["flags"] & po.PX_LEAF) != 0, (o["flags"] & po.PX_DISCARD) != 0
    assert np.count_nonzero(leaf & disc) > 2000  # every leaf pixel: its shadow ray discards


@pytest.mark.parametrize("maxlevel", [0, 1, 2, 5])
def test_shallow_trees(ctx, maxlevel):
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=5000)
    tree = q.octree_from_points(pnt, sc.basecube, maxlevel)
    nodes = tree.nodes()
    _upload_scene(ctx, nodes, col, nrm)
    osc = po.OracleScene(nodes, col, nrm, sc.basecube, maxlevel)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 900.0), (0.7, 0.2), 64, 48, sc.lights, sc.basecube, maxlevel)
    _assert_same(g, o, f"maxlevel {maxlevel}")


def test_empty_tree(ctx):
    col = np.zeros((1, 3), np.float32)
    nodes = np.zeros((1, 12), np.int32)
    _upload_scene(ctx, nodes, col, col)
    osc = po.OracleScene(nodes, col, col, (0, 1800, 1800, 1800), 10)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 900.0), (0.0, 0.0), 32, 32, [[900, 1700, 900]],
                        (0, 1800, 1800, 1800), 10)
    _assert_same(g, o, "empty")
    assert np.all(g["hit"] == -1)


@pytest.mark.parametrize("semantics", [0, 1])
def test_trace_rays(ctx, semantics):
    """Batched octree_trace_line: random rays, plus axis-aligned rays whose
    zero direction components exercise the two plane-test conventions."""
    sc, pnt, nrm, col, nodes = H.scene(1)
    _upload_scene(ctx, nodes, col, nrm)
    osc = H.oracle_scene(1)
    rng = np.random.default_rng(7 + semantics)
    n = 20000
    org = rng.uniform(-200, 2000, (n, 3)).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32) * 10
    k = n // 4
    ax = rng.integers(0, 3, k)
    dirs[:k] = 0
    dirs[np.arange(k), ax] = rng.choice([-10.0, 10.0], k)
    org[:k // 2] = rng.uniform(1, 1799, (k // 2, 3))
    # origins exactly on the y=0 / z=0 edge region (the CPU twin's (0,0,0,FLT_MAX) quirk)
    org[k // 2:k // 2 + 200, 1] = rng.uniform(0, 5, 200)
    org[k // 2:k // 2 + 200, 2] = rng.uniform(0, 5, 200)
    print(f"[case] trace_rays semantics={semantics} n={n}", flush=True)
    g = ctx.trace_rays(org, dirs, sc.basecube, sc.depth, semantics)
    o = osc.trace_many(org, dirs, semantics)
    for key in ("status", "model", "leaf"):
        bad = np.count_nonzero(g[key] != o[key])
        assert bad == 0, f"{key}: {bad} mismatches"
    assert np.array_equal(g["isp"], o["isp"])
    assert np.array_equal(g["tlf"], o["tlf"])
    assert np.count_nonzero(o["status"] == 1) > 1000


def test_range_updates(ctx):
    """In-place zeroing and appended octets through partial uploads
    (modelutil.c:425-500): the device tree equals the edited host tree, and
    the frame rendered after the edits equals the oracle's. The compact layout
    takes them as an in-place patch (cleared live bits, moved child blocks,
    appended subtrees), not a relayout."""
    sc, pnt, nrm, col, _ = H.scene(1, npoints=300000)
    tree = q.octree_from_points(pnt, sc.basecube, sc.depth)
    nodes = tree.nodes()
    cap = tree.len + 50000
    big = np.zeros((cap, 12), np.int32)
    big[:tree.len] = nodes
    # first upload with a generous size, then range updates as the reference does
    ctx.upload(big, q._ffi.TYPE_INT, tree.len * 48, 16, 0, tree.len * 48, q._ffi.BUFFER_STATIC_OCTREE)
    ctx.upload_model(col, nrm)
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    osc0 = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc0, (900.0, 900.0, 900.0), (np.pi, 0.0), 160, 120, sc.lights, sc.basecube,
                        sc.depth)
    _assert_same(g, o, "before edits")
    leaf_before = o["leaf"]
    rel0, pat0, _ = ctx.layout_stats()
    rng = np.random.default_rng(3)
    victims = pnt[rng.choice(pnt.shape[0], 3000, replace=False)]
    for p in victims:  # punch: zero slots, upload the one modified node
        m, o = tree.remove_point(p)
        if o >= 0:
            nv = tree.nodes_view()
            ctx.upload(nv, q._ffi.TYPE_INT, tree.len * 48, 16, o * 48, (o + 1) * 48,
                           q._ffi.BUFFER_STATIC_OCTREE)
    newp = (victims[:2000] + rng.normal(0, 3, (2000, 3))).astype(np.float32).clip(1, 1799)
    for i, p in enumerate(newp):  # re-insert moved points: upload every touched node
        oa = np.zeros(sc.depth + 1, np.int32)
        tree.insert_point(i, p, octindarr=oa)
        nv = tree.nodes_view()
        for o in oa[oa > 0]:
            ctx.upload(nv, q._ffi.TYPE_INT, tree.len * 48, 16, int(o) * 48, (int(o) + 1) * 48,
                           q._ffi.BUFFER_STATIC_OCTREE)
    final = tree.nodes()
    back = ctx.download_octree(tree.len)
    assert np.array_equal(back, final)
    osc = po.OracleScene(final, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 900.0), (np.pi, 0.0), 160, 120, sc.lights, sc.basecube,
                        sc.depth)
    _assert_same(g, o, "after edits")
    assert np.count_nonzero(o["leaf"] != leaf_before) > 0  # the edits are in view
    rel1, pat1, _ = ctx.layout_stats()
    if ctx.tuning()["layout"] == 0 and ctx.tuning()["path"] == 1:
        assert (rel1, pat1) == (rel0, pat0 + 1), "the edits were not patched in place"
    # a second batch on the patched layout: removals of re-inserted points
    # (single-node uploads) and a range of mostly unchanged nodes
    for p in newp[:50]:
        m, o_ = tree.remove_point(p)
        if o_ >= 0:
            ctx.upload(tree.nodes_view(), q._ffi.TYPE_INT, tree.len * 48, 16, o_ * 48, (o_ + 1) * 48,
                       q._ffi.BUFFER_STATIC_OCTREE)
    nv = tree.nodes()
    ctx.upload(nv, q._ffi.TYPE_INT, tree.len * 48, 16, 48, min(tree.len, 4000) * 48, q._ffi.BUFFER_STATIC_OCTREE)
    osc2 = po.OracleScene(nv, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc2, (900.0, 900.0, 900.0), (np.pi, 0.0), 160, 120, sc.lights, sc.basecube,
                        sc.depth)
    _assert_same(g, o, "second edit batch")
    if ctx.tuning()["layout"] == 0 and ctx.tuning()["path"] == 1:
        assert ctx.layout_stats()[:2] == (rel0, pat0 + 2)


def test_patch_fallbacks(ctx):
    """Range uploads the in-place patch does not take go to a relayout: a
    root slot re-pointed at an appended copy of its child (the copy's
    children are still laid out under the detached child), and the slot
    restored (the old child's children are now laid out under the copy).
    Frames equal the oracle in every case."""
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=50000)
    _upload_scene(ctx, nodes, col, nrm)
    osc = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)
    cam, ang = (900.0, 900.0, 900.0), (np.pi, 0.0)
    g, o = _render_both(ctx, osc, cam, ang, 96, 64, sc.lights, sc.basecube, sc.depth)
    _assert_same(g, o, "start")
    n = nodes.copy()
    root = n[0]
    kids = [k for k in range(8) if root[k] > 0]
    # (a) replace root child kids[0] by an appended copy of that subtree's root node
    copy_ = n[root[kids[0]]].copy()
    n2 = np.vstack([n, copy_[None]])
    n2[0, kids[0]] = n.shape[0]
    ctx.upload(n2, q._ffi.TYPE_INT, n2.shape[0] * 48, 16, 0, 48, q._ffi.BUFFER_STATIC_OCTREE)
    ctx.upload(n2, q._ffi.TYPE_INT, n2.shape[0] * 48, 16, n.shape[0] * 48, n2.shape[0] * 48,
               q._ffi.BUFFER_STATIC_OCTREE)
    osc2 = po.OracleScene(n2, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc2, cam, ang, 96, 64, sc.lights, sc.basecube, sc.depth)
    _assert_same(g, o, "root child re-pointed at a copy")
    rel1 = ctx.layout_stats()[0]
    # (b) restore the original slot: a tree again
    ctx.upload(n, q._ffi.TYPE_INT, n2.shape[0] * 48, 16, 0, 48, q._ffi.BUFFER_STATIC_OCTREE)
    n3 = n2.copy()
    n3[0] = n[0]
    osc3 = po.OracleScene(n3, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc3, cam, ang, 96, 64, sc.lights, sc.basecube, sc.depth)
    _assert_same(g, o, "restored")
    if ctx.tuning()["layout"] == 0 and ctx.tuning()["path"] == 1:
        assert ctx.layout_stats()[0] == rel1 + 1


def test_built_tree_needs_no_relayout(ctx):
    """qbt_build_octree writes the compact layout with the node array: renders
    use it with no relayout, equal to the oracle on the downloaded array."""
    sc = q.Scene(1, seed=4, npoints=200000)
    pnt, nrm, col = sc.points()
    ctx.upload_points(pnt)
    ctx.upload_model(col, nrm)
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    rel0, pat0, _ = ctx.layout_stats()
    nn = ctx.build_octree(sc.basecube, sc.depth, pnt.shape[0])
    nodes = ctx.download_octree(nn)
    osc = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth)
    g, o = _render_both(ctx, osc, (900.0, 900.0, 900.0), (2.2, -0.1), 128, 96, sc.lights, sc.basecube, sc.depth)
    _assert_same(g, o, "built tree")
    rel1, pat1, rec = ctx.layout_stats()
    assert (rel1, pat1) == (rel0, pat0)
    if ctx.tuning()["layout"] == 0 and ctx.tuning()["path"] == 1:
        assert ctx.stats()["compact"] == 1 and rec == nn


def test_update_entry_point(ctx):
    """qbt_update derives light/basecube/dimensions like octree_glc_update."""
    sc, pnt, nrm, col, nodes = H.scene(1)
    _upload_scene(ctx, nodes, col, nrm)
    lighta = 0.7
    print("[case] qbt_update", flush=True)
    ctx.update(640.0, 480.0, (900.0, 900.0, 900.0), (np.pi, 0.1, 0.0), lighta, 6, sc.depth, 1800.0, 1)
    g = ctx.read_frame(0)
    libm = ctypes.CDLL("libm.so.6")
    libm.sinf.argtypes = [ctypes.c_float]
    libm.sinf.restype = ctypes.c_float
    s = float(libm.sinf(lighta))  # the host sinf the library uses
    light = np.array([420.0, np.float32(200.0 - s * 20.0), np.float32(680.0 - s * 200.0)], np.float32)
    osc = H.oracle_scene(1)
    o = osc.render((900.0, 900.0, 900.0), (np.pi, 0.1), np.float32(640.0 / 3.0), np.float32(480.0 / 3.0), light,
                   shoot=1)
    assert g["hit"].shape == (160, 213)
    _assert_same(g, o, "update")


def test_invalid_tree_rejected(ctx):
    nodes = np.zeros((3, 12), np.int32)
    nodes[0, 0] = 1
    nodes[1, 3] = 7  # out of range
    col = np.zeros((3, 3), np.float32)
    _upload_scene(ctx, nodes, col, col)
    v = q.view((900, 900, 900), (0, 0), 16, 16, (0, 1800, 1800, 1800), 4)
    with pytest.raises(q.QbtError) as e:
        ctx.render(v, [[900, 1700, 900]])
    assert e.value.code == q._ffi.QBT_ETREE


def test_scattered_range_uploads_checked(ctx):
    """Single-node uploads scattered over the array are checked node range by
    node range: a bad child index in one of them is QBT_ETREE; re-uploading
    that node restores the frame exactly."""
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=20000)
    _upload_scene(ctx, nodes, col, nrm)
    v = q.view((900, 900, 900), (0.3, 0.1), 48, 32, sc.basecube, sc.depth)
    ctx.render(v, sc.lights, q.OUT_ALL)
    ref = ctx.read_frame(0)
    n = nodes.shape[0]
    T = q._ffi
    bad = nodes.copy()
    k = n // 2
    bad[k, 5] = n + 5
    print(f"[case] scattered uploads: n={n}, bad node {k}", flush=True)
    for o in (3, k, n - 2):
        ctx.upload(bad, T.TYPE_INT, n * 48, 16, o * 48, (o + 1) * 48, T.BUFFER_STATIC_OCTREE)
    with pytest.raises(q.QbtError) as e:
        ctx.render(v, sc.lights, q.OUT_ALL)
    assert e.value.code == q._ffi.QBT_ETREE
    ctx.upload(nodes, T.TYPE_INT, n * 48, 16, k * 48, (k + 1) * 48, T.BUFFER_STATIC_OCTREE)
    ctx.render(v, sc.lights, q.OUT_ALL)
    got = ctx.read_frame(0)
    mm = H.compare_frames(got, ref)
    assert all(x == 0 for x in mm.values()), mm


def test_upload_argument_checks(ctx):
    nodes = np.zeros((2, 12), np.int32)
    T = q._ffi
    for args in [(T.TYPE_FLOAT, 96, 16, 0, 96, T.BUFFER_STATIC_OCTREE),   # wrong type
                 (T.TYPE_INT, 96, 12, 0, 96, T.BUFFER_STATIC_OCTREE),     # wrong itemsize
                 (T.TYPE_INT, 96, 16, 32, 16, T.BUFFER_STATIC_OCTREE),    # start > end
                 (T.TYPE_INT, 96, 16, 0, 112, T.BUFFER_STATIC_OCTREE),    # end > size
                 (T.TYPE_INT, 80, 16, 0, 80, T.BUFFER_STATIC_OCTREE),     # partial node
                 (T.TYPE_INT, 96, 16, 8, 96, T.BUFFER_STATIC_OCTREE),     # unaligned start
                 (T.TYPE_INT, 96, 16, 0, 96, 9)]:                         # bad buffer id
        with pytest.raises(q.QbtError) as e:
            ctx.upload(nodes, *args)
        assert e.value.code == T.QBT_EINVAL


def test_render_and_trace_limits(ctx):
    """Argument limits: at most QBT_MAX_LIGHTS_RENDER lights and QBT_MAX_VIEWS
    views per call, views of at least 1x1, maxlevel at most 24 (the LDS
    stack depth), known trace semantics; an empty particle batch is a no-op."""
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=20000)
    _upload_scene(ctx, nodes, col, nrm)
    T = q._ffi
    v = q.view((900, 900, 900), (0.3, 0.1), 16, 16, sc.basecube, sc.depth)
    cases = [
        lambda: ctx.render(v, np.full((9, 3), 900.0, np.float32)),                       # 9 lights
        lambda: ctx.render([v] * 257, sc.lights),                                         # 257 views
        lambda: ctx.render(q.view((900, 900, 900), (0, 0), 0, 16, sc.basecube, sc.depth), sc.lights),
        lambda: ctx.render(q.view((900, 900, 900), (0, 0), 16, 16, sc.basecube, 25), sc.lights),  # deeper than 24
        lambda: ctx.trace_rays(np.zeros((1, 3)), np.ones((1, 3)), sc.basecube, sc.depth, semantics=3),
    ]
    for k, f in enumerate(cases):
        with pytest.raises(q.QbtError) as e:
            f()
        assert e.value.code == T.QBT_EINVAL, k
    ctx.render(v, np.full((8, 3), 900.0, np.float32))  # 8 lights is allowed
    p, s = ctx.particle_update(np.zeros((0, 3)), np.zeros((0, 3)), sc.depth, float(sc.basecube[3]))
    assert p.shape == (0, 3) and s.shape == (0, 3)


def test_merged_launch_matches_two_launches(gpu_ctx):
    """The merged camera-and-shadow launch (k_frame) against the separate
    camera and shadow launches on the same frames: every output, the
    expansion counts and the ray statistics identical, and the oracle on the
    largest frame. Small frames put every camera ray in flight at once, so
    shadow lanes reach most slots before their camera ray ends and hand the
    ray back to the camera lane (stats shadow_deferred): both hand-over
    directions are exercised."""
    ctx = gpu_ctx
    sc, pnt, nrm, col, nodes = H.scene(1, npoints=200000)
    _upload_scene(ctx, nodes, col, nrm)
    lights = np.array([[900.0, 1700.0, 900.0], [300.0, 1500.0, 1200.0]], np.float32)
    pos, ang = sc.poses(3, seed=17)
    deferred = 0
    try:
        for i, (W, H_) in enumerate([(8, 8), (64, 48), (320, 240)]):
            v = q.view(pos[i], ang[i], W, H_, sc.basecube, sc.depth, i % 2)
            out = {}
            for merge in (0, 2):
                ctx.set_tuning(path=1, layout=0, merge=merge)
                print(f"[case] merge={merge} {W}x{H_}", flush=True)
                ctx.render(v, lights, ALL)
                out[merge] = (ctx.read_frame(0), ctx.stats())
            (f0, s0), (f1, s1) = out[0], out[2]
            for k in f0:
                assert np.array_equal(f0[k], f1[k]), f"{W}x{H_}: {k} differs"
            for k in ("primary_rays", "shadow_rays", "hits", "expansions", "expansions_primary"):
                assert s0[k] == s1[k], (k, s0[k], s1[k])
            assert s0["shadow_deferred"] == 0 and s1["shadow_deferred"] <= s1["shadow_rays"]
            deferred += s1["shadow_deferred"]
        o = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth).render(pos[2], ang[2], 320, 240, lights, shoot=0)
        _assert_same(f1, o, "merged 320x240")
        assert deferred > 0
    finally:
        ctx.set_tuning(path=1, layout=0, merge=1)
