"""Pins of the CPU oracle (the restatement of octree_fsh.c / octree.c) that do
not use the oracle's own formulas: brute-force nearest-leaf search, axis
aligned closed forms, containment invariants, the 5-point OCTTEST scene, the
dual-tree reduction and the camera closed form."""
import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po
from tests import brute

BC = np.array([0.0, 1800.0, 1800.0, 1800.0], np.float32)


def _tree(pts, levels, bc=BC):
    return q.octree_from_points(pts, bc, levels).nodes()


def _random_rays(rng, n, inside_frac=0.7):
    org = rng.uniform(1, 1799, (n, 3)).astype(np.float32)
    k = int(n * (1 - inside_frac))
    org[:k] = rng.uniform(-900, 2700, (k, 3)).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32) * rng.uniform(0.5, 50, (n, 1)).astype(np.float32)
    return org, dirs


@pytest.mark.parametrize("levels,npts,seed", [(5, 40, 0), (7, 400, 1), (9, 3000, 2)])
def test_brute_force_first_leaf(levels, npts, seed):
    """Front-to-back traversal returns the occupied leaf with the smallest
    entry parameter along the ray (float64 slab test over all leaves)."""
    rng = np.random.default_rng(seed)
    # clustered points: planes and blobs so that rays cross occupied regions
    u = 1800.0 / 2 ** levels  # a dense patch of a plane (70% of its cells) plus a blob
    cells = np.argwhere(rng.uniform(size=(2 ** levels // 2, 2 ** levels // 2)) < 0.7) + 2 ** levels // 4
    patch = np.stack([(cells[:, 0] + 0.5) * u, np.full(len(cells), 900.3), (cells[:, 1] + 0.5) * u], 1)
    pts = np.concatenate([patch, rng.normal(900, 300, (npts, 3)).clip(1, 1799)]).astype(np.float32)
    nodes = _tree(pts, levels)
    ln, lo, hi = brute.leaves(nodes, BC, levels)
    osc = po.OracleScene(nodes, None, None, BC, levels)
    org, dirs = _random_rays(rng, 1500)
    best, tent, degen = brute.first_hits(org, dirs, lo, hi)
    res = osc.trace_many(org, dirs)
    checked = hits = 0
    for r in range(org.shape[0]):
        if degen[r]:
            continue
        checked += 1
        if best[r] < 0:
            assert res["status"][r] in (po.MISS, po.DISCARD), r
            continue
        hits += 1
        assert res["status"][r] == po.HIT, r
        assert res["leaf"][r] == ln[best[r]], r
        assert res["tlf"][r][0] == lo[best[r]][0] and res["tlf"][r][1] == hi[best[r]][1]
        assert abs(res["isp"][r][3] - tent[r]) <= 1e-4 * max(1.0, tent[r]), r
    assert checked > 1300 and hits > 40


def test_hit_leaf_contains_isp():
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 1800, (20000, 3)).astype(np.float32)
    nodes = _tree(pts, 8)
    osc = po.OracleScene(nodes, None, None, BC, 8)
    org, dirs = _random_rays(rng, 3000)
    res = osc.trace_many(org, dirs)
    h = res["status"] == po.HIT
    assert h.sum() > 300
    isp, tlf = res["isp"][h], res["tlf"][h]
    tol = 1e-3
    assert np.all(isp[:, 0] >= tlf[:, 0] - tol) and np.all(isp[:, 0] <= tlf[:, 0] + tlf[:, 3] + tol)
    assert np.all(isp[:, 1] <= tlf[:, 1] + tol) and np.all(isp[:, 1] >= tlf[:, 1] - tlf[:, 3] - tol)
    assert np.all(isp[:, 2] <= tlf[:, 2] + tol) and np.all(isp[:, 2] >= tlf[:, 2] - tlf[:, 3] - tol)
    assert np.all(isp[:, 3] >= 0)
    # a hit with t > 0 enters through a face: one coordinate equals a face plane exactly
    pos = isp[:, 3] > 0
    onface = ((isp[:, 0] == tlf[:, 0]) | (isp[:, 0] == tlf[:, 0] + tlf[:, 3]) |
              (isp[:, 1] == tlf[:, 1]) | (isp[:, 1] == tlf[:, 1] - tlf[:, 3]) |
              (isp[:, 2] == tlf[:, 2]) | (isp[:, 2] == tlf[:, 2] - tlf[:, 3]))
    assert np.all(onface[pos])


def test_axis_aligned_closed_form():
    """A ray along -z onto a horizontal sheet: the hit leaf is the sheet cell
    under the ray and the entry point is (x0, y0, z_top, (z_top - z0)/dz)."""
    levels = 8
    u = 1800.0 / 2 ** levels
    zc = 700.3
    g = (np.arange(0, 2 ** levels) + 0.5) * u
    xx, yy = np.meshgrid(g[100:140], g[100:140])
    pts = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, zc)], 1).astype(np.float32)
    nodes = _tree(pts, levels)
    osc = po.OracleScene(nodes, None, None, BC, levels)
    rng = np.random.default_rng(0)
    for _ in range(200):
        x0, y0 = (rng.uniform(100.2, 139.8, 2) * u).astype(np.float32)
        z0 = np.float32(rng.uniform(800, 1700))
        dz = np.float32(-rng.uniform(0.1, 20))
        r = osc.trace((x0, y0, z0), (0.0, 0.0, dz))
        assert r.status == po.HIT
        iz = np.floor(zc / u)
        ztop = np.float32((iz + 1) * u)
        assert r.isp[0] == x0 and r.isp[1] == y0 and r.isp[2] == ztop
        assert r.isp[3] == np.float32((ztop - z0) / dz)
        assert r.tlf[0] == np.float32(np.floor(x0 / u) * u)
        # CPU-twin semantics agree here (no d == 0 x-plane case inside: x-plane tests return FLT_MAX)
        rc = osc.trace((x0, y0, z0), (0.0, 0.0, dz), po.SEM_CPU)
        assert rc.status == po.HIT and rc.socti == r.socti


def test_octtest_scene():
    """The 5-point OCTTEST scene (modelutil.c:91-110) at depth 12: a ray from
    the OCTTEST camera (900, 900, 3000) to each point's leaf centre hits that
    point's leaf and reports its model index."""
    pts = np.array([[10, 690, 10], [10, 340, 10], [10, 340, 690], [10, 10, 10], [690, 10, 690]], np.float32)
    nodes = _tree(pts, 12)
    osc = po.OracleScene(nodes, np.ones((5, 3), np.float32), np.tile([[0, 0, -1.0]], (5, 1)), BC, 12)
    cam = np.array([900.0, 900.0, 3000.0], np.float32)
    u = 1800.0 / 4096
    for i, p in enumerate(pts):
        centre = (np.floor(p / u) + 0.5) * u
        d = (centre - cam).astype(np.float32)
        r = osc.trace(cam, d)
        rc = osc.trace(cam, d, po.SEM_CPU)
        if i == 3:
            # (10,10,10) from (900,900,3000): dx == dy and px == py, so at depth 6 the
            # x and y mid-plane hits tie exactly; the flip rule (octree_fsh.c:300-311)
            # then assigns octets 0 and 1 instead of the diagonal 2 -> 6 and the
            # reference misses the leaf. Breaking the symmetry restores the hit.
            assert r.status == po.MISS and rc.status == po.MISS
            d2 = d.copy()
            d2[0] += np.float32(1e-3)
            r2 = osc.trace(cam, d2)
            assert r2.status == po.HIT and r2.ms == i
            continue
        assert r.status == po.HIT and r.ms == i
        assert rc.status == po.HIT and rc.ms == i


def test_start_on_midplane_quirk():
    """A ray starting exactly on a root mid-plane is classified into the
    low-x / high-y / high-z side (strict comparisons, octree_fsh.c:296-298)
    and plane hits at t = 0 are rejected, so a ray from the exact cube centre
    heading into another octant misses although it crosses occupied leaves."""
    rng = np.random.default_rng(8)
    pts = rng.uniform(0, 1800, (60000, 3)).astype(np.float32)
    nodes = _tree(pts, 7)
    ln, lo, hi = brute.leaves(nodes, BC, 7)
    osc = po.OracleScene(nodes, None, None, BC, 7)
    c = np.array([900.0, 900.0, 900.0], np.float32)
    toward_0 = np.array([-1.3, 0.7, 0.9], np.float32)   # octet 0 side: hits like the brute force
    away = np.array([1.3, -0.7, -0.9], np.float32)      # octet 7 side: misses
    b0, _, _ = brute.first_hits(c[None], toward_0[None], lo, hi)
    assert osc.trace(c, toward_0).socti == ln[b0[0]]
    b1, _, _ = brute.first_hits(c[None], away[None], lo, hi)
    assert b1[0] >= 0 and osc.trace(c, away).status == po.MISS
    c2 = c + np.float32(0.01)
    b2, _, _ = brute.first_hits(c2[None], away[None], lo, hi)
    assert osc.trace(c2, away).socti == ln[b2[0]]


def test_empty_dynamic_tree_is_single_tree():
    rng = np.random.default_rng(9)
    pts = rng.uniform(0, 1800, (5000, 3)).astype(np.float32)
    nodes = _tree(pts, 8)
    a = po.OracleScene(nodes, None, None, BC, 8)
    b = po.OracleScene(nodes, None, None, BC, 8, dyn_nodes=np.zeros((1, 12), np.int32))
    org, dirs = _random_rays(rng, 1000)
    ra, rb = a.trace_many(org, dirs), b.trace_many(org, dirs)
    for k in ("status", "leaf", "model", "isp", "tlf"):
        assert np.array_equal(ra[k], rb[k])


def test_dual_tree_equals_union():
    """Lockstep traversal of (static A, dynamic B) reaches the same leaf cell
    as the single traversal of the tree of A union B (octree_fsh.c:313-317)."""
    rng = np.random.default_rng(4)
    levels = 8
    A = rng.uniform(0, 1800, (6000, 3)).astype(np.float32)
    B = rng.normal(900, 250, (6000, 3)).clip(1, 1799).astype(np.float32)
    na, nb, nu = _tree(A, levels), _tree(B, levels), _tree(np.concatenate([A, B]), levels)
    dual = po.OracleScene(na, None, None, BC, levels, dyn_nodes=nb)
    union = po.OracleScene(nu, None, None, BC, levels)
    org, dirs = _random_rays(rng, 1500)
    rd, ru = dual.trace_many(org, dirs), union.trace_many(org, dirs)
    assert np.array_equal(rd["status"], ru["status"])
    assert np.array_equal(rd["tlf"], ru["tlf"])
    assert np.array_equal(rd["isp"], ru["isp"])


def test_cpu_and_shader_semantics_agree_without_zero_components():
    rng = np.random.default_rng(6)
    pts = rng.uniform(0, 1800, (20000, 3)).astype(np.float32)
    nodes = _tree(pts, 9)
    osc = po.OracleScene(nodes, None, None, BC, 9)
    org, dirs = _random_rays(rng, 2000)
    assert np.all(dirs != 0)
    a, b = osc.trace_many(org, dirs, po.SEM_SHADER), osc.trace_many(org, dirs, po.SEM_CPU)
    for k in ("status", "leaf", "isp", "tlf"):
        assert np.array_equal(a[k], b[k])


def test_camera_closed_form():
    """Yaw pi, pitch 0 looks along +z (octree_fsh.c:403-413): pixel (px, py)
    maps to (W/2 - px - 0.5, py + 0.5 - H/2, W/2) up to rounding."""
    W, H = 320, 240
    pix = [(0, 0), (160, 120), (319, 239), (17, 200)]
    d = po.pixel_dirs((np.pi, 0.0), W, H, pix)
    for (px, py), v in zip(pix, d):
        exp = np.array([W / 2 - px - 0.5, py + 0.5 - H / 2, W / 2])
        assert np.allclose(v, exp, atol=1e-3)
    # yaw 0, pitch 0 looks along -z
    d0 = po.pixel_dirs((0.0, 0.0), W, H, [(160, 120)])[0]
    assert np.allclose(d0, [0.5, 0.5, -W / 2], atol=1e-4)


def test_shadow_occluder_closed_form():
    """A lit wall pixel: with an occluder between light and wall the shadow
    bit is 0 and the colour is col*(0.1+0.2c) (b *0.7); without it the bit
    is 1 and the light term is added (octree_fsh.c:431-445)."""
    levels = 9
    u = 1800.0 / 2 ** levels
    g = (np.arange(200, 312) + 0.5) * u
    xx, yy = np.meshgrid(g, g)
    # wall at z ~ 300 facing +z; blocker at z ~ 1200 above the wall centre;
    # camera at z = 1100 (below the blocker) looking along -z at the wall
    wall = np.stack([xx.ravel(), yy.ravel(), np.full(xx.size, 300.1)], 1).astype(np.float32)
    gb = (np.arange(230, 280) + 0.5) * u
    bx, by = np.meshgrid(gb, gb)
    blocker = np.stack([bx.ravel(), by.ravel(), np.full(bx.size, 1200.2)], 1).astype(np.float32)
    light = np.array([[901.7, 898.3, 1750.3]], np.float32)  # off the root mid-planes
    nrm_w = np.tile([[0, 0, 1.0]], (wall.shape[0], 1)).astype(np.float32)
    col_w = np.tile([[0.5, 0.6, 0.7]], (wall.shape[0], 1)).astype(np.float32)
    nodes = _tree(np.concatenate([wall, blocker]), levels)
    cols = np.concatenate([col_w, np.ones((blocker.shape[0], 3), np.float32)])
    nrms = np.concatenate([nrm_w, np.tile([[0, 0, 1.0]], (blocker.shape[0], 1)).astype(np.float32)])
    osc = po.OracleScene(nodes, cols, nrms, BC, levels)
    fr = osc.render((903.3, 897.1, 1100.7), (0.0, 0.0), 64, 64, light)
    c = (32, 32)
    assert fr["flags"][c] & po.PX_SHADED
    assert fr["shadow"][c] == 0  # the blocker sits between the light and the wall
    rgb = fr["rgbaf"][c][:3]
    camv = rgb[0] / 0.5 - 0.1
    assert 0.15 < camv / 0.2 <= 1.0 + 1e-6
    np.testing.assert_allclose(rgb, [0.5 * (0.1 + camv), 0.6 * (0.1 + camv), 0.7 * (0.1 + camv) * 0.7], rtol=1e-5)
    # move the light below the blocker: now the wall pixel is lit
    fr2 = osc.render((903.3, 897.1, 1100.7), (0.0, 0.0), 64, 64, np.array([[901.7, 898.3, 1000.3]], np.float32))
    assert fr2["shadow"][c] == 1
    assert fr2["rgbaf"][c][0] > rgb[0] + 0.1


def test_rgba8_packing():
    lib = po.lib()
    import ctypes as C
    for c, exp in [((0.0, 0.5, 1.0, 1.0), (0, 128, 255, 255)), ((-1.0, 2.0, np.nan, 0.2), (0, 255, 0, 51))]:
        a = np.array(c, np.float32)
        v = lib.orc_pack_rgba8(a.ctypes.data_as(C.POINTER(C.c_float)))
        assert [(v >> (8 * k)) & 255 for k in range(4)] == list(exp)
