"""Closed-form pins of the oracle's pixel-program sub-rules
(src/qubatron/shaders/octree_fsh.c), computed here in float64 from the scene
geometry, not from the oracle's formulas:

* the light sprite (octree_fsh.c:415-418, 452-459): cone test and the
  `resv.x / lghtv.x > 1` rule;
* the dynamic attribute override (octree_fsh.c:233-241): colour and normal
  come from the dynamic model only where the dynamic leaf's model index is
  above 0;
* the multi-light Lambert sum: 0.1 + 0.2 camv + 0.7 * sum(lam_l * bit_l).

Each test was run once against a planted mistake in oracle/qbt_oracle.c and
failed: `>=` for `>` in the sprite ratio test (the ratio-exactly-1 case),
`md >= 0` for `md > 0` (the md == 0 leaves), and the light term assigned
per light instead of summed (the cases with a lit light before the last).
"""
import numpy as np
import pytest

import qubatron_amd as q
from oracle import pyoracle as po

BC = np.array([0.0, 1800.0, 1800.0, 1800.0], np.float32)
LEVELS = 9
U = 1800.0 / 2 ** LEVELS  # 3.515625, exact in fp32


def _tree(pts, levels=LEVELS):
    return q.octree_from_points(np.asarray(pts, np.float32), BC, levels).nodes()


def _sheet_z(zc, cx, cy, half=40.0):
    """points at the centres of the leaf cells of the plane z = zc around (cx, cy)"""
    i0, i1 = int((cx - half) // U), int((cx + half) // U)
    j0, j1 = int((cy - half) // U), int((cy + half) // U)
    g = np.stack(np.meshgrid((np.arange(i0, i1) + 0.5) * U, (np.arange(j0, j1) + 0.5) * U), -1).reshape(-1, 2)
    return np.column_stack([g, np.full(len(g), zc)]).astype(np.float32)


# ---------------------------------------------------------------- sprite ----

W = H = 256
CAM = np.array([880.0, 905.0, 600.0], np.float32)
LZ = 400 * U  # 1406.25: a leaf boundary in z


def _cone_pixels(cam, light, W_, H_):
    """(in-cone mask, ambiguous mask) of a yaw pi / pitch 0 view: pixel
    (px, py) looks along (W/2 - px - 0.5, py + 0.5 - H/2, W/2); the cone
    holds the pixels within 0.02 rad of the camera-to-light direction."""
    px, py = np.meshgrid(np.arange(W_), np.arange(H_))
    d = np.stack([W_ / 2 - px - 0.5, py + 0.5 - H_ / 2, np.full(px.shape, W_ / 2)], -1).astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    lv = (np.asarray(light, np.float64) - np.asarray(cam, np.float64))
    lv /= np.linalg.norm(lv)
    ang = np.arccos(np.clip(d @ lv, -1.0, 1.0))
    return ang < 0.02, np.abs(ang - 0.02) < 1e-4


def _sprite_frame(nodes, light):
    osc = po.OracleScene(nodes, np.full((max(1, nodes.shape[0]), 3), 0.25, np.float32),
                         np.tile([[0.0, 0.0, -1.0]], (max(1, nodes.shape[0]), 1)).astype(np.float32), BC, LEVELS)
    return osc.render(CAM, (np.pi, 0.0), W, H, np.asarray([light], np.float32))


def _check_sprite(fr, light, white):
    cone, amb = _cone_pixels(CAM, light, W, H)
    assert cone.sum() >= 12  # the cone covers a few pixels around the light
    spr = (fr["flags"] & po.PX_SPRITE) != 0
    ok = ~amb
    if white:
        assert np.array_equal(spr[ok], cone[ok])
        assert np.all(fr["rgbaf"][cone & ok] == 1.0)
    else:
        assert not spr.any()


@pytest.mark.parametrize("case,zsheet,white", [
    ("sheet at the light: ratio exactly 1", LZ + 0.5 * U, False),
    ("sheet in front of the light", LZ - 19.5 * U, False),
    ("sheet behind the light", LZ + 20.5 * U, True),
])
def test_sprite_ratio_rule(case, zsheet, white):
    """The sprite trace runs from the camera along light - camfp, so resv.x /
    lghtv.x is its hit parameter: above 1 only when the first leaf lies
    beyond the light. With the sheet's leaves starting exactly at z = L.z
    the entry is t = 806.25 / 806.25 = 1 and resv.x = lghtv.x = 10 exactly:
    1 > 1 is false, so no sprite (a `>=` would paint it)."""
    light = np.array([890.0, 910.0, LZ], np.float32)
    nodes = _tree(_sheet_z(zsheet, 890.0, 910.0))
    if case.startswith("sheet at"):  # the entry point the rule compares
        osc = po.OracleScene(nodes, None, None, BC, LEVELS)
        r = osc.trace(CAM, light - CAM)
        assert r.status == po.HIT and list(r.isp) == [890.0, 910.0, LZ, 1.0]
    _check_sprite(_sprite_frame(nodes, light), light, white)


@pytest.mark.parametrize("lx,white", [(890.0, False), (870.0, True)])
def test_sprite_empty_tree(lx, white):
    """Through an empty tree the sprite trace misses and its entry point is
    0, so resv.x / lghtv.x = -camfp.x / (L.x - camfp.x): -88 for a light at
    larger x (no sprite), +88 for one at smaller x (sprite) - a quirk of the
    rule the closed form states."""
    light = np.array([lx, 910.0, LZ], np.float32)
    fr = _sprite_frame(np.zeros((1, 12), np.int32), light)
    assert not (fr["flags"] & po.PX_LEAF).any()
    _check_sprite(fr, light, white)


# ------------------------------------------------- dynamic attribute override ----

def test_dynamic_override_only_above_zero():
    """The leaf's colour and normal come from the static model of slot 8 and
    are replaced by the dynamic model's where the dynamic leaf's slot 8 is
    above 0 (octree_fsh.c:233-241). Dynamic point 0 creates its leaf with
    model index 0, so that leaf keeps the static attributes; dynamic point 1
    overrides. A dynamic-only leaf of dynamic point 0 shows static model 0."""
    zc = 800.0 + 0.5 * U
    stat_pts = _sheet_z(zc, 900.0, 900.0, half=30.0)
    n_s = stat_pts.shape[0]
    col_s = np.tile([[0.2, 0.4, 0.6]], (n_s, 1)).astype(np.float32)
    col_s[0] = (0.9, 0.1, 0.1)  # static model 0
    nrm_s = np.tile([[0.0, 0.0, -1.0]], (n_s, 1)).astype(np.float32)
    a = stat_pts[n_s // 3]
    b = stat_pts[2 * n_s // 3]
    c = np.array([300.0 + 0.5 * U, 300.0 + 0.5 * U, zc], np.float32)  # no static leaf here
    for case, dyn_pts, probe, exp_col, exp_md in [
        ("md == 0 on a static leaf", [a, b], a, (0.2, 0.4, 0.6), 0),
        ("md == 1 on a static leaf", [a, b], b, (0.7, 0.7, 0.0), 1),
        ("md == 0, dynamic leaf only", [c, b], c, (0.9, 0.1, 0.1), 0),
    ]:
        dyn = np.asarray(dyn_pts, np.float32)
        col_d = np.array([[0.0, 0.8, 0.8], [0.7, 0.7, 0.0]], np.float32)
        nrm_d = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
        osc = po.OracleScene(_tree(stat_pts), col_s, nrm_s, BC, LEVELS, dyn_nodes=_tree(dyn), dyn_col=col_d,
                             dyn_nrm=nrm_d)
        org = np.array([probe[0] + 0.01, probe[1] - 0.02, 100.0], np.float32)
        r = osc.trace(org, np.asarray(probe, np.float32) - org)
        assert r.status == po.HIT, case
        assert r.md == exp_md, case
        np.testing.assert_array_equal(np.array(r.col[:3], np.float32), np.array(exp_col, np.float32), case)
        exp_n = nrm_d[1] if exp_md > 0 else (0.0, 0.0, -1.0)
        np.testing.assert_array_equal(np.array(r.nrm[:3], np.float32), np.array(exp_n, np.float32), case)


# ------------------------------------------------- multi-light Lambert sum ----

WALL_Z = 228 * U  # 801.5625, a leaf boundary


def _wall_scene():
    """A wall of leaves at z in [WALL_Z, WALL_Z + u] facing -z (towards the
    camera at z = 300 looking +z and the lights at z = 50.75), colour
    (0.5, 0.6, 0.7); a blocker sheet at z ~ 501 across the path of L_BLOCK
    only (its path crosses that plane near x = 1062, L_OPEN's near x = 822)."""
    wall = _sheet_z(WALL_Z + 0.5 * U, 900.0, 900.0, half=120.0)
    blocker = _sheet_z(142.5 * U, 1062.0, 900.0, half=60.0)
    pts = np.concatenate([wall, blocker])
    col = np.tile([[0.5, 0.6, 0.7]], (pts.shape[0], 1)).astype(np.float32)
    nrm = np.tile([[0.0, 0.0, -1.0]], (pts.shape[0], 1)).astype(np.float32)
    return po.OracleScene(_tree(pts), col, nrm, BC, LEVELS)


CAM2 = np.array([903.25, 897.5, 300.0], np.float32)
L_OPEN = np.array([701.5, 1003.25, 50.75], np.float32)   # straight path to the wall centre
L_BLOCK = np.array([1301.5, 905.25, 50.75], np.float32)  # the blocker sits on its path


def _expected(fr, px, py, lights, bits):
    """float64 closed form of one shaded pixel: rgb = col * (0.1 + 0.2 camv +
    0.7 sum_l lam_l bit_l), blue * 0.7; camv from the pixel direction of a
    yaw pi view (W/2 - px - 0.5, py + 0.5 - H/2, W/2), lam_l = cos of the
    angle between the wall normal (0, 0, -1) and isp -> light."""
    W_, H_ = fr["hit"].shape[1], fr["hit"].shape[0]
    isp = fr["isp"][py, px, :3].astype(np.float64)
    d = np.array([W_ / 2 - px - 0.5, py + 0.5 - H_ / 2, W_ / 2])
    n = np.array([0.0, 0.0, -1.0])
    camv = max(float(np.dot(-d / np.linalg.norm(d), n)), 0.0)
    f = 0.1 + 0.2 * camv
    for L, b in zip(lights, bits):
        lv = np.asarray(L, np.float64) - isp
        f += 0.7 * b * max(float(np.dot(lv / np.linalg.norm(lv), n)), 0.0)
    return np.array([0.5 * f, 0.6 * f, 0.7 * f * 0.7])


@pytest.mark.parametrize("lights,bits", [
    ((L_OPEN, L_BLOCK), (1, 0)),
    ((L_BLOCK, L_OPEN), (0, 1)),
    ((L_OPEN, L_OPEN + np.float32([40.0, -30.0, 0.0])), (1, 1)),
])
def test_two_light_lambert_sum(lights, bits):
    """Two lights, pixels at the wall centre: the shadow bits follow the
    blocker and the colour is the closed-form sum (rgbaf is before clamping,
    so a sum above 1 stays visible: the two-lit case reaches about 1.6)."""
    osc = _wall_scene()
    Wd = Hd = 64
    lights = np.asarray(lights, np.float32)
    fr = osc.render(CAM2, (np.pi, 0.0), Wd, Hd, lights)
    for px, py in [(32, 32), (30, 35), (36, 29)]:
        assert fr["flags"][py, px] & po.PX_SHADED
        assert fr["isp"][py, px, 2] == WALL_Z  # entered through the wall's face
        assert fr["shadow"][py, px] == bits[0] | (bits[1] << 1)
        exp = _expected(fr, px, py, lights, bits)
        assert exp[0] / 0.5 > 0.3 + 0.65 * sum(bits)  # the light terms are large: lam is about 0.95
        np.testing.assert_allclose(fr["rgbaf"][py, px, :3], exp, rtol=2e-6)
