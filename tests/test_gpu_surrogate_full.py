"""The learned shadow surrogate at the config-5 workload, through the C ABI:
the config-2 scene (10M points, depth 12, one light), trained online on the
GPU by the library's own forward, backward and Adam kernels with 16M shadow
samples collected from 256 exact 1920x1080 frames, then evaluated on held-out
1920x1080 frames with exact-trace fallback for predictions in [0.1, 0.9].

Checked on every evaluated frame: primary outputs (hit, leaf, entry point)
identical to the exact render; every query the surrogate leaves undecided
traced exactly (shadow bit and colour equal to the exact frame); every
decided query following the network's own probability (computed again by
qbt_surrogate_predict from the hit point, the hit leaf's normal and the
light). The exact frame itself equals the oracle on one of them. What the
trained network gains over the exact form: every decided query is one exact
shadow ray fewer (checked per frame, and more than a fifth of them overall);
reported beside it: shadow-bit mismatch, decided fraction, and the per-frame
time of the shadow side (fallback traces + decision pass) against the exact
shadow pass on the same frames.
"""
import os
import time

import numpy as np
import pytest

import qubatron_amd as q
from qubatron_amd import surrogate as S
from oracle import pyoracle as po
from tests import helpers as H

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
W, H_ = 1920, 1080
KIND_PRIMARY, KIND_SHADOW, KIND_SURROGATE = 0, 1, 2


@pytest.fixture(scope="module")
def trained():
    t0 = time.time()
    sc = q.Scene(5)
    pnt, nrm, col = sc.points(nthreads=NT)
    ctx = q.Context(0)
    ctx.upload_points(pnt)
    ctx.upload_model(col, nrm)
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)
    n = ctx.build_octree(sc.basecube, sc.depth, sc.npoints)
    print(f"[scene] config 5: {sc.npoints} points, {n} nodes in {time.time() - t0:.1f}s", flush=True)
    tpos, tang = sc.poses(256, seed=7920)
    tviews = [q.view(tpos[i], tang[i], W, H_, sc.basecube, sc.depth) for i in range(256)]
    ctx.render(tviews[0], sc.lights, q.OUT_FLAGS)
    queries = ctx.stats()["shadow_rays"]
    stride = max(1, int(round(256 * queries / (1 << 24))))
    t0 = time.time()
    info = S.train_online(ctx, tviews, sc.lights, sc.basecube, nsamples=1 << 24, stride=stride,
                          steps_per_frame=32, batch=4096, lr=2e-3, seed=0, final_steps=2000)
    ctx.synchronize()
    print(f"[surrogate] trained in {time.time() - t0:.1f}s: {info}", flush=True)
    yield sc, ctx, pnt, nrm, col, n, info
    ctx.close()


def _shadow_side_ms(ctx, views, lights, mask):
    """per-frame traversal time of the shadow side: shadow kernel + decision pass"""
    for v in views[:1]:
        ctx.render(v, lights, mask)
    ctx.profile_begin()
    for v in views:
        ctx.render(v, lights, mask)
    p = ctx.profile_end()
    k = len(views)
    return (p["kind_ms"][KIND_SHADOW] + p["kind_ms"][KIND_SURROGATE]) / k, p["kind_ms"][KIND_PRIMARY] / k


def test_config5_training_collected_16m_samples(trained):
    sc, ctx, pnt, nrm, col, n, info = trained
    assert info["samples"] == 1 << 24
    assert info["frames"] == 256
    f = info["lit_fraction"]
    entropy = -(f * np.log(f) + (1 - f) * np.log(1 - f))  # loss of the best constant predictor
    assert info["loss_last"] < 0.8 * entropy, info


def test_config5_surrogate_frames(trained):
    sc, ctx, pnt, nrm, col, n, info = trained
    lights = np.asarray(sc.lights, np.float32)
    pos, ang = sc.poses(6, seed=1)
    views = [q.view(pos[i], ang[i], W, H_, sc.basecube, sc.depth) for i in range(6)]
    ctx.surrogate_set_band(0.1, 0.9)
    bad = tot = pred = saved = traced_all = 0
    for i, v in enumerate(views):
        print(f"[case] config 5 surrogate frame {i}", flush=True)
        ctx.render(v, lights, q.OUT_ALL)
        ex = ctx.read_frame(0)
        traced_exact = ctx.stats()["shadow_rays"]
        ctx.render(v, lights, q.OUT_ALL | q.RENDER_SURROGATE)
        su = ctx.read_frame(0)
        st = ctx.stats()
        pred += st["predicted"]
        # the work the network saves: every query it decides is one exact shadow ray fewer
        assert st["shadow_rays"] + st["predicted"] == traced_exact, (st, traced_exact)
        saved += st["predicted"]
        traced_all += traced_exact
        for k in ("hit", "leaf", "isp"):
            np.testing.assert_array_equal(su[k], ex[k])
        keep = (su["flags"] & q.PX_PREDICTED) == 0
        np.testing.assert_array_equal(su["shadow"][keep], ex["shadow"][keep])
        np.testing.assert_array_equal(su["rgba8"][keep], ex["rgba8"][keep])
        sh = ((ex["flags"] & q.PX_SHADED) != 0) & ((ex["flags"] & q.PX_DISCARD) == 0)
        ii = np.nonzero(sh.ravel())[0]
        pr = ctx.surrogate_predict(ex["isp"].reshape(-1, 4)[ii, :3], nrm[ex["hit"].ravel()[ii]],
                                   np.repeat(lights[:1], ii.size, 0), sc.basecube)
        bit = su["shadow"].ravel()[ii] & 1
        assert np.all(bit[pr < 0.099] == 0) and np.all(bit[pr > 0.901] == 1)
        mid = (pr > 0.101) & (pr < 0.899)
        np.testing.assert_array_equal(bit[mid], ex["shadow"].ravel()[ii][mid] & 1)
        b, t = S.shadow_mismatch(ex, su, 1)
        bad, tot = bad + b, tot + t
        if i == 0:  # the exact frame against the oracle
            nodes = ctx.download_octree(n)
            o = po.OracleScene(nodes, col, nrm, sc.basecube, sc.depth).render(pos[i], ang[i], W, H_, lights,
                                                                               nthreads=NT)
            mm = H.compare_frames(ex, o)
            assert all(x == 0 for x in mm.values()), mm
            del nodes, o
    mismatch, frac = bad / tot, pred / tot
    exact_shadow, exact_prim = _shadow_side_ms(ctx, views, lights, q.OUT_DEFAULT)
    sur_shadow, sur_prim = _shadow_side_ms(ctx, views, lights, q.OUT_DEFAULT | q.RENDER_SURROGATE)
    print(f"[surrogate] config 5: mismatch {mismatch:.4f}, decided {frac:.3f}; exact shadow rays traced "
          f"{traced_all} without the network, {traced_all - saved} with it; per frame: primary "
          f"{exact_prim:.3f} / {sur_prim:.3f} ms, shadow side exact {exact_shadow:.3f} ms, with the network "
          f"{sur_shadow:.3f} ms (gain {exact_shadow / sur_shadow:.3f}x)", flush=True)
    assert saved > 0.2 * traced_all
    assert tot > 1_000_000
    assert mismatch < 0.02 and frac > 0.2, (mismatch, frac)


def test_config5_calibrated_band(trained):
    """The band calibrated on 4 held-out exact frames to 1% mismatch
    (qubatron_amd.surrogate.calibrate_band), used on 3 other frames: decisions
    follow the calibrated band, the mismatch stays near the target (the
    frames differ from the calibration frames, so 2% is allowed) and the
    network decides a part of the queries. How large a part depends on the
    view: two thirds of the calibration frames' queries, 15% of these three
    frames' (0.8% mismatch), 39% of the bench's ten."""
    sc, ctx, pnt, nrm, col, n, info = trained
    lights = np.asarray(sc.lights, np.float32)
    cpos, cang = sc.poses(4, seed=3)
    cal = S.calibrate_band(ctx, [q.view(cpos[i], cang[i], W, H_, sc.basecube, sc.depth) for i in range(4)],
                           lights, sc.basecube, target=0.01)
    lo, hi = cal["band"]
    print(f"[surrogate] calibrated band [{lo:.3f}, {hi:.3f}]: {cal}", flush=True)
    assert 0.0 < lo <= 0.5 <= hi < 1.0 and cal["held_out_mismatch"] <= 0.01
    pos, ang = sc.poses(3, seed=11)
    bad = tot = pred = 0
    for i in range(3):
        v = q.view(pos[i], ang[i], W, H_, sc.basecube, sc.depth)
        print(f"[case] calibrated band frame {i}", flush=True)
        ctx.render(v, lights, q.OUT_ALL)
        ex = ctx.read_frame(0)
        ctx.render(v, lights, q.OUT_ALL | q.RENDER_SURROGATE)
        su = ctx.read_frame(0)
        pred += ctx.stats()["predicted"]
        sh = ((ex["flags"] & q.PX_SHADED) != 0) & ((ex["flags"] & q.PX_DISCARD) == 0)
        ii = np.nonzero(sh.ravel())[0]
        pr = ctx.surrogate_predict(ex["isp"].reshape(-1, 4)[ii, :3], nrm[ex["hit"].ravel()[ii]],
                                   np.repeat(lights[:1], ii.size, 0), sc.basecube)
        decided = (su["flags"].ravel()[ii] & q.PX_PREDICTED) != 0
        assert not np.any(decided & (pr > lo + 1e-3) & (pr < hi - 1e-3))  # inside the band: traced
        assert np.all(decided[(pr < lo - 1e-3) | (pr > hi + 1e-3)])       # outside: decided
        b, t = S.shadow_mismatch(ex, su, 1)
        bad, tot = bad + b, tot + t
    print(f"[surrogate] calibrated band on new frames: mismatch {bad / tot:.4f}, decided {pred / tot:.3f}", flush=True)
    assert bad / tot < 0.02 and pred / tot > 0.05
