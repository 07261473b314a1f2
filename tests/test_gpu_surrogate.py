"""The learned shadow surrogate on the GPU (libqbt.so through its C ABI):
MLP forward, fused inference, gradients and Adam against the CPU oracle,
sample collection against the oracle's shadow trace, and the surrogate render
path against the exact one.

Tolerances. The GPU sums MFMA products in its own order, so a pre-activation
can round to the neighbouring bf16 value where the oracle's sequential sum
does not: probabilities agree to 1e-4 in the median and 2e-2 at worst;
gradients to 3e-2 in relative L2 norm per parameter block (cosine > 0.999);
the first Adam step from the GPU's own gradient to 1e-6.
"""
import functools

import numpy as np
import pytest

import qubatron_amd as q
from qubatron_amd import surrogate as S
from oracle import pyoracle as po
from tests.test_mlp_oracle import OFF, rand_params

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"), ("b3", "W4"), ("W4", "b4")]


def _rand_features(seed, n, bc=(0.0, 1800.0, 1800.0, 1800.0)):
    rng = np.random.default_rng(seed)
    isp = rng.uniform(0, 1800, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    light = rng.uniform(0, 1800, (n, 3)).astype(np.float32)
    return isp, nrm, light, po.mlp_features(isp, nrm, light, bc)


def test_params_roundtrip_and_forward(gpu_ctx):
    p = rand_params(21)
    gpu_ctx.surrogate_set_params(p)
    np.testing.assert_array_equal(gpu_ctx.surrogate_params(), p)
    _, _, _, x = _rand_features(22, 1000)  # 1000: not a multiple of 16
    print("[case] forward n=1000", flush=True)
    pg = gpu_ctx.surrogate_forward(x)
    pr, _ = po.mlp_forward(p, x, bf16=True)
    err = np.abs(pg - pr)
    assert np.median(err) < 1e-4 and err.max() < 2e-2, (np.median(err), err.max())


def test_fused_predict_matches_oracle(gpu_ctx):
    """The render path's inference (features computed in the kernel)."""
    p = rand_params(23)
    gpu_ctx.surrogate_set_params(p)
    bc = (100.0, 1900.0, 1850.0, 1800.0)
    rng = np.random.default_rng(24)
    isp = (rng.uniform(0, 1800, (777, 3)) + [100, 100, 50]).astype(np.float32)
    nrm = np.eye(3, dtype=np.float32)[rng.integers(0, 3, 777)]
    light = np.repeat(np.array([[880.25, 1700.0, 915.75]], np.float32), 777, 0)
    print("[case] predict n=777", flush=True)
    pg = gpu_ctx.surrogate_predict(isp, nrm, light, bc)
    pr, _ = po.mlp_forward(p, po.mlp_features(isp, nrm, light, bc), bf16=True)
    err = np.abs(pg - pr)
    assert np.median(err) < 1e-4 and err.max() < 2e-2, (np.median(err), err.max())


@pytest.mark.parametrize("n", [40, 1000, 3000])
def test_gradient_matches_oracle(gpu_ctx, n):
    """n = 40: one K slice; 1000 and 3000: split-K atomics and padded columns."""
    p = rand_params(25 + n)
    gpu_ctx.surrogate_set_params(p)
    _, _, _, x = _rand_features(26 + n, n)
    y = (np.random.default_rng(27 + n).uniform(size=n) < 0.4).astype(np.float32)
    print(f"[case] grad n={n}", flush=True)
    loss, g = gpu_ctx.surrogate_grad(x, y)
    lo, go = po.mlp_backward(p, x, y, bf16=True)
    assert abs(loss - lo) < 1e-3 * max(1.0, lo)
    for a, b in BLOCKS:
        s = slice(OFF[a], OFF[b])
        rel = np.linalg.norm(g[s] - go[s]) / np.linalg.norm(go[s])
        assert rel < 3e-2, (a, rel)
    assert abs(g[OFF["b4"]] - go[OFF["b4"]]) < 1e-2 * abs(go[OFF["b4"]]) + 1e-6
    cos = float(g @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    assert cos > 0.999


def test_adam_step_matches_oracle(gpu_ctx):
    p = rand_params(31)
    gpu_ctx.surrogate_set_params(p)  # resets the Adam state
    _, _, _, x = _rand_features(32, 512)
    y = (x[:, 0] > 0).astype(np.float32)
    _, g = gpu_ctx.surrogate_grad(x, y)
    gpu_ctx.surrogate_step(x, y, 1e-3)
    got = gpu_ctx.surrogate_params()
    ref = p.copy()
    m, v = np.zeros_like(p), np.zeros_like(p)
    po.mlp_adam(ref, g, m, v, 1, lr=1e-3)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


def test_training_learns_a_separable_rule(gpu_ctx):
    gpu_ctx.surrogate_init(5)
    _, _, _, x = _rand_features(33, 4096)
    y = (x[:, 3] + x[:, 40] > 0.5).astype(np.float32)
    first = gpu_ctx.surrogate_step(x, y, 3e-3)
    for _ in range(300):
        last = gpu_ctx.surrogate_step(x, y, 3e-3)
    p = gpu_ctx.surrogate_forward(x)
    acc = float(np.mean((p > 0.5) == (y > 0.5)))
    assert last < 0.3 * first and acc > 0.97, (first, last, acc)


DEPTH = 8  # leaf edge 7.0: closed surfaces at 1M points (mean spacing ~4.4)


@functools.lru_cache(maxsize=1)
def _small_scene():
    """The config-2 generator at 1M points on a depth-8 tree."""
    sc = q.Scene(2, seed=0, npoints=1000000)
    pnt, nrm, col = sc.points()
    nodes = q.octree_from_points(pnt, sc.basecube, DEPTH).nodes()
    return sc, nrm, col, nodes


def _upload(ctx, nodes, col, nrm):
    ctx.upload_octree(nodes)
    ctx.upload_model(col, nrm)
    ctx.upload_octree(np.zeros((1, 12), np.int32), dynamic=True)


def test_collected_labels_match_oracle_shadow_trace(gpu_ctx):
    sc, nrm, col, nodes = _small_scene()
    _upload(gpu_ctx, nodes, col, nrm)
    gpu_ctx.surrogate_init(1)
    gpu_ctx.surrogate_reserve(100000, 7)
    pos, ang = sc.poses(3, seed=5)
    for i in range(3):
        print(f"[case] collect view {i}", flush=True)
        gpu_ctx.render(q.view(pos[i], ang[i], 128, 96, sc.basecube, DEPTH), sc.lights,
                       q.OUT_FLAGS | q.RENDER_COLLECT)
    n, lit = gpu_ctx.surrogate_samples()
    assert n > 1000 and 0 < lit < n
    smp = gpu_ctx.surrogate_read_samples()
    assert smp.shape == (n, 10) and int(smp[:, 9].sum()) == lit
    np.testing.assert_array_equal(smp[:, 6:9], np.repeat(np.asarray(sc.lights, np.float32)[:1], n, 0))
    osc = po.OracleScene(nodes, col, nrm, sc.basecube, DEPTH)
    pick = np.random.default_rng(6).choice(n, 300, replace=False)
    L = smp[pick, 6:9]
    ip = smp[pick, 0:3]
    res = osc.trace_many(L, ip - L)
    li = np.where((res["status"] == 1)[:, None], res["isp"][:, :3], 0.0).astype(np.float32)
    d = li - ip
    sqr = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    label = np.where(np.float32(15.0) < sqr, 0.0, 1.0)
    assert np.all(res["status"] != -1)
    np.testing.assert_array_equal(smp[pick, 9], label)
    # the normal is the hit leaf's model normal
    assert np.all(np.abs(np.linalg.norm(smp[:, 3:6], axis=1) - 1) < 1e-3)


def test_surrogate_render(gpu_ctx):
    """Train online on exact renders, then render held-out views with the
    surrogate: primary outputs identical; every query the surrogate does not
    decide is traced exactly; its decisions follow its own probabilities; the
    trained network decides most queries with few mismatches, where the
    untrained one is wrong far more often."""
    sc, nrm, col, nodes = _small_scene()
    _upload(gpu_ctx, nodes, col, nrm)
    pos, ang = sc.poses(56, seed=9)
    views = [q.view(pos[i], ang[i], 160, 120, sc.basecube, DEPTH) for i in range(56)]
    print("[case] online training, 48 frames", flush=True)
    gpu_ctx.surrogate_init(3)
    untrained = gpu_ctx.surrogate_params()
    info = S.train_online(gpu_ctx, views[:48], sc.lights, sc.basecube, nsamples=1 << 20, stride=2,
                          steps_per_frame=24, batch=2048, lr=2e-3, seed=3, init=False, final_steps=400)
    print("[surrogate] training", info, flush=True)
    f = info["lit_fraction"]
    entropy = -(f * np.log(f) + (1 - f) * np.log(1 - f))  # loss of the best constant predictor
    assert info["loss_last"] < 0.5 * entropy, info
    trained = gpu_ctx.surrogate_params()
    lights = np.asarray(sc.lights, np.float32)
    stats = {}
    for name, params in (("trained", trained), ("untrained", untrained)):
        gpu_ctx.surrogate_set_params(params)
        bad = tot = pred = 0
        for v in views[48:]:
            print(f"[case] {name} surrogate render", flush=True)
            gpu_ctx.render(v, lights, q.OUT_ALL)
            ex = gpu_ctx.read_frame(0)
            gpu_ctx.render(v, lights, q.OUT_ALL | q.RENDER_SURROGATE)
            su = gpu_ctx.read_frame(0)
            pred += gpu_ctx.stats()["predicted"]
            for k in ("hit", "leaf", "dyn", "isp"):
                np.testing.assert_array_equal(su[k], ex[k])
            # undecided pixels are exact
            keep = (su["flags"] & q.PX_PREDICTED) == 0
            np.testing.assert_array_equal(su["shadow"][keep], ex["shadow"][keep])
            np.testing.assert_array_equal(su["rgba8"][keep], ex["rgba8"][keep])
            # decided pixels follow the network's probability
            sh = ((ex["flags"] & q.PX_SHADED) != 0) & ((ex["flags"] & q.PX_DISCARD) == 0)
            ii = np.nonzero(sh.ravel())[0]
            isp = ex["isp"].reshape(-1, 4)[ii, :3]
            pn = nrm[ex["hit"].ravel()[ii]]
            pr = gpu_ctx.surrogate_predict(isp, pn, np.repeat(lights[:1], ii.size, 0), sc.basecube)
            bit = su["shadow"].ravel()[ii] & 1
            ebit = ex["shadow"].ravel()[ii] & 1
            assert np.all(bit[pr < 0.099] == 0) and np.all(bit[pr > 0.901] == 1)
            mid = (pr > 0.101) & (pr < 0.899)
            np.testing.assert_array_equal(bit[mid], ebit[mid])
            b, t = S.shadow_mismatch(ex, su, 1)
            bad, tot = bad + b, tot + t
        stats[name] = (bad / tot, pred / tot)
    print("[surrogate]", info, stats, flush=True)
    mis_t, frac_t = stats["trained"]
    mis_u, frac_u = stats["untrained"]
    assert frac_t > 0.6 and mis_t < 0.03, stats
    assert mis_t * frac_u < 0.5 * mis_u * frac_t or mis_u > 0.1, stats


def test_surrogate_render_three_lights(gpu_ctx):
    """Three lights with an untrained network and a band at its 30% and 70%
    quantiles: every light keeps a different set of undecided queries (the
    fallback lists differ in length and content), and the one shadow launch
    serves all three lists. Per light: undecided queries equal the exact
    render, decided ones follow the network's probability."""
    sc, nrm, col, nodes = _small_scene()
    _upload(gpu_ctx, nodes, col, nrm)
    lights = np.array([sc.lights[0], (300.0, 1500.0, 400.0), (1400.0, 900.0, 1500.0)], np.float32)
    pos, ang = sc.poses(3, seed=21)
    views = [q.view(pos[i], ang[i], 160, 120, sc.basecube, DEPTH) for i in range(3)]
    gpu_ctx.surrogate_init(11)
    try:
        _three_lights(gpu_ctx, sc, nrm, lights, views)
    finally:
        gpu_ctx.surrogate_set_band(0.1, 0.9)


def _three_lights(gpu_ctx, sc, nrm, lights, views):
    checked = 0
    for vi, v in enumerate(views):
        print(f"[case] three-light surrogate render, view {vi}", flush=True)
        gpu_ctx.render(v, lights, q.OUT_ALL)
        ex = gpu_ctx.read_frame(0)
        sh = ((ex["flags"] & q.PX_SHADED) != 0) & ((ex["flags"] & q.PX_DISCARD) == 0)
        ii = np.nonzero(sh.ravel())[0]
        isp = ex["isp"].reshape(-1, 4)[ii, :3]
        pn = nrm[ex["hit"].ravel()[ii]]
        pr = np.stack([gpu_ctx.surrogate_predict(isp, pn, np.repeat(lights[l:l + 1], ii.size, 0), sc.basecube)
                       for l in range(3)])
        lo, hi = (float(x) for x in np.quantile(pr, [0.3, 0.7]))
        gpu_ctx.surrogate_set_band(lo, hi)
        gpu_ctx.render(v, lights, q.OUT_ALL | q.RENDER_SURROGATE)
        su = gpu_ctx.read_frame(0)
        for k in ("hit", "leaf", "dyn", "isp"):
            np.testing.assert_array_equal(su[k], ex[k])
        undecided = []
        for l in range(3):
            bit = (su["shadow"].ravel()[ii] >> l) & 1
            ebit = (ex["shadow"].ravel()[ii] >> l) & 1
            p = pr[l]
            eps = 1e-3 * (hi - lo) + 1e-4
            assert np.all(bit[p < lo - eps] == 0) and np.all(bit[p > hi + eps] == 1), f"light {l}"
            mid = (p > lo + eps) & (p < hi - eps)
            np.testing.assert_array_equal(bit[mid], ebit[mid], err_msg=f"light {l}")
            undecided.append(int(np.count_nonzero(mid)))
            checked += int(mid.sum())
        assert len(set(undecided)) > 1, undecided  # the fallback lists differ
    assert checked > 1000


def test_surrogate_requires_parameters(gpu_ctx):
    with q.Context(0) as c2:
        sc, nrm, col, nodes = _small_scene()
        _upload(c2, nodes, col, nrm)
        v = q.view((700.0, 800.0, 600.0), (0.5, 0.1), 32, 32, sc.basecube, DEPTH)
        with pytest.raises(q.QbtError):
            c2.render(v, sc.lights, q.OUT_ALL | q.RENDER_SURROGATE)
        with pytest.raises(q.QbtError):
            c2.surrogate_set_band(0.0, 0.9)
        with pytest.raises(q.QbtError):
            c2.surrogate_forward(np.zeros((4, 42), np.float32))
