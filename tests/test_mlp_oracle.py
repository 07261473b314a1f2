"""Pins of the oracle's shadow-visibility MLP (forward, backward, Adam,
features) against independent references: torch autograd in float64,
central finite differences, torch.optim.Adam, closed forms of the feature
encoding. CPU only."""
import math

import numpy as np
import pytest
import torch

from oracle import pyoracle as po

N_W1, N_B1 = 64 * 42, 64
OFF = {"W1": 0, "b1": 2688, "W2": 2752, "b2": 6848, "W3": 6912, "b3": 11008, "W4": 11072, "b4": 11136}


def rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(po.MLP_NPARAM, np.float32)
    p[OFF["W1"]:OFF["b1"]] = rng.uniform(-1, 1, 64 * 42) * math.sqrt(6 / 42) * scale
    p[OFF["b1"]:OFF["W2"]] = rng.normal(0, 0.1, 64)
    p[OFF["W2"]:OFF["b2"]] = rng.uniform(-1, 1, 4096) * math.sqrt(6 / 64) * scale
    p[OFF["b2"]:OFF["W3"]] = rng.normal(0, 0.1, 64)
    p[OFF["W3"]:OFF["b3"]] = rng.uniform(-1, 1, 4096) * math.sqrt(6 / 64) * scale
    p[OFF["b3"]:OFF["W4"]] = rng.normal(0, 0.1, 64)
    p[OFF["W4"]:OFF["b4"]] = rng.uniform(-1, 1, 64) * math.sqrt(6 / 64) * scale
    p[OFF["b4"]] = 0.05
    return p


def torch_model(p, dtype=torch.float64):
    """Parameter tensors (requires_grad) in the flat layout's shapes."""
    t = torch.tensor(p.astype(np.float64), dtype=dtype)
    W1 = t[OFF["W1"]:OFF["b1"]].reshape(64, 42).clone().requires_grad_()
    b1 = t[OFF["b1"]:OFF["W2"]].clone().requires_grad_()
    W2 = t[OFF["W2"]:OFF["b2"]].reshape(64, 64).clone().requires_grad_()
    b2 = t[OFF["b2"]:OFF["W3"]].clone().requires_grad_()
    W3 = t[OFF["W3"]:OFF["b3"]].reshape(64, 64).clone().requires_grad_()
    b3 = t[OFF["b3"]:OFF["W4"]].clone().requires_grad_()
    W4 = t[OFF["W4"]:OFF["b4"]].reshape(1, 64).clone().requires_grad_()
    b4 = t[OFF["b4"]:].clone().requires_grad_()
    return [W1, b1, W2, b2, W3, b3, W4, b4]


def torch_forward(ps, x):
    W1, b1, W2, b2, W3, b3, W4, b4 = ps
    h = torch.relu(x @ W1.T + b1)
    h = torch.relu(h @ W2.T + b2)
    h = torch.relu(h @ W3.T + b3)
    return (h @ W4.T + b4).squeeze(1)


def flat_grad(ps):
    return np.concatenate([q.grad.detach().numpy().ravel() for q in ps])


def data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 42)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.5).astype(np.float32)
    return x, y


def test_forward_fp32_matches_torch():
    p = rand_params(1)
    x, _ = data(2, 64)
    prob, logit = po.mlp_forward(p, x, bf16=False)
    ps = torch_model(p)
    ref = torch_forward(ps, torch.tensor(x, dtype=torch.float64)).detach().numpy()
    np.testing.assert_allclose(logit, ref, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(prob, 1 / (1 + np.exp(-ref)), rtol=1e-5, atol=1e-6)


def test_backward_fp32_matches_torch_autograd():
    p = rand_params(3)
    x, y = data(4, 48)
    loss, g = po.mlp_backward(p, x, y, bf16=False)
    ps = torch_model(p)
    logit = torch_forward(ps, torch.tensor(x, dtype=torch.float64))
    ref_loss = torch.nn.functional.binary_cross_entropy_with_logits(logit, torch.tensor(y, dtype=torch.float64))
    ref_loss.backward()
    ref = flat_grad(ps)
    assert abs(loss - ref_loss.item()) < 1e-5
    scale = np.abs(ref).max()
    np.testing.assert_allclose(g, ref, rtol=1e-4, atol=1e-6 * scale)
    # every block is exercised (a dropped term would leave one at zero)
    for name in OFF:
        lo = OFF[name]
        hi = {"W1": OFF["b1"], "b1": OFF["W2"], "W2": OFF["b2"], "b2": OFF["W3"], "W3": OFF["b3"],
              "b3": OFF["W4"], "W4": OFF["b4"], "b4": OFF["b4"] + 1}[name]
        assert np.abs(g[lo:hi]).max() > 0, name


def test_backward_fp32_finite_differences():
    p = rand_params(5)
    x, y = data(6, 8)
    _, g = po.mlp_backward(p, x, y, bf16=False)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(OFF["b1"], 6, replace=False), OFF["b1"] + rng.choice(64, 3),
                          OFF["W2"] + rng.choice(4096, 6), OFF["W3"] + rng.choice(4096, 6),
                          OFF["W4"] + rng.choice(64, 4), [OFF["b4"]]])

    def loss_at(q):  # float64 loss of the fp32 oracle forward
        prob, logit = po.mlp_forward(q, x, bf16=False)
        lg = logit.astype(np.float64)
        return float(np.mean(np.logaddexp(0, lg) - y * lg))

    l0 = loss_at(p)
    checked = 0
    for i in idx:
        h = 4e-3 * max(1.0, abs(float(p[i])))
        qp, qm = p.copy(), p.copy()
        qp[i] += h
        qm[i] -= h
        lp, lm = loss_at(qp), loss_at(qm)
        fwd, bwd, fd = (lp - l0) / h, (l0 - lm) / h, (lp - lm) / (2 * h)
        if abs(fwd - bwd) > 2e-2 * abs(fd) + 2e-5:
            continue  # a ReLU kink inside [p - h, p + h]: the one-sided slopes differ
        if abs(g[i]) < 1e-4 and abs(fd) < 1e-4:
            continue
        assert abs(fd - g[i]) <= 5e-3 * max(abs(fd), abs(g[i])) + 2e-5, (int(i), fd, float(g[i]))
        checked += 1
    assert checked >= 12


def bf16(a):
    return torch.tensor(np.asarray(a, np.float32)).bfloat16().float().numpy().astype(np.float64)


def test_bf16_rounding_is_round_to_nearest_even():
    rng = np.random.default_rng(8)
    vals = np.concatenate([rng.normal(0, 10, 2000), [0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 65504.0,
                                                       1e-40, -3.3e38]]).astype(np.float32)
    ours = np.array([po.bf16_round(v) for v in vals], np.float32)
    ref = torch.tensor(vals).bfloat16().float().numpy()
    np.testing.assert_array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert po.bf16_round(1 + 2 ** -8) == 1.0  # tie to even
    assert po.bf16_round(1 + 3 * 2 ** -8) == 1 + 2 ** -6


def test_forward_bf16_matches_rounded_float64():
    """bf16 mode = inputs, weights and post-ReLU activations rounded to bf16,
    float accumulation, fp32 biases."""
    p = rand_params(9)
    x, _ = data(10, 64)
    _, logit = po.mlp_forward(p, x, bf16=True)
    W1 = bf16(p[OFF["W1"]:OFF["b1"]]).reshape(64, 42)
    W2 = bf16(p[OFF["W2"]:OFF["b2"]]).reshape(64, 64)
    W3 = bf16(p[OFF["W3"]:OFF["b3"]]).reshape(64, 64)
    W4 = bf16(p[OFF["W4"]:OFF["b4"]]).reshape(1, 64)
    h = bf16(np.maximum(bf16(x) @ W1.T + p[OFF["b1"]:OFF["W2"]], 0))
    h = bf16(np.maximum(h @ W2.T + p[OFF["b2"]:OFF["W3"]], 0))
    h = bf16(np.maximum(h @ W3.T + p[OFF["b3"]:OFF["W4"]], 0))
    ref = (h @ W4.T).ravel() + p[OFF["b4"]]
    # a rounding-boundary flip of one activation moves the logit by ~2^-8 |w|
    err = np.abs(logit - ref)
    assert np.median(err) < 1e-5 and err.max() < 2e-2


def test_backward_bf16_close_to_fp32():
    p = rand_params(11)
    x, y = data(12, 256)
    l32, g32 = po.mlp_backward(p, x, y, bf16=False)
    l16, g16 = po.mlp_backward(p, x, y, bf16=True)
    assert abs(l32 - l16) < 0.02 * l32
    cos = float(g32 @ g16 / (np.linalg.norm(g32) * np.linalg.norm(g16)))
    assert cos > 0.99


def test_adam_matches_torch():
    rng = np.random.default_rng(13)
    p = rand_params(14)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.tensor(p.copy(), dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.normal(0, 1e-2, p.size).astype(np.float32)
        po.mlp_adam(p, g, m, v, t, lr=3e-3)
        tp.grad = torch.tensor(g)
        opt.step()
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-7)


def test_adam_first_step_closed_form():
    p = np.ones(po.MLP_NPARAM, np.float32)
    g = np.full(po.MLP_NPARAM, 0.5, np.float32)
    g[::2] = -0.25
    m, v = np.zeros_like(p), np.zeros_like(p)
    po.mlp_adam(p, g, m, v, 1, lr=0.01)
    # m_hat = g, v_hat = g^2: the first step is lr * sign(g)
    np.testing.assert_allclose(p, 1 - 0.01 * np.sign(g), rtol=0, atol=1e-6)


def test_features_closed_form():
    bc = [0.0, 1800.0, 1800.0, 1800.0]  # cube [0,1800]^3
    isp = np.array([[0, 0, 0], [900, 450, 1350], [1800, 1800, 1800]], np.float32)
    nrm = np.array([[0, 0, 1], [0.6, 0.8, 0], [1, 0, 0]], np.float32)
    light = np.array([[900, 1700, 900]], np.float32)
    f = po.mlp_features(isp, nrm, light, bc).astype(np.float64)
    u = isp / 1800.0
    for k in range(6):
        for a in range(3):
            np.testing.assert_allclose(f[:, 6 * k + 2 * a], np.sin(2 ** k * np.pi * u[:, a]), atol=2e-5)
            np.testing.assert_allclose(f[:, 6 * k + 2 * a + 1], np.cos(2 ** k * np.pi * u[:, a]), atol=2e-5)
    np.testing.assert_array_equal(f[:, 36:39], nrm)
    np.testing.assert_allclose(f[:, 39:42], np.repeat([[0.5, 1700 / 1800, 0.5]], 3, 0), atol=1e-7)
    # the cube's lower corner is (x, y - w, z - w): a shifted cube gives the same features
    f2 = po.mlp_features(isp + 100, nrm, light + 100, [100.0, 1900.0, 1900.0, 1800.0])
    np.testing.assert_allclose(f2, f, atol=2e-4)
