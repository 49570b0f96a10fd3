"""Pins of the surrogate oracle (oracle/surrogate.py) on the CPU."""

import numpy as np
import pytest
import torch

from oracle import surrogate as o

SHAPE = (7, 16, 3)


def _case(batch=9, seed=1):
    rng = np.random.default_rng(seed)
    params = o.init(*SHAPE, seed=seed).astype(np.float64)
    # Nonzero biases so the bias rows are exercised.
    for w in o.split(params, *SHAPE):
        w[-1] = rng.normal(0, 0.3, w.shape[1])
    x = rng.normal(size=(batch, SHAPE[0]))
    t = rng.normal(size=(batch, SHAPE[2]))
    return params, x, t


def test_forward_backward_match_torch_autograd():
    params, x, t = _case()
    W = [torch.tensor(w, dtype=torch.float64, requires_grad=True) for w in o.split(params, *SHAPE)]
    xt = torch.tensor(x)
    h = xt
    for i, w in enumerate(W):
        h = h @ w[:-1] + w[-1]
        if i < 2:
            h = torch.relu(h)
    loss = torch.mean((h - torch.tensor(t)) ** 2)
    loss.backward()
    value, grad = o.backward(params, x, t, *SHAPE)
    assert value == pytest.approx(loss.item(), rel=1e-12)
    np.testing.assert_allclose(o.forward(params, x, *SHAPE)[2], h.detach().numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(grad, np.concatenate([w.grad.numpy().ravel() for w in W]), rtol=1e-10, atol=1e-12)


def test_backward_matches_finite_differences():
    params, x, t = _case(batch=5, seed=2)
    _, grad = o.backward(params, x, t, *SHAPE)
    rng = np.random.default_rng(3)
    eps = 1e-6
    for i in rng.choice(len(params), 40, replace=False):
        e = np.zeros_like(params)
        e[i] = eps
        fd = (o.loss(params + e, x, t, *SHAPE) - o.loss(params - e, x, t, *SHAPE)) / (2 * eps)
        assert grad[i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_adam_closed_forms():
    rng = np.random.default_rng(4)
    w = rng.normal(size=50)
    g = rng.normal(size=50)
    lr = 1e-3
    # First step: m_hat = g, v_hat = g^2, so the update is -lr sign(g) (up to eps).
    w1, m, v = o.adam(w, g, np.zeros(50), np.zeros(50), 1, lr)
    np.testing.assert_allclose(w1 - w, -lr * np.sign(g) * np.abs(g) / (np.abs(g) + 1e-8), rtol=1e-12)
    # Constant gradient: the bias-corrected moments stay g and g^2 at every step.
    for step in range(2, 6):
        prev = w1
        w1, m, v = o.adam(w1, g, m, v, step, lr)
        np.testing.assert_allclose(w1 - prev, -lr * g / (np.abs(g) + 1e-8), rtol=1e-9)
    # Zero gradient: only the decoupled decay acts, w <- (1 - lr decay) w.
    w2, _, _ = o.adam(w, np.zeros(50), np.zeros(50), np.zeros(50), 3, lr, weight_decay=0.1)
    np.testing.assert_allclose(w2, (1 - lr * 0.1) * w, rtol=1e-12)


def test_bf16_matches_torch():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(size=1000) * 10.0 ** rng.integers(-6, 6, 1000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38])]).astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(o.bf16(x), ref)


def test_patch_matches_brute_force():
    rng = np.random.default_rng(6)
    h, half = 0.003, (0.0005, 0.02, 0.01)
    n = 400
    x = rng.uniform(-0.03, 0.03, size=(n, 3))
    damage = rng.uniform(0, 1, n)
    instance = rng.integers(0, 2, n)
    pos = np.array([[0.001, -0.002, 0.012], [0.0, 0.0, 0.008]])
    ang = 0.3
    rot = np.array([[np.sin(ang / 2), 0, 0, np.cos(ang / 2)], [0, 0, 0, 1.0]])
    out, _ = o.patch(x, damage, instance, pos, rot, h, half)
    # Brute force: rotation matrix of each knife, every particle against every cell's bounds.
    expected = np.zeros_like(out)
    for i in range(2):
        qx, qy, qz, qw = rot[i]
        R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                      [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                      [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
        for p in np.nonzero(instance == i)[0]:
            local = R.T @ (x[p] - pos[i])
            for a in range(8):
                for b in range(8):
                    for c in range(8):
                        lo = np.array([-4 * h + a * h, -half[1] + b * half[1] / 4, -half[2] - 4 * h + c * h])
                        hi = lo + np.array([h, half[1] / 4, h])
                        if (local >= lo).all() and (local < hi).all():
                            expected[i, 0, a, b, c] += 1
                            expected[i, 1, a, b, c] += damage[p]
    assert expected[:, 0].sum() > 20
    np.testing.assert_allclose(out, expected, atol=1e-12)
