"""Pins of the dense-network oracle (``oracle/mlp.py``): finite differences of the FP64
forward against the backward, bf16 rounding against torch and hand cases, AdamW against its
textbook form on a one-dimensional quadratic."""

import numpy as np
import pytest
import torch

from oracle import mlp


def _net(sizes, seed):
    rng = np.random.default_rng(seed)
    params = [(rng.standard_normal((c, a)) / np.sqrt(a), rng.standard_normal(c) * 0.1)
              for a, c in zip(sizes[:-1], sizes[1:])]
    acts = [1] * (len(params) - 1) + [0]
    return params, acts


def test_backward_matches_finite_differences():
    sizes = [5, 7, 6, 3]
    params, acts = _net(sizes, 1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5))
    t = rng.standard_normal((4, 3))

    def loss(ps, xx):
        y, _ = mlp.forward(ps, xx, acts, bf16=False)
        return 0.5 * np.sum((y - t) ** 2)

    y, cache = mlp.forward(params, x, acts, bf16=False)
    grads, gx = mlp.backward(params, cache, acts, y - t, bf16=False)
    h = 1e-6
    for k, (w, b) in enumerate(params):
        for arr, g in ((w, grads[k][0]), (b, grads[k][1])):
            num = np.zeros_like(arr)
            for idx in np.ndindex(arr.shape):
                old = arr[idx]
                arr[idx] = old + h
                lp = loss(params, x)
                arr[idx] = old - h
                lm = loss(params, x)
                arr[idx] = old
                num[idx] = (lp - lm) / (2 * h)
            np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-8)
    num = np.zeros_like(x)
    for idx in np.ndindex(x.shape):
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        num[idx] = (loss(params, xp) - loss(params, xm)) / (2 * h)
    np.testing.assert_allclose(gx, num, rtol=1e-6, atol=1e-8)


def test_silu_grad_is_the_derivative():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    np.testing.assert_allclose(mlp.silu_grad(z), (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_bf16_rounding_matches_torch_and_hand_cases():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        [0.0, -0.0, np.inf, -np.inf, 1e-40, 3.4e38]]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mlp.bf16_round(x), want)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: goes to the even mantissa (1)
    assert mlp.bf16_round(np.float32(1 + 2 ** -8)) == 1.0
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert mlp.bf16_round(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)
    assert np.isnan(mlp.bf16_round(np.float32(np.nan)))


def test_adamw_textbook_on_a_quadratic():
    w, m, v = np.array([3.0]), np.zeros(1), np.zeros(1)
    for step in range(1, 2001):
        w, m, v = mlp.adamw(w, 2 * w, m, v, step, lr=1e-2)
    assert abs(w[0]) < 1e-2
    # first step moves by lr against the gradient sign (bias correction makes |m_hat/sqrt(v_hat)| = 1)
    w1, _, _ = mlp.adamw(np.array([3.0]), np.array([6.0]), np.zeros(1), np.zeros(1), 1, lr=0.1)
    assert w1[0] == pytest.approx(2.9, abs=1e-7)
    w1, _, _ = mlp.adamw(np.array([3.0]), np.array([6.0]), np.zeros(1), np.zeros(1), 1, lr=0.1, wd=0.5)
    assert w1[0] == pytest.approx(3.0 * (1 - 0.05) - 0.1, abs=1e-7)
