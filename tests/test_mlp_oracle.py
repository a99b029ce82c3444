"""Pins of the host restatement of the BiCG warm-start network (oracle/mlp_oracle.py).

Each check uses an independent route: torch's own float32 -> bfloat16 conversion for the
rounding, torch.nn.functional on the CPU for the forward pass, central finite differences of
the loss for the gradient.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def _params(seed=3):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed).astype(np.float64)
    p += rng.normal(0.0, 0.05, p.size)  # non-zero biases, generic weights
    return p.astype(np.float32)


def _rows(n=200, seed=4):
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, 10), dtype=np.float32)
    rows[:, :5] = rng.normal(0.0, 1.0, (n, 5))
    rows[:, 5] = rng.normal(1.0, 0.3, n)
    return rows


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e4, 2000),
                        rng.normal(0, 1e-6, 2000)]).astype(np.float32)
    # exact ties: low 16 bits = 0x8000 with even and odd kept parts
    ties = (np.arange(200, dtype=np.uint32) << 16 | 0x8000).astype(np.uint32) + 0x3F800000
    x = np.concatenate([x, ties.view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16_round(x), ref)


def test_forward_matches_torch_functional():
    p = _params()
    rows = _rows()
    y, _ = mo.forward(p, rows)

    def bf(t):
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)

    W1 = torch.from_numpy(p[0:384].reshape(6, 64)).double()
    W2 = torch.from_numpy(p[384:384 + 4096].reshape(64, 64)).double()
    b2 = torch.from_numpy(p[4480:4544]).double()
    w3 = torch.from_numpy(p[4544:4608]).double()
    b3 = float(p[4608])
    x = bf(torch.from_numpy(rows[:, :5]).double())
    x = torch.cat([x, torch.ones(x.shape[0], 1, dtype=torch.float64)], 1)
    h1 = bf(torch.nn.functional.relu(torch.nn.functional.linear(x, bf(W1).T)))
    h2 = bf(torch.nn.functional.relu(torch.nn.functional.linear(h1, bf(W2).T, b2)))
    yt = torch.nn.functional.linear(h2, bf(w3)[None, :]).squeeze(1) + b3
    np.testing.assert_allclose(y, yt.numpy(), rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    p = _params(5).astype(np.float64)
    rows = _rows(64, 6)
    loss, g = mo.loss_grad(p, rows, rounding=False)
    rng = np.random.default_rng(7)
    # parameters from every block: W1 (incl. bias row), W2, b2, w3, b3
    pick = np.concatenate([rng.choice(np.arange(0, 384), 20, replace=False),
                           rng.choice(np.arange(384, 4480), 40, replace=False),
                           rng.choice(np.arange(4480, 4544), 10, replace=False),
                           rng.choice(np.arange(4544, 4608), 10, replace=False), [4608]])
    h = 1e-6
    for q in pick:
        pp = p.copy()
        pm = p.copy()
        pp[q] += h
        pm[q] -= h
        fd = (mo.loss_grad(pp, rows, rounding=False)[0] -
              mo.loss_grad(pm, rows, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])) + 1e-8, (q, fd, g[q])


def test_planted_gradient_mistake_is_caught():
    """A dropped ReLU mask in the backward pass must fail the finite-difference check."""
    p = _params(8).astype(np.float64)
    rows = _rows(32, 9)
    W1, W2, b2, w3, b3 = mo.unpack(p)
    y, (inp, z1, h1, z2, h2) = mo.forward(p, rows, rounding=False)
    dy = 2.0 * (y - rows[:, 5]) / rows.shape[0]
    dz2_wrong = dy[:, None] * w3[None, :]  # mask (z2 > 0) dropped
    g_wrong = (h1.T @ dz2_wrong).reshape(-1)
    _, g = mo.loss_grad(p, rows, rounding=False)
    assert np.max(np.abs(g_wrong - g[mo.OFF_W2:mo.OFF_B2])) > 1e-6
