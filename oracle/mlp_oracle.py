"""Host restatement of the BiCG warm-start network (test infrastructure only).

Forward pass, loss and straight-through gradient of the network defined in
bubbleensemble_implicitextension_amd/csrc/bex_mlp.h, in float64 with the same bfloat16
rounding points as the device kernels (inputs, weights of W1/W2/w3, both hidden activations).
The reference has no counterpart: it always starts BiCG from the uncoupled acceleration
(SourceCodes/CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh:494-499).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline import this module; it is the
checker, never the product path.

Pins (tests/test_mlp_oracle.py): bf16 rounding against torch's float32 -> bfloat16 conversion;
the forward pass against torch.nn.functional on CPU; the gradient (rounding off) against
central finite differences of the loss.
"""
from __future__ import annotations

import numpy as np

IN = 5
H = 64
OFF_W1 = 0
OFF_W2 = 6 * H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1
TARGET_CLAMP = 5.0


def bf16_round(x):
    """Nearest bfloat16 (ties to even) of float32(x), as float64."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_W2].reshape(6, H)
    W2 = p[OFF_W2:OFF_B2].reshape(H, H)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, W2, b2, w3, b3


def init_params(seed=0):
    """He-uniform initialisation of the two hidden layers, small output layer."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    lim1 = np.sqrt(6.0 / 6.0)
    p[OFF_W1:OFF_W2] = rng.uniform(-lim1, lim1, 6 * H) * 0.5
    p[OFF_W1 + 5 * H:OFF_W2] = 0.0  # first-layer bias
    lim2 = np.sqrt(6.0 / H)
    p[OFF_W2:OFF_B2] = rng.uniform(-lim2, lim2, H * H)
    p[OFF_W3:OFF_B3] = rng.uniform(-1.0, 1.0, H) / np.sqrt(H)
    p[OFF_B3] = 1.0  # y = 1 is the plain guess scale x0 = b
    return p


def forward(p, feats, rounding=True):
    """y for feature rows [n, >=5]; returns y and the intermediates."""
    rnd = bf16_round if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(feats, dtype=np.float64)[:, :IN]
    inp = np.concatenate([rnd(x), np.ones((x.shape[0], 1))], axis=1)
    z1 = inp @ rnd(W1)
    h1 = rnd(np.maximum(z1, 0.0))
    z2 = b2 + h1 @ rnd(W2)
    h2 = rnd(np.maximum(z2, 0.0))
    y = b3 + h2 @ rnd(w3)
    return y, (inp, z1, h1, z2, h2)


def loss_grad(p, rows, rounding=True):
    """Mean squared error of y against column 5 and its straight-through gradient."""
    rows = np.asarray(rows, dtype=np.float64)
    rnd = bf16_round if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, b2, w3, b3 = unpack(p)
    y, (inp, z1, h1, z2, h2) = forward(p, rows, rounding)
    n = rows.shape[0]
    e = y - rows[:, 5]
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = np.zeros(NPARAM)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = h2.T @ dy
    dz2 = (dy[:, None] * rnd(w3)[None, :]) * (z2 > 0)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).reshape(-1)
    dz1 = (dz2 @ rnd(W2).T) * (z1 > 0)
    g[OFF_W1:OFF_W2] = (inp.T @ dz1).reshape(-1)
    return loss, g
