"""FP64 restatement of the continuation-value network (3 -> 64 -> 64 -> 64 -> 1, tanh) and its
gradient, with optional emulation of the BF16 operand rounding of the GPU kernels.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the HIP forward / backward / Adam
kernels. There is no reference counterpart: the reference's only regression method is the quadratic
LSM fit of ``payoff/lsm.rs:73-107``; this network is a north-star addition.

Parameter vector (float32 master copy on the GPU, same order here), N_PARAMS = 8641:
  W1 [64][4]  (columns: inputs u0, u1, u2, then the layer-1 bias)   offset 0
  W2 [64][64] (out, in)                                              offset 256
  b2 [64]                                                            offset 4352
  W3 [64][64]                                                        offset 4416
  b3 [64]                                                            offset 8512
  w4 [64]                                                            offset 8576
  b4 [1]                                                             offset 8640

Rounding points of the GPU kernels (matrix-core operands are BF16, round-to-nearest-even of the
float32 value; products accumulate in float32):
  forward:  bf(u), bf(W1), bf(W2), bf(W3), bf(h1), bf(h2); b2, b3, w4, b4 and h3 stay float32.
  backward: bf(dz3), bf(dz2), bf(dz1) enter every product they feed (weight gradients and the
            back-propagated dh), paired with bf(h2), bf(h1), bf(u) and bf(W3), bf(W2).
With ``emulate_bf16=False`` nothing is rounded and the gradient is the exact FP64 derivative of
the loss, which the finite-difference pin in tests checks.
"""
from __future__ import annotations

import numpy as np

H = 64
OFF_W1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, OFF_W4, OFF_B4 = 0, 256, 4352, 4416, 8512, 8576, 8640
N_PARAMS = 8641


def bf16(x):
    """Round to the nearest BF16 value (ties to even) of the float32 value of x; float64 result."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    th = np.asarray(theta, dtype=np.float64)
    return (th[OFF_W1:OFF_W2].reshape(H, 4), th[OFF_W2:OFF_B2].reshape(H, H), th[OFF_B2:OFF_W3],
            th[OFF_W3:OFF_B3].reshape(H, H), th[OFF_B3:OFF_W4], th[OFF_W4:OFF_B4], th[OFF_B4])


def forward(theta, u, emulate_bf16=True):
    """u: [n][3] inputs. Returns (y [n], cache)."""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    W1, W2, b2, W3, b3, w4, b4 = unpack(theta)
    n = u.shape[0]
    ub = np.empty((4, n))
    ub[:3] = q(np.asarray(u, dtype=np.float64).T)
    ub[3] = 1.0
    z1 = q(W1) @ ub
    h1 = np.tanh(z1)
    h1b = q(h1)
    z2 = q(W2) @ h1b + b2[:, None]
    h2 = np.tanh(z2)
    h2b = q(h2)
    z3 = q(W3) @ h2b + b3[:, None]
    h3 = np.tanh(z3)
    y = w4 @ h3 + b4
    return y, dict(ub=ub, h1=h1, h1b=h1b, h2=h2, h2b=h2b, h3=h3, q=q)


def loss_and_grad(theta, u, target, mask, batch, emulate_bf16=True):
    """Loss = sum(mask (y - target)^2) / batch and its gradient with respect to theta."""
    W1, W2, b2, W3, b3, w4, b4 = unpack(theta)
    y, c = forward(theta, u, emulate_bf16)
    q = c["q"]
    res = (y - target) * mask
    loss = float(np.sum(res * (y - target)) / batch)
    dy = 2.0 * res / batch
    g = np.zeros(N_PARAMS)
    g[OFF_W4:OFF_B4] = c["h3"] @ dy
    g[OFF_B4] = dy.sum()
    dz3 = w4[:, None] * dy[None, :] * (1.0 - c["h3"] ** 2)
    g[OFF_B3:OFF_W4] = dz3.sum(axis=1)
    dz3b = q(dz3)
    g[OFF_W3:OFF_B3] = (dz3b @ c["h2b"].T).ravel()
    dh2 = q(W3).T @ dz3b
    dz2 = dh2 * (1.0 - c["h2"] ** 2)
    g[OFF_B2:OFF_W3] = dz2.sum(axis=1)
    dz2b = q(dz2)
    g[OFF_W2:OFF_B2] = (dz2b @ c["h1b"].T).ravel()
    dh1 = q(W2).T @ dz2b
    dz1 = dh1 * (1.0 - c["h1"] ** 2)
    dz1b = q(dz1)
    g[OFF_W1:OFF_W2] = (dz1b @ c["ub"].T).ravel()
    return loss, g


def adam(theta, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in float32 arithmetic; step counts from 1."""
    f = np.float32
    theta, m, v, grad = (np.asarray(x, dtype=np.float32) for x in (theta, m, v, grad))
    m = f(beta1) * m + f(1.0 - beta1) * grad
    v = f(beta2) * v + f(1.0 - beta2) * grad * grad
    mh = m / f(1.0 - beta1 ** step)
    vh = v / f(1.0 - beta2 ** step)
    theta = theta - f(lr) * mh / (np.sqrt(vh) + f(eps))
    return theta, m, v
