"""FP64 restatement of the continuation-value network (3 -> 64 -> 64 -> 64 -> 1, tanh) and its
gradient, with optional emulation of the BF16 operand rounding of the GPU kernels.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the HIP forward / backward / Adam
kernels. There is no reference counterpart: the reference's only regression method is the quadratic
LSM fit of ``payoff/lsm.rs:73-107``; this network is an addition of this library (the Bermudan
exposure surrogate).

Parameter vector (float32 master copy on the GPU, same order here), N_PARAMS = 8641:
  W1 [64][4]  (columns: inputs u0, u1, u2, then the layer-1 bias)   offset 0
  W2 [64][64] (out, in)                                              offset 256
  b2 [64]                                                            offset 4352
  W3 [64][64]                                                        offset 4416
  b3 [64]                                                            offset 8512
  w4 [64]                                                            offset 8576
  b4 [1]                                                             offset 8640

Forward as the GPU kernels evaluate it. With c = 2 log2(e) (float32) the activation is carried as
r = 1 / (1 + 2^(c z)), so that tanh(z) = 1 - 2r; the next layer absorbs r -> 1 - 2r:
  z1s = bf(c W1) [bf(u); 1]                        r1 = 1 / (1 + 2^z1s)
  z2s = bf(-2c W2) bf(r1) + c (b2 + W2 1)          r2 = ...        (likewise layer 3)
  y   = (b4 + sum w4) + (-2 w4) . r3
bf() rounds the float32 value to BF16 (ties to even); products accumulate in float32 on the GPU and
in float64 here; the folded biases are float32 on the GPU. Backward (in tanh space, h = 1 - 2r):
  dy = 2 mask (y - t) / batch, dz3 = w4 dy (1 - h3^2), dW3 = bf(dz3) bf(h2)^T,
  dh2 = bf(W3)^T bf(dz3), dz2 = dh2 (1 - h2^2), dW2 = bf(dz2) bf(h1)^T, dh1 = bf(W2)^T bf(dz2),
  dz1 = dh1 (1 - h1^2), dW1 = bf(dz1) [bf(u); 1]^T, db2/db3 = sums of dz2/dz3, dw4 = h3 dy.
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


ACT_SCALE = float(np.float32(2.8853900817779268))  # 2 log2(e) as the kernels hold it


def _scaled(q, w, scale, emulate):
    if not emulate:
        return scale * np.asarray(w, dtype=np.float64)
    return q(np.float32(scale) * np.asarray(w, dtype=np.float32))


def forward(theta, u, emulate_bf16=True):
    """u: [n][3] inputs. Returns (y [n], cache)."""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    c = ACT_SCALE
    W1, W2, b2, W3, b3, w4, b4 = unpack(theta)
    n = u.shape[0]
    ub = np.empty((4, n))
    ub[:3] = q(np.asarray(u, dtype=np.float64).T)
    ub[3] = 1.0
    r1 = 1.0 / (1.0 + np.exp2(_scaled(q, W1, c, emulate_bf16) @ ub))
    z2s = _scaled(q, W2, -2.0 * c, emulate_bf16) @ q(r1) + (c * (b2 + W2.sum(axis=1)))[:, None]
    r2 = 1.0 / (1.0 + np.exp2(z2s))
    z3s = _scaled(q, W3, -2.0 * c, emulate_bf16) @ q(r2) + (c * (b3 + W3.sum(axis=1)))[:, None]
    r3 = 1.0 / (1.0 + np.exp2(z3s))
    y = (b4 + w4.sum()) + (-2.0 * w4) @ r3
    h1, h2, h3 = 1.0 - 2.0 * r1, 1.0 - 2.0 * r2, 1.0 - 2.0 * r3
    return y, dict(ub=ub, h1=h1, h2=h2, h3=h3, q=q)


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
    g[OFF_W3:OFF_B3] = (dz3b @ q(c["h2"]).T).ravel()
    dh2 = q(W3).T @ dz3b
    dz2 = dh2 * (1.0 - c["h2"] ** 2)
    g[OFF_B2:OFF_W3] = dz2.sum(axis=1)
    dz2b = q(dz2)
    g[OFF_W2:OFF_B2] = (dz2b @ q(c["h1"]).T).ravel()
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
