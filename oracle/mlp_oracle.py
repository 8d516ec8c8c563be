"""Warm-start network restated in numpy: forward pass, analytic backward pass and Adam.

TEST INFRASTRUCTURE ONLY: the tests use it to check the device kernels (akua_mlp_forward,
akua_mlp_gradient, akua_mlp_adam_step); the product never imports it.  There is no reference
counterpart (the reference has no learned component), so it is pinned by central finite
differences of its own FP64 loss and by closed forms on tiny networks.

Two modes:
  exact=True   FP64 throughout, the mathematical network;
  exact=False  the device's rounding points emulated in FP64: normalised features, weights,
               hidden activations and backpropagated gradients rounded to BF16 (round to nearest
               even) exactly where the kernels round them, everything else unrounded.  The kernels'
               FP32 accumulation order is not reproduced, so results agree to accumulation error
               (and rarely one BF16 ulp of an activation), not bit for bit.

Parameter blob (floats): W1[16][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4] fmean[16] finv[16]
ymul[4]; weights are input-major.  Loss: (1/n) sum_rows sum_o w_o (y_o - clip(t_o / ymul_o))^2
with y the normalised outputs and clip to [-8, 8].
"""
from __future__ import annotations

import numpy as np

NIN, NHID, NOUT = 16, 64, 4
OFF_W1 = 0
OFF_B1 = OFF_W1 + NIN * NHID
OFF_W2 = OFF_B1 + NHID
OFF_B2 = OFF_W2 + NHID * NHID
OFF_W3 = OFF_B2 + NHID
OFF_B3 = OFF_W3 + NHID * NOUT
TRAINABLE = OFF_B3 + NOUT
OFF_MEAN = TRAINABLE
OFF_INV = OFF_MEAN + NIN
OFF_YMUL = OFF_INV + NIN
NPARAMS = OFF_YMUL + NOUT
assert (TRAINABLE, NPARAMS) == (5508, 5544)
TARGET_CLIP = 8.0  # normalised targets are clipped to [-8, 8] in the loss


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even); returns float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NIN, NHID), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(NHID, NHID), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(NHID, NOUT), b3=p[OFF_B3:TRAINABLE],
                mean=p[OFF_MEAN:OFF_INV], inv=p[OFF_INV:OFF_YMUL], ymul=p[OFF_YMUL:NPARAMS])


def _net(params, exact):
    q = unpack(params)
    if not exact:
        for k in ("W1", "W2", "W3"):
            q[k] = bf16(q[k])
    return q


def _inputs(q, F, exact):
    F = np.asarray(F, dtype=np.float64)
    if exact:
        return (F - q["mean"]) * q["inv"]
    # the kernel forms (f - mean) * inv in FP32, then rounds to BF16
    xn = (f32(f32(F) - f32(q["mean"])) * f32(q["inv"])).astype(np.float32)
    return bf16(xn)


def forward(params, F, exact: bool = True, cache: bool = False):
    """Denormalised outputs (n, 4); with cache, also the intermediates of the backward pass."""
    q = _net(params, exact)
    x = _inputs(q, F, exact)
    z1 = x @ q["W1"] + q["b1"]
    h1 = np.maximum(z1, 0.0)
    if not exact:
        h1 = bf16(h1)
    z2 = h1 @ q["W2"] + q["b2"]
    h2 = np.maximum(z2, 0.0)
    if not exact:
        h2 = bf16(h2)
    y = h2 @ q["W3"] + q["b3"]
    out = y * q["ymul"]
    if cache:
        return out, dict(q=q, x=x, z1=z1, h1=h1, z2=z2, h2=h2, y=y)
    return out


def loss_and_grad(params, F, T, weights=(1.0, 1.0, 1.0, 1.0), exact: bool = True):
    """(loss, gradient over the TRAINABLE parameters) of the mean weighted squared error."""
    n = len(F)
    w = np.asarray(weights, dtype=np.float64)
    _, c = forward(params, F, exact=exact, cache=True)
    q = c["q"]
    tn = np.clip(np.asarray(T, dtype=np.float64) / q["ymul"], -TARGET_CLIP, TARGET_CLIP)
    diff = c["y"] - tn
    loss = float((w * diff * diff).sum() / n)
    G3 = 2.0 * w * diff / n
    G3p = G3 if exact else bf16(G3)
    G2 = (G3p @ q["W3"].T) * (c["z2"] > 0)
    G2p = G2 if exact else bf16(G2)
    G1 = (G2p @ q["W2"].T) * (c["z1"] > 0)
    G1p = G1 if exact else bf16(G1)
    g = np.zeros(TRAINABLE)
    g[OFF_W1:OFF_B1] = (c["x"].T @ G1p).ravel()
    g[OFF_B1:OFF_W2] = G1.sum(0)
    g[OFF_W2:OFF_B2] = (c["h1"].T @ G2p).ravel()
    g[OFF_B2:OFF_W3] = G2.sum(0)
    g[OFF_W3:OFF_B3] = (c["h2"].T @ G3p).ravel()
    g[OFF_B3:TRAINABLE] = G3.sum(0)
    return loss, g


def loss_only(params, F, T, weights=(1.0, 1.0, 1.0, 1.0)) -> float:
    """FP64 loss (for finite differences)."""
    q = unpack(params)
    y = forward(params, F, exact=True) / q["ymul"]
    d = y - np.clip(np.asarray(T, dtype=np.float64) / q["ymul"], -TARGET_CLIP, TARGET_CLIP)
    return float((np.asarray(weights) * d * d).sum() / len(F))


def adam(params, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (FP64) on the trainable prefix; returns new (params, m, v)."""
    p = np.array(params, dtype=np.float64)
    m = b1 * np.asarray(m, np.float64) + (1 - b1) * g
    v = b2 * np.asarray(v, np.float64) + (1 - b2) * g * g
    c1 = 1.0 / (1.0 - b1 ** t)
    c2 = 1.0 / (1.0 - b2 ** t)
    p[:TRAINABLE] -= lr * (m * c1) / (np.sqrt(v * c2) + eps)
    return p, m, v


def random_params(rng: np.random.Generator, features=None, targets=None) -> np.ndarray:
    """He-normal weights with small biases; normalisation from data when given."""
    p = np.zeros(NPARAMS)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2 / NIN), NIN * NHID)
    p[OFF_B1:OFF_W2] = rng.normal(0, 0.1, NHID)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2 / NHID), NHID * NHID)
    p[OFF_B2:OFF_W3] = rng.normal(0, 0.1, NHID)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1 / NHID), NHID * NOUT)
    p[OFF_B3:TRAINABLE] = rng.normal(0, 0.1, NOUT)
    if features is not None:
        F = np.asarray(features, np.float64)
        p[OFF_MEAN:OFF_INV] = F.mean(0)
        sd = F.std(0)
        p[OFF_INV:OFF_YMUL] = np.where(sd > 0, 1.0 / np.where(sd > 0, sd, 1.0), 1.0)
    else:
        p[OFF_INV:OFF_YMUL] = 1.0
    if targets is not None:
        r = np.sqrt((np.asarray(targets, np.float64) ** 2).mean(0))
        p[OFF_YMUL:NPARAMS] = np.where(r > 0, r, 1.0)
    else:
        p[OFF_YMUL:NPARAMS] = 1.0
    return p.astype(np.float32).astype(np.float64)
