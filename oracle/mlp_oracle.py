"""ORACLE of the warm-start network -- test infrastructure, not product code.

A numpy restatement of the MLP that replica_exchange_wang_landau_amd/csrc/mlp.hip trains and
evaluates: 6 features (padded to 16) -> 128 -> 128 -> 128 -> 1, SiLU hidden activations,
mean-squared-error loss. The reference has no counterpart (it has no learned part), so the
restatement is pinned by central finite differences of its FP64 form and by closed forms.

Two precisions:
- `forward64` / `gradient64`: plain FP64, no rounding (the mathematics);
- `forward_bf16` / `gradient_bf16`: the kernels' arithmetic: BF16 (round to nearest even)
  weights, inputs, activations and pre-activation gradients as matrix-product operands, FP32
  everything else. The kernels accumulate in a different order, so comparisons against them
  use relative-norm tolerances.
"""
from __future__ import annotations

import numpy as np

FEATURES, IN, WIDTH = 6, 16, 128
SHAPES = [("W0", (WIDTH, IN)), ("b0", (WIDTH,)), ("W1", (WIDTH, WIDTH)), ("b1", (WIDTH,)),
          ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)), ("w3", (WIDTH,)), ("b3", (1,))]
PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, at = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = p[at:at + size].reshape(shape)
        at += size
    return out


def pack(parts, dtype=np.float64):
    return np.concatenate([np.asarray(parts[name], dtype=dtype).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """float32 values rounded to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_prime(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def padded(x):
    x = np.asarray(x).reshape(-1, FEATURES)
    out = np.zeros((len(x), IN), dtype=x.dtype)
    out[:, :FEATURES] = x
    return out


def forward64(p, x):
    w = unpack(np.asarray(p, dtype=np.float64))
    a = padded(np.asarray(x, dtype=np.float64))
    zs, hs = [], [a]
    for W, b in (("W0", "b0"), ("W1", "b1"), ("W2", "b2")):
        z = hs[-1] @ w[W].T + w[b]
        zs.append(z)
        hs.append(silu(z))
    y = hs[-1] @ w["w3"] + w["b3"][0]
    return y, zs, hs


def loss64(p, x, t):
    y, _, _ = forward64(p, x)
    return float(np.mean((y - t) ** 2))


def gradient64(p, x, t):
    w = unpack(np.asarray(p, dtype=np.float64))
    y, zs, hs = forward64(p, x)
    n = len(y)
    dy = 2.0 * (y - t) / n
    g = {"b3": np.array([dy.sum()]), "w3": dy @ hs[3]}
    dz = dy[:, None] * w["w3"][None, :] * silu_prime(zs[2])
    for layer in (2, 1, 0):
        g[f"b{layer}"] = dz.sum(axis=0)
        g[f"W{layer}"] = dz.T @ hs[layer]
        if layer:
            dz = (dz @ w[f"W{layer}"]) * silu_prime(zs[layer - 1])
    return float(np.mean((y - t) ** 2)), pack(g)


def forward_bf16(p, x):
    w = unpack(np.asarray(p, dtype=np.float32))
    a = bf16(padded(np.asarray(x, dtype=np.float32)))
    zs, hs = [], [a]
    for W, b in (("W0", "b0"), ("W1", "b1"), ("W2", "b2")):
        z = (hs[-1].astype(np.float32) @ bf16(w[W]).T).astype(np.float32) + w[b]
        zs.append(z.astype(np.float32))
        hs.append(bf16(silu(z).astype(np.float32)))
    y = (hs[-1] @ bf16(w["w3"])).astype(np.float32) + w["b3"][0]
    return y.astype(np.float32), zs, hs


def gradient_bf16(p, x, t):
    w = unpack(np.asarray(p, dtype=np.float32))
    y, zs, hs = forward_bf16(p, x)
    t = np.asarray(t, dtype=np.float32)
    n = len(y)
    dy = (2.0 * (y - t) / n).astype(np.float32)
    g = {"b3": np.array([dy.sum()]), "w3": dy @ hs[3]}
    dz = (dy[:, None] * bf16(w["w3"])[None, :] * silu_prime(zs[2])).astype(np.float32)
    for layer in (2, 1, 0):
        g[f"b{layer}"] = dz.sum(axis=0)
        dzb = bf16(dz)
        g[f"W{layer}"] = dzb.T @ hs[layer]
        if layer:
            dz = ((dzb @ bf16(w[f"W{layer}"])) * silu_prime(zs[layer - 1])).astype(np.float32)
    return float(np.mean((y - t) ** 2)), pack(g, np.float32)
