"""The warm-start network's restatement (oracle/mlp_oracle.py), pinned on the CPU: central
finite differences of the FP64 loss against its analytic gradient, closed forms, the BF16
rounding rule, and the BF16 restatement against the FP64 one."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402


def sample(seed=3, n=40):
    rng = np.random.default_rng(seed)
    p = np.zeros(mo.PARAMS)
    parts = mo.unpack(p)
    for name, shape in mo.SHAPES:
        fan = shape[1] if len(shape) == 2 else mo.WIDTH
        parts[name][...] = rng.uniform(-1, 1, size=shape) * np.sqrt(3.0 / fan)
    parts["W0"][:, mo.FEATURES:] = 0.0
    x = rng.normal(size=(n, mo.FEATURES))
    t = rng.normal(size=n)
    return mo.pack(parts), x, t


def test_gradient_matches_central_differences():
    p, x, t = sample()
    _, g = mo.gradient64(p, x, t)
    rng = np.random.default_rng(7)
    offsets = np.cumsum([0] + [int(np.prod(s)) for _, s in mo.SHAPES])
    picks = []
    for a, b in zip(offsets[:-1], offsets[1:]):  # every block, including the bias of layer 3
        picks += list(rng.integers(a, b, size=min(12, b - a)))
    for q in picks:
        if q < offsets[1] and (q % mo.IN) >= mo.FEATURES:
            assert g[q] == 0.0  # padded inputs carry no gradient
            continue
        h = 1e-6 * max(1.0, abs(p[q]))
        up, dn = p.copy(), p.copy()
        up[q] += h
        dn[q] -= h
        fd = (mo.loss64(up, x, t) - mo.loss64(dn, x, t)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1e-3, abs(g[q])) + 1e-9, q


def test_closed_forms():
    # all weights zero: every hidden unit is SiLU(b), y = b3, and dL/db3 = 2 mean(b3 - t)
    p = np.zeros(mo.PARAMS)
    parts = mo.unpack(p)
    parts["b3"][0] = 0.75
    parts["b2"][:] = 1.0
    x = np.random.default_rng(1).normal(size=(5, mo.FEATURES))
    t = np.arange(5.0)
    y, _, hs = mo.forward64(p, x)
    assert np.allclose(y, 0.75)
    assert np.allclose(hs[3], 1.0 / (1.0 + np.exp(-1.0)))
    loss, g = mo.gradient64(p, x, t)
    gp = mo.unpack(g)
    assert np.isclose(loss, np.mean((0.75 - t) ** 2))
    assert np.isclose(gp["b3"][0], 2.0 * np.mean(0.75 - t))
    assert np.allclose(gp["w3"], 2.0 * np.mean(0.75 - t) * hs[3][0])
    assert not gp["W1"].any() and not gp["W0"].any()  # w3 = 0 blocks the deeper layers


def test_bf16_rounding_is_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 spacing at 1
    cases = {1.0 + 2.0 ** -8: 1.0,                 # halfway, even neighbour below
             1.0 + 3 * 2.0 ** -8: 1.0 + 2 * ulp,   # halfway, even neighbour above
             1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + ulp,
             -3.0: -3.0, 0.0: 0.0}
    for v, want in cases.items():
        assert mo.bf16(np.array([v], dtype=np.float32))[0] == np.float32(want), v
    assert one + ulp != one


def test_bf16_restatement_tracks_fp64():
    p, x, t = sample(5, 64)
    l64, g64 = mo.gradient64(p, x, t)
    l16, g16 = mo.gradient_bf16(p, x, t)
    assert abs(l16 - l64) <= 2e-2 * l64
    assert np.linalg.norm(g16 - g64) <= 3e-2 * np.linalg.norm(g64)
