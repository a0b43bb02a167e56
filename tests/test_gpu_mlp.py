"""The warm-start network's kernels (csrc/mlp.hip: BF16 matrix-core products, FP32
accumulate) against the numpy restatement in oracle/mlp_oracle.py, and the warm start of a
WL run.

Tolerances: the kernels sum in another order than numpy, and a value that lands near a BF16
rounding boundary can round the other way, so outputs and gradients are compared by relative
norms (1e-2) rather than bit for bit.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

from harness import GOLDEN, c1_config, c1_geometry
from replica_exchange_wang_landau_amd import abi
from replica_exchange_wang_landau_amd.engine import Engine, InvalidArgument, System
from replica_exchange_wang_landau_amd.warmstart import Mlp, warm_start

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, mo.FEATURES)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2]).astype(np.float32)
    return x, t


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_parameters_and_initialisation(gpu):
    a, b, c = Mlp(7, gpu), Mlp(7, gpu), Mlp(8, gpu)
    assert a.param_count == mo.PARAMS
    pa = a.params()
    assert np.array_equal(pa, b.params()) and not np.array_equal(pa, c.params())
    w = mo.unpack(pa)
    assert not w["W0"][:, mo.FEATURES:].any() and not w["b1"].any()
    bound = np.sqrt(6.0 / mo.WIDTH)
    assert np.abs(w["W1"]).max() <= bound and np.abs(w["W1"]).std() > 0.3 * bound
    q = pa * 0.5
    a.set_params(q)
    assert np.array_equal(a.params(), q)
    with pytest.raises(InvalidArgument):
        a.set_params(np.full(mo.PARAMS, np.nan, dtype=np.float32))


def test_forward_and_gradient_match_restatement(gpu):
    mlp = Mlp(3, gpu)
    p = mlp.params()
    for n in (1, 16, 37, 500):  # partial tiles included
        x, t = data(n, n)
        y = mlp.predict(x)
        want, _, _ = mo.forward_bf16(p, x)
        assert rel(y, want) < 1e-2, n
        loss, g = mlp.gradient(x, t)
        l_ref, g_ref = mo.gradient_bf16(p, x, t)
        assert abs(loss - l_ref) <= 1e-2 * l_ref, n
        offsets = np.cumsum([0] + [int(np.prod(s)) for _, s in mo.SHAPES])
        for (name, _), a, b in zip(mo.SHAPES, offsets[:-1], offsets[1:]):
            assert rel(g[a:b], g_ref[a:b]) < 2e-2, (n, name)
        # and the FP64 mathematics, to BF16 accuracy
        _, g64 = mo.gradient64(p.astype(np.float64), x, t)
        assert rel(g, g64) < 5e-2, n


def test_training_reduces_the_loss(gpu):
    mlp = Mlp(11, gpu)
    x, t = data(5, 4096)
    before, _ = mlp.gradient(x, t)
    losses = mlp.train(x, t, steps=600, batch=256, learning_rate=2e-3, seed=4)
    after, _ = mlp.gradient(x, t)
    assert losses[0] == pytest.approx(losses[0]) and after < 0.1 * before, (before, after)
    # deterministic: the same seeds give the same weights
    again = Mlp(11, gpu)
    again.train(x, t, steps=600, batch=256, learning_rate=2e-3, seed=4)
    assert np.array_equal(again.params(), mlp.params())


def test_warm_start_from_exact_dos_meets_the_same_bound(gpu):
    """C1 plain WL warm-started from the exact log g at ln f = 1e-4 converges to ln f = 1e-6
    within the flat start's bound (0.03 on max |dlog g|), in fewer attempts."""
    from replica_exchange_wang_landau_amd import analysis
    exact = json.loads((GOLDEN / "exact_c1.json").read_text())["counts"]
    occupied = sorted(int(b) for b in exact)
    results = {}
    for mode in ("flat", "warm"):
        e = Engine(System(*c1_geometry(), backend=gpu), c1_config(rng=abi.RNG_PHILOX))
        e.initialize()
        if mode == "warm":
            log_g = np.zeros(e.bins)
            for b in occupied:
                log_g[b] = np.log(exact[str(b)])
            warm_start(e, log_g, 1e-4)
            assert e.read_walker(0).factor == 1e-4
        st = e.run()
        assert st.converged
        f = analysis.summarize_window([e.read_walker(0), e.read_walker(1)], e.bins)
        d = np.array([f.log_g[b] - np.log(exact[str(b)]) for b in occupied])
        results[mode] = (st.attempted, float(np.max(np.abs(d - d.mean()))))
    assert results["warm"][1] < 0.03 and results["flat"][1] < 0.03
    assert results["warm"][0] < results["flat"][0], results
    with pytest.raises(InvalidArgument):
        warm_start(e, np.zeros(e.bins), 0.0)
