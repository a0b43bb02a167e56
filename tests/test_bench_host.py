"""Host-side pieces of bench.py that need no GPU: the one-GPU refusal, the counter summary it
reads the traffic figure from, the warm-start ln f rule and the C2 run configurations."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_traffic_figures_come_from_the_committed_counter_summary():
    for kernel in ("lookahead_fast_kernel<8>", "lookahead_xo_kernel<8>"):
        value = bench.traffic_bytes(kernel)
        assert value is not None and 1e6 < value < 1e9
    assert bench.traffic_bytes("no_such_kernel<1>") is None


@pytest.mark.parametrize("rms,n,expected", [(1e-6, 512, 1e-4), (1e-4, 512, (3e-4 * 512) ** 2),
                                            (1e-3, 512, 1.0)])
def test_warm_start_factor_is_the_squared_error_bound_clamped(rms, n, expected):
    assert bench.warm_initial_factor(rms, n) == pytest.approx(expected, rel=1e-12)


def test_c2_configurations():
    plain = bench.c2_config(rng=1)
    many = bench.c2_config(rng=1, seed=77, windows=48)
    assert (plain.windows, plain.walkers_per_window, plain.seed) == (8, 16, 1234)
    assert (many.windows, many.walkers_per_window, many.seed) == (48, 16, 77)
    for c in (plain, many):
        assert (c.emin, c.emax, c.bin_width) == bench.C2_GRID
        assert c.exchange_interval_attempts == 512 and c.overlap == 0.75
        assert round((c.emax - c.emin) / c.bin_width) == 1255


def test_window_rms_removes_each_window_mean():
    """A linear error d = k e, sampled densely, leaves k w / sqrt(12) inside windows of width w
    once each window's mean is removed; a constant error leaves nothing."""
    import numpy as np

    e = np.linspace(0.0, 4.0, 400001)[:-1]
    k = 2e-3
    for w in (1.0, 0.5, 0.25):
        rms = bench.window_rms([(e, k * e)], w)
        assert rms == pytest.approx(k * w / np.sqrt(12.0), rel=1e-3)
    assert bench.window_rms([(e, np.full_like(e, 0.7))], 0.5) == pytest.approx(0.0, abs=1e-12)
