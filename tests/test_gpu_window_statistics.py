"""Pilot statistics (collect_window_statistics) on the device against the oracle.

The statistics are recorded by the one-flip-at-a-time kernels on either proposal stream:
per-bin squared accepted displacements and attempt counts (wang_landau.cpp:272-276), round
trips between the window's margins (wang_landau.cpp:317-330) and the nine representative
configurations (wang_landau.cpp:107-118, 332-342). Bar: bit-exact, with the walker states.
The oracle's statistics are pinned against the compiled reference walker in
test_oracle_pinning.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from harness import assert_walkers_equal, c2_config
from replica_exchange_wang_landau_amd import abi
from replica_exchange_wang_landau_amd.engine import (Engine, InvalidArgument, LogicError, RunConfig,
                                                     System, WlParameters, simple_cubic)
from test_oracle_pinning import assert_sampling_equal

pytestmark = pytest.mark.gpu


def pair(gpu, orc, geometry, cfg, initial=None, supplied=None):
    eg = Engine(System(*geometry, backend=gpu), cfg)
    eo = Engine(System(*geometry, backend=orc), cfg)
    for e in (eg, eo):
        e.initialize(initial, supplied)
    return eg, eo


def assert_engines_equal(eg, eo):
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
        assert_sampling_equal(eg.read_sampling(k), eo.read_sampling(k))


def narrow_config(rng_mode, margin=0.15):
    """3x3x3 periodic, 2 windows x 2 walkers on [-60, 80): windows narrow enough for many
    round trips per walker."""
    return RunConfig(emin=-60.0, emax=80.0, bin_width=0.5, windows=2, walkers_per_window=2,
                     seed=5, exchange_interval_attempts=54, max_attempts=0, rng=rng_mode,
                     wl=WlParameters(final_factor=1e-3, check_interval_attempts=540,
                                     collect_window_statistics=True,
                                     round_trip_margin_fraction=margin))


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_narrow_windows_statistics_bit_exact(gpu, orc, rng_mode):
    geo = simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True)
    eg, eo = pair(gpu, orc, geo, narrow_config(rng_mode))
    assert_engines_equal(eg, eo)  # after initialisation
    for epochs in (1, 7, 120):
        sg, so = eg.run(epochs), eo.run(epochs)
        assert (sg.epoch, sg.attempted, sg.accepted, sg.exchange_accepted) == (
            so.epoch, so.attempted, so.accepted, so.exchange_accepted)
        assert_engines_equal(eg, eo)
    stats = [eg.read_sampling(k) for k in range(eg.walkers)]
    assert sum(s["round_trips"] for s in stats) > 0
    assert all(sum(r is not None for r in s["representatives"]) >= 5 for s in stats)
    # every attempt lands in exactly one bin of its walker's window
    for k, s in enumerate(stats):
        w = eg.read_walker(k)
        assert int(s["displacement_samples"].sum()) == int(w.attempted)
        outside = np.ones(eg.bins, dtype=bool)
        outside[w.window_begin:w.window_end] = False
        assert not s["displacement_samples"][outside].any()
    # a representative's spins reproduce its energy
    for s in stats:
        for r in s["representatives"]:
            if r is not None:
                _, energy = eg.system.fields(r[1][None, :].astype(np.int8))
                assert abs(energy[0] - r[0]) <= 1e-10 * max(1.0, abs(r[0]))


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_c2_statistics_bit_exact(gpu, orc, rng_mode):
    """C2 at full size (128 walkers) for 30 exchange epochs with statistics collected."""
    cfg = c2_config(max_attempts=512 * 30, rng=rng_mode)
    cfg.wl.collect_window_statistics = True
    cfg.wl.round_trip_margin_fraction = 0.1
    eg, eo = pair(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    assert_engines_equal(eg, eo)


def test_supplied_starts_statistics_bit_exact(gpu, orc):
    """Supplied starting configurations (the adaptive planner's representatives take this
    path): statistics from the placement onwards."""
    geo = simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True)
    cfg = narrow_config(abi.RNG_XOSHIRO)
    rng = np.random.default_rng(3)
    initial = rng.choice(np.array([-1, 1], dtype=np.int8), (4, 27))
    supplied = np.array([1, 0, 1, 1], dtype=np.uint8)
    eg, eo = pair(gpu, orc, geo, cfg, initial, supplied)
    assert_engines_equal(eg, eo)
    eg.run(30)
    eo.run(30)
    assert_engines_equal(eg, eo)


def test_statistics_errors(gpu):
    geo = simple_cubic(2, 2, 2, 1.0, (0, 0, 1), False)
    cfg = RunConfig(emin=-20.0, emax=20.0, bin_width=0.25, seed=1)
    e = Engine(System(*geo, backend=gpu), cfg)
    e.initialize()
    with pytest.raises(LogicError, match="not collected"):
        e.read_sampling(0)
    bad = RunConfig(emin=-20.0, emax=20.0, bin_width=0.25, seed=1, rng=abi.RNG_PHILOX,
                    fields=abi.FIELDS_FP32_KAHAN,
                    wl=WlParameters(collect_window_statistics=True))
    with pytest.raises(InvalidArgument, match="FP64 fields only"):
        Engine(System(*geo, backend=gpu), bad)
