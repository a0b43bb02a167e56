"""The LDS walker kernel (WLX_KERNEL_LDS: fields and spins in LDS, CSR row scatter or dense row
stream) and the kernel autotuner, against the oracle.

The oracle applies a flip with the reference's CsrCouplings::add_flip_delta
(physics.cpp:231-236). Bar: bit-exact walkers (spins, fields, energies, log_g, histograms,
counters, RNG words) on both proposal streams, including the per-attempt traces.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from harness import assert_walkers_equal
from replica_exchange_wang_landau_amd import abi
from replica_exchange_wang_landau_amd.engine import (Engine, InvalidArgument, RunConfig, System,
                                                     WlParameters, simple_cubic)

pytestmark = pytest.mark.gpu


def dilute(n, seed):
    """C3-style system: n sites uniform in a periodic cube at number density 0.3, axes
    uniform on the sphere (seed-determined)."""
    rng = np.random.default_rng(seed)
    L = (n / 0.3) ** (1 / 3)
    pos = rng.uniform(0.0, L, size=(n, 3))
    axes = rng.normal(size=(n, 3))
    axes /= np.sqrt((axes * axes).sum(axis=1))[:, None]
    return pos, axes, np.array([L, L, L]), True


def grid_for(system, n):
    """Energy range from a short pilot sweep, padded by 20 %, on bins of width 1.0."""
    lo, hi = system.pilot(1234, 20 * n)
    pad = 0.2 * (hi - lo) + 2.0
    return math.floor(lo - pad), math.ceil(hi + pad)


def pair(gpu, orc, geometry, cfg_kw, cutoff=3.0):
    sg = System(*geometry, backend=gpu, cutoff=cutoff)
    so = System(*geometry, backend=orc, cutoff=cutoff)
    emin, emax = grid_for(sg, len(geometry[0]))
    cfg = RunConfig(emin=float(emin), emax=float(emax), bin_width=1.0, **cfg_kw)
    eg, eo = Engine(sg, cfg), Engine(so, cfg)
    eg.initialize()
    eo.initialize()
    eg.set_kernel(abi.KERNEL_LDS)
    return eg, eo


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
@pytest.mark.parametrize("n,seed", [(64, 1), (333, 2), (1000, 3), (4096, 4)])
def test_csr_kernel_bit_exact(gpu, orc, rng_mode, n, seed):
    windows = min(16, max(4, round(4 + 12 * (n - 64) / 4032)))  # C3's windows per system
    eg, eo = pair(gpu, orc, dilute(n, seed),
                  dict(windows=windows, walkers_per_window=2, seed=17,
                       exchange_interval_attempts=n, max_attempts=n * 10, rng=rng_mode,
                       wl=WlParameters(check_interval_attempts=4 * n)))
    family, name = eg.kernel()
    assert family == abi.KERNEL_LDS and name.startswith("lds_sample_kernel")
    sg, so = eg.run(), eo.run()
    assert (sg.epoch, sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.epoch, so.attempted, so.accepted, so.exchange_accepted)
    assert sg.accepted > 0
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def test_csr_kernel_trace_forced_nalivaiko_bit_exact(gpu, orc):
    """Per-attempt traces with forced acceptance and the Nalivaiko mask on."""
    n = 200
    eg, eo = pair(gpu, orc, dilute(n, 9),
                  dict(windows=2, walkers_per_window=2, seed=3, exchange_interval_attempts=n,
                       wl=WlParameters(check_interval_attempts=n, force_accept_after_attempts=5,
                                       nalivaiko_mod=True, inverse_time_enabled=False,
                                       flatness=0.5, minimum_visits=2)))
    for e in (eg, eo):
        e.trace(3000)
        e.sample(3000)
        e.check()
    for k in range(eg.walkers):
        a, b = eg.read_trace(k), eo.read_trace(k)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), k
        assert_walkers_equal(eg, eo, k)
    assert sum(eg.read_walker(k).forced_accepted for k in range(eg.walkers)) > 0


def test_autotune_keeps_the_trajectory(gpu, orc):
    """Autotuning runs trial epochs with each kernel family inside the live run; the run stays
    bit-identical to the oracle's uninterrupted one."""
    n = 512
    eg, eo = pair(gpu, orc, dilute(n, 5),
                  dict(windows=4, walkers_per_window=4, seed=23, exchange_interval_attempts=n,
                       wl=WlParameters(check_interval_attempts=10 * n)))
    chosen, seconds = eg.autotune(3)
    assert chosen in (abi.KERNEL_LOOKAHEAD, abi.KERNEL_SEQUENTIAL, abi.KERNEL_LDS)
    assert all(seconds[k] > 0 for k in (1, 2, 3))
    done = eg.status().epoch
    assert done == 3 * 4  # three families, one warm-up epoch and three timed epochs each
    sg = eg.run(20)
    so = eo.run(done + 20)
    assert (sg.epoch, sg.attempted, sg.accepted) == (so.epoch, so.attempted, so.accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def test_kernel_selection_errors(gpu):
    e = Engine(System(*simple_cubic(2, 2, 2, 1.0, (0, 0, 1), False), backend=gpu),
               RunConfig(emin=-20.0, emax=20.0, bin_width=0.25, seed=1))
    with pytest.raises(InvalidArgument, match="Unknown walker kernel"):
        e.set_kernel(7)
    e.initialize()
    e.set_kernel(abi.KERNEL_SEQUENTIAL)
    assert e.kernel() == (abi.KERNEL_SEQUENTIAL, "sample_kernel<1,false>")


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
@pytest.mark.parametrize("geometry", [(8, 8, 8, 1.0, (0, 0, 1), True),
                                      (5, 7, 3, 1.0, (0.6, 0.0, 0.8), True)])
def test_lds_kernel_dense_rows_bit_exact(gpu, orc, rng_mode, geometry):
    """The LDS walker on dense systems streams whole rows (DenseCouplings::add_flip_delta)."""
    from harness import c2_config
    n = geometry[0] * geometry[1] * geometry[2]
    cfg = c2_config(max_attempts=n * 20, rng=rng_mode, exchange_interval_attempts=n)
    if n != 512:
        cfg.emin, cfg.emax, cfg.bin_width = -120.0, 120.0, 0.3
    eg = Engine(System(*simple_cubic(*geometry), backend=gpu), cfg)
    eo = Engine(System(*simple_cubic(*geometry), backend=orc), cfg)
    eg.initialize()
    eo.initialize()
    eg.set_kernel(abi.KERNEL_LDS)
    assert eg.kernel()[1].startswith("lds_sample_kernel")
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
