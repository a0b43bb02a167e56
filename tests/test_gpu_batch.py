"""Batches of uneven systems in one launch (wlx_batch): every engine of a batch ends up
bit-identical to the same engine run alone, and to the oracle; HIP-graph replays equal plain
launches on every replay; engines that finish early stop on the device while the others go on.
"""
from __future__ import annotations

import numpy as np
import pytest

from harness import assert_walkers_equal
from replica_exchange_wang_landau_amd import Batch, abi
from replica_exchange_wang_landau_amd.engine import (Engine, InvalidArgument, RunConfig, System,
                                                     WlParameters, simple_cubic)
from test_gpu_lds_walker import dilute, grid_for

pytestmark = pytest.mark.gpu

SIZES = [(64, 1), (300, 2), (1000, 3), (2500, 4), (4096, 5)]


def c3_engines(backend, rng_mode=abi.RNG_PHILOX, max_epochs=0, sizes=SIZES, grids=None,
               walkers=2):
    """C3-shaped engines: dilute CSR systems (cutoff 3), windows in proportion to N, bin 1.0,
    exchange every MCS. `grids` reuses energy ranges (they come from a GPU pilot)."""
    engines, out_grids = [], []
    for i, (n, seed) in enumerate(sizes):
        system = System(*dilute(n, seed), backend=backend, cutoff=3.0)
        emin, emax = grids[i] if grids else grid_for(system, n)
        out_grids.append((emin, emax))
        windows = min(16, max(4, round(4 + 12 * (n - 64) / 4032)))
        cfg = RunConfig(emin=float(emin), emax=float(emax), bin_width=1.0, windows=windows,
                        walkers_per_window=walkers, seed=1234 + seed,
                        exchange_interval_attempts=n,
                        max_attempts=n * max_epochs if max_epochs else 10**12, rng=rng_mode,
                        wl=WlParameters(check_interval_attempts=5 * n))
        e = Engine(system, cfg)
        e.initialize()
        if backend.__class__.__name__ == "GpuBackend":
            e.set_kernel(abi.KERNEL_LDS)
        engines.append(e)
    return engines, out_grids


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_batch_equals_engines_alone_and_oracle(gpu, orc, rng_mode):
    batched, grids = c3_engines(gpu, rng_mode)
    alone, _ = c3_engines(gpu, rng_mode, grids=grids)
    batch = Batch(batched)
    epochs, done = batch.run(12)
    assert (epochs, done) == (12, 0)
    for e in alone:
        e.run(12)
    for a, b in zip(batched, alone):
        sa, sb = a.status(), b.status()
        assert (sa.epoch, sa.attempted, sa.accepted, sa.exchange_accepted) == (
            sb.epoch, sb.attempted, sb.accepted, sb.exchange_accepted)
        for k in range(a.walkers):
            assert_walkers_equal(a, b, k)
    oracle, _ = c3_engines(orc, rng_mode, sizes=SIZES[:2], grids=grids[:2])
    for a, o in zip(batched, oracle):
        o.run(12)
        for k in range(a.walkers):
            assert_walkers_equal(a, o, k)


def test_graph_replays_equal_plain_launches(gpu):
    graphed, grids = c3_engines(gpu)
    plain, _ = c3_engines(gpu, grids=grids)
    bg, bp = Batch(graphed), Batch(plain)
    bg.set_graph(4)
    for _ in range(3):  # three replays, host reads in between
        assert bg.run(4) == (4, 0)
        assert bp.run(4) == (4, 0)
        for a, b in zip(graphed, plain):
            for k in range(a.walkers):
                assert_walkers_equal(a, b, k)
    assert graphed[0].status().epoch == 12


def test_finished_engines_stop_on_the_device(gpu):
    """Engines with attempt limits of 3, 6 and 9 MCS: the batch runs until all are done and
    each stops at its own limit, like run() of the engine alone."""
    sizes = SIZES[:3]
    batched, grids = [], []
    for limit, size in zip((3, 6, 9), sizes):
        e, g = c3_engines(gpu, max_epochs=limit, sizes=[size])
        batched += e
        grids += g
    alone = []
    for limit, size, g in zip((3, 6, 9), sizes, grids):
        alone += c3_engines(gpu, max_epochs=limit, sizes=[size], grids=[g])[0]
    batch = Batch(batched)
    batch.set_graph(2)
    epochs, done = batch.run(100)
    assert done == 3 and epochs <= 12
    for a, b in zip(batched, alone):
        sb = b.run()
        sa = a.status()
        assert sa.done and sb.done and sa.epoch == sb.epoch and sa.attempted == sb.attempted
        for k in range(a.walkers):
            assert_walkers_equal(a, b, k)


def test_batch_errors(gpu):
    a, grids = c3_engines(gpu, abi.RNG_PHILOX, sizes=SIZES[:1])
    b, _ = c3_engines(gpu, abi.RNG_XOSHIRO, sizes=SIZES[:1], grids=grids)
    with pytest.raises(InvalidArgument, match="share one proposal stream"):
        Batch(a + b)
    with pytest.raises(InvalidArgument, match="listed twice"):
        Batch(a + a)
    f32 = Engine(System(*simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True), backend=gpu),
                 RunConfig(emin=-60.0, emax=80.0, bin_width=0.5, rng=abi.RNG_PHILOX,
                           fields=abi.FIELDS_FP32_KAHAN))
    f32.initialize()
    with pytest.raises(InvalidArgument, match="FP64 fields"):
        Batch(a + [f32])


def test_batched_pilot_equals_single_pilots(gpu):
    """wlx_systems_pilot (one launch, LDS fields, CSR or dense rows) against the per-system
    wl_run --pilot sweep, whose output is pinned to the reference's wl_run."""
    systems = [System(*dilute(n, seed), backend=gpu, cutoff=3.0) for n, seed in SIZES]
    systems.append(System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu))
    systems.append(System(*simple_cubic(5, 7, 3, 1.0, (0.6, 0.0, 0.8), True), backend=gpu))
    attempts = [20 * s.n for s in systems]
    lo, hi = System.pilot_many(systems, 1234, attempts)
    for k, s in enumerate(systems):
        a, b = s.pilot(1234, attempts[k])
        assert (lo[k].hex(), hi[k].hex()) == (a.hex(), b.hex()), k


def replicas(backend, lx, grid, seeds, walkers=4, max_epochs=0, rng=abi.RNG_PHILOX):
    """Dense simple-cubic replicas that differ only in their seed (default kernel)."""
    out = []
    for seed in seeds:
        system = System(*simple_cubic(lx, lx, lx, 1.0, (0, 0, 1), True), backend=backend)
        cfg = RunConfig(emin=grid[0], emax=grid[1], bin_width=grid[2], windows=4,
                        walkers_per_window=walkers, overlap=0.75, seed=seed,
                        exchange_interval_attempts=system.n,
                        max_attempts=system.n * max_epochs if max_epochs else 0,
                        rng=rng, wl=WlParameters(check_interval_attempts=8 * system.n))
        e = Engine(system, cfg)
        e.initialize()
        out.append(e)
    return out


GRID_8 = (-273.0, 354.5, 0.5)    # 8x8x8 dipolar cube, D = 1


def grid_4(gpu):
    """A 4x4x4 grid from a pilot sweep, padded 20 %, bins of 0.5."""
    system = System(*simple_cubic(4, 4, 4, 1.0, (0, 0, 1), True), backend=gpu)
    emin, emax = grid_for(system, system.n)
    return float(emin), float(emax), 0.5


@pytest.mark.parametrize("rng,kernel", [(abi.RNG_PHILOX, "lookahead_fast_kernel"),
                                        (abi.RNG_XOSHIRO, "lookahead_xo_kernel")])
@pytest.mark.parametrize("lx,npl", [(4, 1), (8, 8)])
def test_replica_batch_runs_the_lookahead_walker_bit_exact(gpu, lx, npl, rng, kernel):
    """Replicas of one dense system on the default kernel of either stream batch onto that
    lookahead kernel (with the batch's slot list) and end up identical to each replica run
    alone, over plain launches and graph replays."""
    seeds = [11, 12, 13]
    grid = GRID_8 if lx == 8 else grid_4(gpu)
    batched = replicas(gpu, lx, grid, seeds, rng=rng)
    alone = replicas(gpu, lx, grid, seeds, rng=rng)
    # alone, an N = 512 Philox engine runs the staged walker, which batches do not use
    staged = rng == abi.RNG_PHILOX and npl == 8
    assert alone[0].kernel()[1] == (f"lookahead_staged_kernel<{npl}>" if staged
                                    else f"{kernel}<{npl}>")
    batch = Batch(batched)
    assert batch.kernel() == f"{kernel}<{npl}>"
    assert batch.run(5) == (5, 0)
    batch.set_graph(3)
    assert batch.run(6) == (6, 0)
    for e in alone:
        e.run(11)
    for a, b in zip(batched, alone):
        sa, sb = a.status(), b.status()
        assert (sa.epoch, sa.attempted, sa.accepted, sa.exchange_accepted) == (
            sb.epoch, sb.attempted, sb.accepted, sb.exchange_accepted)
        assert sa.accepted > 0
        for k in range(a.walkers):
            assert_walkers_equal(a, b, k)


def test_mixed_layouts_batch_onto_the_lds_walker(gpu):
    """Replicas of two sizes (register layouts 1 and 8) fall back to the LDS walker, whose
    trajectories are the same."""
    g4 = grid_4(gpu)
    small, big = replicas(gpu, 4, g4, [3]), replicas(gpu, 8, GRID_8, [4])
    small_alone, big_alone = replicas(gpu, 4, g4, [3]), replicas(gpu, 8, GRID_8, [4])
    batch = Batch(small + big)
    assert batch.kernel() == "lds_sample_kernel<true>"
    assert batch.run(4) == (4, 0)
    for a, b in zip(small + big, small_alone + big_alone):
        b.run(4)
        for k in range(a.walkers):
            assert_walkers_equal(a, b, k)


def test_smaller_engine_keeps_a_larger_engines_lds_limit(gpu):
    """Kernels are shared by every engine: creating an engine that needs less LDS (here one
    window of 5120 bins, ~92 KB) after one that needs more (8192 bins, ~148 KB) must leave
    the larger engine runnable, alone and in a batch."""
    def engine(width, seed):
        system = System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu)
        e = Engine(system, RunConfig(emin=-280.0, emax=360.0, bin_width=width, windows=1,
                                     walkers_per_window=2, seed=seed,
                                     exchange_interval_attempts=512, rng=abi.RNG_PHILOX))
        e.initialize()
        return e

    big = engine(0.078125, 1)
    assert big.bins == 8192
    small = engine(0.125, 2)
    assert small.bins == 5120
    small.run(2)
    assert big.run(2).epoch == 2
    batch = Batch([big, engine(0.078125, 3)])
    engine(0.125, 4)
    assert batch.run(2) == (2, 0)


def test_sampled_launch_timing(gpu):
    """wlx_engine_set_timing(k): events around every k-th sampling launch; the mean of the
    sampled launches is close to the mean of all of them, and 0 switches timing off."""
    e = replicas(gpu, 8, GRID_8, [5])[0]
    e.run(10)
    e.set_timing(True)
    e.run(40)
    every_ms, every_n = e.kernel_time()
    e.set_timing(True, every=4)
    e.run(40)
    sampled_ms, sampled_n = e.kernel_time()
    assert (every_n, sampled_n) == (40, 10)
    assert every_ms > 0.0 and abs(sampled_ms - every_ms) < 0.25 * every_ms
    e.set_timing(False)
    e.run(5)
    assert e.kernel_time() == (0.0, 0)
    with pytest.raises(InvalidArgument):
        e.set_timing(True, every=0)


@pytest.mark.parametrize("rng", [abi.RNG_PHILOX, abi.RNG_XOSHIRO])
def test_engine_graph_replays_equal_plain_launches(gpu, rng):
    """wlx_engine_set_graph(k): run() replays k captured epochs per graph launch. Replays with
    host reads in between, a run length that is not a multiple of k, a kernel change (the
    graph is recaptured) and a run that finishes inside a graph all end bit-identical to plain
    launches of the same engine."""
    graphed = replicas(gpu, 8, GRID_8, [21], rng=rng, max_epochs=40)[0]
    plain = replicas(gpu, 8, GRID_8, [21], rng=rng, max_epochs=40)[0]
    graphed.set_graph(4)
    for epochs in (8, 5, 4):
        sg, sp = graphed.run(epochs), plain.run(epochs)
        assert (sg.epoch, sg.attempted, sg.accepted, sg.exchange_accepted) == (
            sp.epoch, sp.attempted, sp.accepted, sp.exchange_accepted)
        for k in range(graphed.walkers):
            assert_walkers_equal(graphed, plain, k)
    for e in (graphed, plain):
        e.set_kernel(abi.KERNEL_SEQUENTIAL)
    assert graphed.run(6).epoch == plain.run(6).epoch == 23
    for e in (graphed, plain):
        e.set_kernel(abi.KERNEL_AUTO)
    sg, sp = graphed.run(100), plain.run(100)  # the attempt limit falls at epoch 40
    assert sg.done and sp.done and sg.epoch == sp.epoch == 40
    assert sg.attempted == sp.attempted
    for k in range(graphed.walkers):
        assert_walkers_equal(graphed, plain, k)
    with pytest.raises(InvalidArgument):
        graphed.set_graph(5000)
