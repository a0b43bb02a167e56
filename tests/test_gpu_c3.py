"""The batched configuration as configured: 64 uneven dilute systems (N uniform in
{64..4096}, uniform random placement at density 0.3, axes uniform on the sphere) with FP32
CSR couplings (cutoff 3), 4-16 windows of 8 walkers each, sampled together in one launch per
epoch. Windows too wide for LDS keep their bins in HBM.

Checked: FP32 coupling builds equal the oracle's (FP64 build rounded to binary32) bit for
bit; the LDS walker on FP32 CSR values and on HBM-resident bins equals the oracle walker;
every engine of the 64-system batch equals the same engine run alone; the smallest and the
largest system equal the oracle run on the same FP32 couplings.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from harness import assert_walkers_equal
from replica_exchange_wang_landau_amd import Batch, abi, workloads
from replica_exchange_wang_landau_amd.engine import (Engine, InvalidArgument, RunConfig, System,
                                                     WlParameters)

pytestmark = pytest.mark.gpu

PILOT_MCS = 200  # the range only has to be plausible here; the bench pilots 10^3 MCS


@pytest.mark.parametrize("n,seed", [(97, 3), (1500, 8)])
def test_fp32_couplings_equal_the_oracle(gpu, orc, n, seed):
    """wlx_system_create(precision = FP32): CSR values and the dense expansion hold each FP64
    coupling rounded to binary32, equal to the oracle's; rows keep their columns."""
    geo = workloads.geometry(n, seed)
    g = System(*geo, cutoff=3.0, backend=gpu, precision=abi.COUPLINGS_FP32)
    o = System(*geo, cutoff=3.0, backend=orc, precision=abi.COUPLINGS_FP32)
    g64 = System(*geo, cutoff=3.0, backend=gpu)
    (go, gc, gv), (oo, oc, ov), (_, _, v64) = g.csr(), o.csr(), g64.csr()
    assert go.tolist() == oo.tolist() and gc.tolist() == oc.tolist()
    assert gv.tobytes() == ov.tobytes()
    assert gv.tobytes() == v64.astype(np.float32).astype(np.float64).tobytes()
    assert np.any(gv != v64)  # the rounding changed something
    dense = g.dense()
    assert dense.tobytes() == o.dense().tobytes()
    assert np.array_equal(dense, dense.T)


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_lds_walker_fp32_csr_bit_exact(gpu, orc, rng_mode):
    """The LDS walker streaming 4-byte CSR values (plus 4-byte columns) equals the oracle
    walker on the same rounded couplings: the float converts exactly to the rounded double."""
    geo = workloads.geometry(700, 21)
    sg = System(*geo, cutoff=3.0, backend=gpu, precision=abi.COUPLINGS_FP32)
    so = System(*geo, cutoff=3.0, backend=orc, precision=abi.COUPLINGS_FP32)
    lo, hi = sg.pilot(1234, PILOT_MCS * sg.n)
    cfg = workloads.config(sg, lo, hi, seed=5, walkers_per_window=2,
                           max_attempts=sg.n * 10)
    cfg.rng = rng_mode
    eg, eo = Engine(sg, cfg), Engine(so, cfg)
    for e in (eg, eo):
        e.initialize()
    eg.set_kernel(abi.KERNEL_LDS)
    sa, sb = eg.run(), eo.run()
    assert (sa.attempted, sa.accepted, sa.exchange_accepted) == (
        sb.attempted, sb.accepted, sb.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_hbm_bins_bit_exact(gpu, orc, rng_mode):
    """Windows of ~12 000 bins (too wide for LDS) keep their bins in HBM: initialisation,
    sampling, flatness checks, iterations and exchanges work on them in place and equal the
    oracle; register kernels are refused for such an engine."""
    geo = workloads.geometry(400, 9)
    sg = System(*geo, cutoff=3.0, backend=gpu, precision=abi.COUPLINGS_FP32)
    so = System(*geo, cutoff=3.0, backend=orc, precision=abi.COUPLINGS_FP32)
    lo, hi = sg.pilot(1234, PILOT_MCS * sg.n)
    span = hi - lo + 40.0
    width = 2.0 ** -math.floor(math.log2(30000.0 / span))  # a power of two: exact edges
    bins = math.ceil(span / width)
    emin = math.floor(lo) - 20.0
    cfg = RunConfig(emin=emin, emax=emin + bins * width, bin_width=width, windows=2,
                    walkers_per_window=2, overlap=0.75, seed=11,
                    exchange_interval_attempts=400, max_attempts=400 * 40, rng=rng_mode,
                    wl=WlParameters(check_interval_attempts=800, minimum_visits=1,
                                    flatness=0.01))
    eg, eo = Engine(sg, cfg), Engine(so, cfg)
    lds, hbm = eg.layout()
    assert hbm and lds == ((400 + 15) // 16 * 16) * 9
    with pytest.raises(InvalidArgument, match="LDS walker only"):
        eg.set_kernel(abi.KERNEL_LOOKAHEAD)
    for e in (eg, eo):
        e.initialize()
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
    sa, sb = eg.run(), eo.run()
    assert (sa.attempted, sa.accepted, sa.exchange_attempted, sa.exchange_accepted) == (
        sb.attempted, sb.accepted, sb.exchange_attempted, sb.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
        st_g, st_o = eg.statistics(k), eo.statistics(k)
        assert (st_g.active_bins, st_g.covered_bins, st_g.minimum) == (
            st_o.active_bins, st_o.covered_bins, st_o.minimum)


@pytest.fixture(scope="module")
def c3_setup(gpu):
    """The 64 systems, their pilot ranges (one launch for all pilots) and configurations
    (windows over the occupied energy intervals)."""
    systems = workloads.systems(64, 1234, backend=gpu)
    lo, hi = System.pilot_many(systems, 1234, [PILOT_MCS * s.n for s in systems])
    configs = [workloads.config(system, a, b, seed=1234 + s, max_attempts=system.n * 5,
                                energies=workloads.sample_energies(system, 2048, s))
               for s, (system, a, b) in enumerate(zip(systems, lo, hi))]
    return systems, configs


def c3_engines(systems, configs):
    engines = []
    for system, cfg in zip(systems, configs):
        e = Engine(system, cfg)
        e.initialize()
        e.set_kernel(abi.KERNEL_LDS)
        engines.append(e)
    return engines


def test_occupied_windows_avoid_energy_gaps(c3_setup):
    """Systems whose close pairs split the density of states get windows over the occupied
    intervals; the others get the reference's partition."""
    systems, configs = c3_setup
    gapped = [c for c in configs if c.explicit_windows is not None]
    assert 0 < len(gapped) < len(configs)
    for c in gapped:
        assert len(c.explicit_windows) == c.windows
        assert all(0 <= b < e <= c.bins for b, e in c.explicit_windows)


def test_c3_batch_as_configured_equals_engines_alone(c3_setup, orc):
    """64 systems, 5 epochs in one launch each (plus graph replays): every walker of every
    engine equals the same engine run alone; the smallest and largest systems equal the
    oracle on the same FP32 couplings. Some windows live in HBM, some in LDS."""
    systems, configs = c3_setup
    sizes = [s.n for s in systems]
    assert len(systems) == 64 and min(sizes) >= 64 and max(sizes) <= 4096
    batched = c3_engines(systems, configs)
    alone = c3_engines(systems, configs)
    layouts = [e.layout()[1] for e in batched]
    assert any(layouts) and not all(layouts), "expected both LDS and HBM bins"
    assert sum(e.walkers for e in batched) == sum(8 * workloads.windows_for(n) for n in sizes)
    batch = Batch(batched)
    assert batch.kernel() == "lds_sample_kernel<true>"
    assert batch.run(2) == (2, 0)
    batch.set_graph(3)
    epochs, done = batch.run(3)
    assert (epochs, done) == (3, 64)
    for e in alone:
        assert e.run().epoch == 5
    for a, b in zip(batched, alone):
        sa, sb = a.status(), b.status()
        assert (sa.epoch, sa.attempted, sa.accepted, sa.exchange_accepted) == (
            sb.epoch, sb.attempted, sb.accepted, sb.exchange_accepted)
        for k in range(a.walkers):
            assert_walkers_equal(a, b, k)
    for idx in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        system = systems[idx]
        so = System(system.positions, system.axes, system.box, True, cutoff=3.0, backend=orc,
                    precision=abi.COUPLINGS_FP32)
        eo = Engine(so, configs[idx])
        eo.initialize()
        so_status = eo.run()
        assert so_status.attempted == batched[idx].status().attempted
        for k in range(eo.walkers):
            assert_walkers_equal(batched[idx], eo, k)


def test_lds_budget_and_autotune_leave_trajectories_unchanged(c3_setup):
    """Launch groups by LDS class, an LDS budget that moves more windows' bins to HBM, and the
    batch autotuner trying budgets on the live run all give the trajectories of plain
    launches of the engines alone."""
    import dataclasses

    systems, configs = c3_setup
    pick = [int(i) for i in np.argsort([s.n for s in systems])[::5]]  # 13 sizes, small to large
    sub_sys = [systems[i] for i in pick]
    sub_cfg = [dataclasses.replace(configs[i], max_attempts=systems[i].n * 40) for i in pick]
    tight, tuned, alone = (c3_engines(sub_sys, sub_cfg) for _ in range(3))
    bt = Batch(tight)
    default_layout = bt.layout()
    bt.set_lds_budget(32 * 1024)
    layout = bt.layout()
    assert layout["bins_in_hbm"] > default_layout["bins_in_hbm"]
    assert layout["max_lds_bytes"] <= 48 * 1024 and layout["groups"] >= 2
    bu = Batch(tuned)
    bu.set_graph(2)
    chosen, seconds = bu.autotune(2)
    assert chosen in (160 * 1024, 64 * 1024, 32 * 1024) and all(s > 0 for s in seconds)
    ran = tuned[0].status().epoch
    assert ran == 3 * (2 + 2)
    assert bt.run(ran) == (ran, 0)
    for e in alone:
        e.run(ran)
    for x, y, z in zip(tight, tuned, alone):
        for k in range(x.walkers):
            assert_walkers_equal(x, z, k)
            assert_walkers_equal(y, z, k)
