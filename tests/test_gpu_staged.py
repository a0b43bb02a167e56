"""The staged lookahead walker (WLX_KERNEL_STAGED, lookahead_staged_kernel): each batch's 16
candidate rows J[s_q, :] are copied into LDS while the previous batch decides, the window's
log_g carries a +inf entry on either side, and the decisions, first visits, 1/t factors and
freezes are those of the one-flip-at-a-time walker. Every test compares with the oracle (or
with the register lookahead kernel, which the oracle pins) bit for bit: spins, fields,
energies, log_g, histograms, counters, stages."""
from __future__ import annotations

import math

import numpy as np
import pytest

from harness import assert_walkers_equal, c2_config
from replica_exchange_wang_landau_amd import Engine, System, abi
from replica_exchange_wang_landau_amd.engine import (InvalidArgument, RunConfig, WlParameters,
                                                     simple_cubic)
from test_gpu_parity import pair_engines

pytestmark = pytest.mark.gpu

SIZES = [
    ((5, 5, 5, 1.0, (0.0, 0.0, 1.0), True), 2),  # N = 125: 2 sites per lane
    ((6, 6, 6, 1.0, (0.6, 0.0, 0.8), True), 4),  # N = 216: 4 sites per lane
    ((8, 8, 8, 1.0, (0.0, 0.0, 1.0), True), 8),  # N = 512 (C2's lattice): 8 sites per lane
]


def pilot_grid(gpu, geometry, bin_width=1.0):
    """Energy range of a short pilot sweep, padded by 20 %, snapped to the bin width."""
    system = System(*simple_cubic(*geometry), backend=gpu)
    lo, hi = system.pilot(1234, 20 * system.n)
    pad = 0.2 * (hi - lo) + 2.0
    return (math.floor((lo - pad) / bin_width) * bin_width,
            math.ceil((hi + pad) / bin_width) * bin_width, bin_width)


def stages_config(grid, n, max_epochs, seed=5, windows=2):
    return RunConfig(emin=grid[0], emax=grid[1], bin_width=grid[2], windows=windows,
                     walkers_per_window=2, seed=seed, exchange_interval_attempts=n,
                     max_attempts=n * max_epochs, rng=abi.RNG_PHILOX,
                     wl=WlParameters(final_factor=1e-3, check_interval_attempts=10 * n))


def test_kernel_choice():
    """AUTO runs the staged walker where it applies; a lookahead width or the LOOKAHEAD
    family selects the register kernel; ineligible engines refuse STAGED."""
    from replica_exchange_wang_landau_amd.engine import default_backend

    gpu = default_backend()
    c2 = Engine(System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu),
                c2_config(rng=abi.RNG_PHILOX))
    assert c2.kernel() == (abi.KERNEL_STAGED, "lookahead_staged_kernel<8>")
    c2.set_kernel(abi.KERNEL_LOOKAHEAD)
    assert c2.kernel()[1] == "lookahead_fast_kernel<8>"
    c2.set_kernel(abi.KERNEL_STAGED)
    assert c2.kernel()[1] == "lookahead_staged_kernel<8>"
    c2.set_kernel(abi.KERNEL_AUTO)
    c2.set_lookahead_width(16)
    assert c2.kernel()[1] == "lookahead_fast_kernel<8>"
    xo = Engine(System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu),
                c2_config(rng=abi.RNG_XOSHIRO))
    assert xo.kernel()[1] == "lookahead_xo_kernel<8>"
    with pytest.raises(InvalidArgument, match="staged walker"):
        xo.set_kernel(abi.KERNEL_STAGED)
    small = Engine(System(*simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True), backend=gpu),
                   stages_config((-60.0, 80.0, 0.5), 27, 10))
    assert small.kernel()[1] == "lookahead_fast_kernel<1>"  # 1 site per lane: not staged
    with pytest.raises(InvalidArgument, match="staged walker"):
        small.set_kernel(abi.KERNEL_STAGED)


def test_c2_epochs_bit_exact(gpu, orc):
    """C2 (128 walkers in 8 windows) for 20 MCS on the staged walker against the oracle."""
    cfg = c2_config(max_attempts=512 * 20, rng=abi.RNG_PHILOX)
    eg, eo = pair_engines(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    eg.set_kernel(abi.KERNEL_STAGED)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    assert sg.ambiguous_decisions == 0
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def test_c2_many_windows_bit_exact(gpu, orc):
    """C2's grid in 48 windows of 2 walkers (narrow windows: proposals often leave the window,
    the lowest and highest windows end at the grid's edges) for 30 MCS."""
    cfg = c2_config(max_attempts=512 * 30, rng=abi.RNG_PHILOX, windows=48,
                    walkers_per_window=2)
    eg, eo = pair_engines(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    eg.set_kernel(abi.KERNEL_STAGED)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


@pytest.mark.parametrize("geometry,npl", SIZES)
def test_stages_through_freeze(gpu, orc, geometry, npl):
    """WL -> 1/t -> frozen on 2, 4 and 8 sites per lane, compared after chunks of 5 to 300
    epochs until every walker is frozen."""
    grid = pilot_grid(gpu, geometry)
    n = geometry[0] * geometry[1] * geometry[2]
    eg, eo = pair_engines(gpu, orc, simple_cubic(*geometry), stages_config(grid, n, 12000))
    eg.set_kernel(abi.KERNEL_STAGED)
    assert eg.kernel()[1] == f"lookahead_staged_kernel<{npl}>"
    stages = set()
    for chunk in [5, 10, 20, 40, 80] + [300] * 40:
        sg, so = eg.run(chunk), eo.run(chunk)
        assert (sg.attempted, sg.accepted, sg.exchange_accepted, sg.epoch) == (
            so.attempted, so.accepted, so.exchange_accepted, so.epoch)
        now = set()
        for k in range(eg.walkers):
            assert_walkers_equal(eg, eo, k)
            now.add(int(eg.read_walker(k).stage))
        stages |= now
        if now == {abi.STAGE_FROZEN} or sg.done:
            break
    assert stages == {abi.STAGE_WANG_LANDAU, abi.STAGE_INVERSE_TIME, abi.STAGE_FROZEN}, stages


@pytest.mark.parametrize("geometry,npl", SIZES[:2])
def test_first_visits_in_inverse_time_and_partial_batches(gpu, orc, geometry, npl):
    """Walkers restored into the 1/t stage with visited bins cleared rediscover them (the 1/t
    factors of later attempts change mid-batch); calls of 5000, 777 and 20000 attempts end in
    partial batches."""
    grid = pilot_grid(gpu, geometry)
    n = geometry[0] * geometry[1] * geometry[2]
    eg, eo = pair_engines(gpu, orc, simple_cubic(*geometry), stages_config(grid, n, 0))
    eg.set_kernel(abi.KERNEL_STAGED)
    eg.run(300)
    eo.run(300)
    rng = np.random.default_rng(7)
    for k in range(eg.walkers):
        snap = eg.read_walker(k)
        cur = int(np.floor((snap.energy - grid[0]) / grid[2]))
        lo, hi = int(snap.window_begin), int(snap.window_end)
        visited = [b for b in range(lo, hi) if snap.active[b] and b != cur]
        for b in rng.choice(visited, size=min(6, len(visited)), replace=False):
            snap.active[b] = 0
            snap.histogram[b] = 0
        snap.info.active_bins = int(snap.active[lo:hi].sum())
        snap.info.stage = abi.STAGE_INVERSE_TIME
        snap.info.factor = snap.info.active_bins / snap.info.attempted
        eg.write_walker(k, snap)
        eo.write_walker(k, snap)
    for chunk in (5000, 777, 20000, 13):
        eg.sample(chunk)
        eo.sample(chunk)
        for k in range(eg.walkers):
            assert_walkers_equal(eg, eo, k)


def test_long_c2_run_equals_register_kernel(gpu):
    """1500 C2 epochs (HIP graphs of 16 epochs) on the staged walker and on the register
    lookahead kernel end in identical walkers."""
    engines = []
    for kernel in (abi.KERNEL_STAGED, abi.KERNEL_LOOKAHEAD):
        e = Engine(System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu),
                   c2_config(rng=abi.RNG_PHILOX, seed=99))
        e.initialize()
        e.set_kernel(kernel)
        e.set_graph(16)
        e.run(1500)
        engines.append(e)
    a, b = engines
    sa, sb = a.status(), b.status()
    assert (sa.epoch, sa.attempted, sa.accepted, sa.exchange_accepted) == (
        sb.epoch, sb.attempted, sb.accepted, sb.exchange_accepted)
    for k in range(a.walkers):
        assert_walkers_equal(a, b, k)
