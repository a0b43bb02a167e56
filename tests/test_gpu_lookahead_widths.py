"""Lookahead widths of the Philox common-configuration kernel (wlx_engine_set_lookahead_width):
4, 8, 16 and 32 attempts per batch must all give the one-flip-at-a-time walker's trajectory,
through WL -> 1/t -> frozen, first visits that change later 1/t factors within a batch, and
register layouts from 1 to 64 sites per lane."""
from __future__ import annotations

import numpy as np
import pytest

from harness import assert_walkers_equal, c2_config
from replica_exchange_wang_landau_amd import abi
from replica_exchange_wang_landau_amd.engine import (InvalidArgument, RunConfig, WlParameters,
                                                     simple_cubic)
from test_gpu_parity import _stages_config, pair_engines

pytestmark = pytest.mark.gpu

WIDTHS = [4, 8, 16, 32]


@pytest.mark.parametrize("width", WIDTHS)
def test_width_stages_through_freeze(gpu, orc, width):
    eg, eo = pair_engines(gpu, orc, simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True),
                          _stages_config(4000))
    eg.set_lookahead_width(width)
    name = eg.kernel()[1]
    assert name == ("lookahead_fast_kernel<1>" if width == 16
                    else f"lookahead_fast_kernel<1,{width}>")
    stages = set()
    for _ in range(4):
        sg, so = eg.run(1000), eo.run(1000)
        assert (sg.attempted, sg.accepted, sg.exchange_accepted, sg.epoch) == (
            so.attempted, so.accepted, so.exchange_accepted, so.epoch)
        for k in range(eg.walkers):
            assert_walkers_equal(eg, eo, k)
            stages.add(int(eg.read_walker(k).stage))
    assert stages == {abi.STAGE_WANG_LANDAU, abi.STAGE_INVERSE_TIME, abi.STAGE_FROZEN}


@pytest.mark.parametrize("width", WIDTHS)
def test_width_first_visits_in_inverse_time(gpu, orc, width):
    eg, eo = pair_engines(gpu, orc, simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True),
                          _stages_config(0))
    eg.set_lookahead_width(width)
    eg.run(300)
    eo.run(300)
    rng = np.random.default_rng(7)
    for k in range(eg.walkers):
        snap = eg.read_walker(k)
        cur = int(np.floor((snap.energy + 60.0) / 0.5))
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
    for chunk in (5000, 777, 20000):
        eg.sample(chunk)
        eo.sample(chunk)
        for k in range(eg.walkers):
            assert_walkers_equal(eg, eo, k)


@pytest.mark.parametrize("width", WIDTHS)
def test_width_c2_epochs(gpu, orc, width):
    cfg = c2_config(max_attempts=512 * 20, rng=abi.RNG_PHILOX)
    eg, eo = pair_engines(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    eg.set_lookahead_width(width)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("geometry,grid", [
    ((16, 16, 4, 1.0, (0.0, 0.0, 1.0), True), (-700.0, 900.0, 2.0)),    # N=1024, 16 per lane
    ((16, 16, 8, 1.0, (0.0, 0.0, 1.0), True), (-1400.0, 1800.0, 4.0)),  # N=2048, 32 per lane
    ((15, 15, 15, 1.0, (0.0, 0.6, 0.8), True), (-2600.0, 3200.0, 5.0)),  # N=3375, 64 per lane
])
def test_width_uneven_sizes(gpu, orc, width, geometry, grid):
    """Width 32 was shown exact up to 16 sites per lane only and is refused beyond."""
    cfg = RunConfig(emin=grid[0], emax=grid[1], bin_width=grid[2], windows=2,
                    walkers_per_window=2, seed=31, exchange_interval_attempts=300,
                    max_attempts=300 * 12, rng=abi.RNG_PHILOX,
                    wl=WlParameters(check_interval_attempts=600))
    eg, eo = pair_engines(gpu, orc, simple_cubic(*geometry), cfg)
    if width == 32 and eg.n > 1024:
        with pytest.raises(InvalidArgument, match="at most 1024 sites"):
            eg.set_lookahead_width(width)
        return
    eg.set_lookahead_width(width)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
