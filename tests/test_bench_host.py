"""Host-side pieces of bench.py that need no GPU: the one-GPU refusal, the counter summary it
reads the traffic figure from, the time-to-accuracy criterion, the C2 run configurations and the
batched workload's windows."""
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


def test_traffic_figures_come_from_the_committed_counter_summaries():
    for kernel in ("lookahead_staged_kernel<8>", "lookahead_fast_kernel<8>",
                   "lookahead_xo_kernel<8>"):
        value = bench.traffic_bytes(kernel)
        assert value is not None and 1e6 < value < 1e9
    assert bench.traffic_bytes("no_such_kernel<1>") is None


def test_stays_below_takes_the_first_snapshot_of_the_last_run_below():
    """Time to accuracy counts from the first snapshot after which every later one is below
    the threshold: a dip that comes back up does not count, and a run that ends above it has
    no time."""
    t = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    assert bench.stays_below(t, [0.5, 0.009, 0.02, 0.008, 0.007, 0.006], 0.01) == 4.0
    assert bench.stays_below(t, [0.005] * 6, 0.01) == 1.0
    assert bench.stays_below(t, [0.5, 0.009, 0.008, 0.007, 0.006, 0.011], 0.01) is None
    assert bench.stays_below([], [], 0.01) is None
    assert bench.stays_below(t, [0.5, 0.2, 0.1, 0.05, 0.02, 0.0099], 0.01) == 6.0


def test_c2_configurations():
    plain = bench.c2_config(rng=1)
    many = bench.c2_config(rng=1, seed=77, windows=48)
    assert (plain.windows, plain.walkers_per_window, plain.seed) == (8, 16, 1234)
    assert (many.windows, many.walkers_per_window, many.seed) == (48, 16, 77)
    for c in (plain, many):
        assert (c.emin, c.emax, c.bin_width) == bench.C2_GRID
        assert c.exchange_interval_attempts == 512 and c.overlap == 0.75
        assert round((c.emax - c.emin) / c.bin_width) == 1255


def test_partition_matches_partition_windows(orc):
    """workloads.partition restates partition_windows (wang_landau.cpp:40-61) for explicit
    windows: same windows as the oracle's for many bin counts, window counts and overlaps."""
    import ctypes

    from replica_exchange_wang_landau_amd import workloads
    for bins in (2, 7, 57, 100, 1255, 28490, 99999):
        for count in (1, 2, 3, 5, 8, 16):
            if count > bins:
                continue
            for overlap in (0.0, 0.5, 0.75, 0.9):
                be = (ctypes.c_uint64 * (2 * count))()
                if orc.lib.orc_partition_windows(ctypes.c_uint64(bins), ctypes.c_uint64(count),
                                                 ctypes.c_double(overlap), be) != 0:
                    continue  # windows that do not overlap are refused
                want = [(be[2 * k], be[2 * k + 1]) for k in range(count)]
                assert workloads.partition(bins, count, overlap) == want, (bins, count, overlap)


def test_occupied_windows_skip_gaps():
    """Energies in two lumps 1000 units apart on a 3000-bin grid: the 8 windows cover the two
    lumps (plus 10 % of the gap on their inner side) and none lies inside the gap; without a
    gap the windows are partition_windows over the whole grid."""
    import numpy as np

    from replica_exchange_wang_landau_amd import workloads
    rng = np.random.default_rng(1)
    energies = np.concatenate([rng.uniform(-1500, -500, 500), rng.uniform(500, 1500, 300)])
    windows = workloads.occupied_windows(-1500.0, 1500.0, 8, energies)
    assert len(windows) == 8
    gap_lo, gap_hi = int(energies[energies < 0].max() + 1500), int(energies[energies > 0].min() + 1500)
    for b, e in windows:
        assert e <= gap_lo + 0.1 * (gap_hi - gap_lo) + 2 or b >= gap_hi - 0.1 * (gap_hi - gap_lo) - 2
    assert windows[0][0] == 0 and windows[-1][1] == 3000
    lower = [w for w in windows if w[1] <= 1500]
    assert len(lower) == 4  # widths ~1100 and ~1100 bins share 8 windows
    flat = workloads.occupied_windows(-1500.0, 1500.0, 8, rng.uniform(-1500, 1500, 800))
    assert flat == workloads.partition(3000, 8, 0.75)


def test_occupied_windows_ignore_samples_outside_the_grid():
    """Sample energies beyond the pilot grid do not create intervals outside it; every window
    lies inside the grid and is at least two bins wide."""
    import numpy as np

    from replica_exchange_wang_landau_amd import workloads
    rng = np.random.default_rng(2)
    energies = np.concatenate([rng.uniform(-900, -400, 400), rng.uniform(450, 700, 50),
                               rng.uniform(2000, 2600, 100)])  # a lump above the grid
    windows = workloads.occupied_windows(-1000.0, 1000.0, 12, energies)
    assert len(windows) == 12
    assert all(0 <= b and b + 2 <= e <= 2000 for b, e in windows)
