"""The host's adaptive-window planner (host/adaptive.cpp) against the reference's own
adapt_energy_windows and select_adaptive_initial_configurations (analysis.cpp:183-416),
compiled from its sources into oracle/_ref.

Both are fed the same pilot data: random windows, fragments with invalid holes (NaN log g),
per-bin squared displacements and samples, round trips, and representatives with repeated
energies (the candidate sort breaks such ties by position, so the order of the pool matters).
Bar: identical windows, identical starting configurations per walker, identical counts and
identical error texts.
"""
from __future__ import annotations

import ctypes
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from harness import _ptr

ROOT = Path(__file__).resolve().parents[1]
SELFTEST = ROOT / "tests" / "host" / "_bin" / "host_checks"

pytestmark = pytest.mark.skipif(not SELFTEST.exists(), reason="tests/host/_bin/host_checks not built")

DEFAULTS = dict(smoothing_width=0.0, minimum_width=0.0, diffusivity_floor_fraction=0.05,
                curvature_weight=0.25, round_trip_target=2.0, maximum_round_trip_penalty=3.0)
ORDER = list(DEFAULTS)


def partition(bins, count, overlap):
    """Input generation only: evenly spaced overlapping windows over the grid."""
    if count == 1:
        return [(0, bins)]
    size = max(2, min(bins, math.ceil(bins / (1.0 + (count - 1) * (1.0 - overlap)))))
    out = [(round(w * (bins - size) / (count - 1)), 0) for w in range(count)]
    out = [(b, b + size) for b, _ in out]
    out[-1] = (out[-1][0], bins)
    return out


def scenario(seed):
    rng = np.random.default_rng(seed)
    bins = int(rng.integers(40, 400))
    count = int(rng.integers(2, 7))
    overlap = float(rng.choice([0.0, 0.3, 0.5, 0.75]))
    width = float(rng.choice([0.25, 0.5, 1.0]))
    emin = -0.5 * bins * width
    grid = (emin, emin + bins * width, width)
    windows = partition(bins, count, max(overlap, 0.3))
    per_window = int(rng.integers(1, 5))
    params = dict(DEFAULTS)
    if rng.random() < 0.5:
        params["smoothing_width"] = float(rng.uniform(0.5, 6.0)) * width
    if rng.random() < 0.3:
        params["minimum_width"] = float(rng.uniform(1.0, 3.0)) * width
    params["curvature_weight"] = float(rng.choice([0.0, 0.25, 2.0]))
    params["round_trip_target"] = float(rng.choice([0.0, 2.0, 5.0]))
    params["maximum_round_trip_penalty"] = float(rng.choice([1.0, 3.0]))
    frags = []
    for b, e in windows:
        lg = np.full(bins, np.nan)
        valid = np.zeros(bins, dtype=np.uint8)
        shape = np.cumsum(rng.normal(0.0, 1.0, e - b)) + 0.02 * (np.arange(e - b) - (e - b) / 2) ** 2
        ok = rng.random(e - b) > 0.1
        if rng.random() < 0.1:
            ok[:] = False  # a window without a common active bin
        lg[b:e] = np.where(ok, shape, np.nan)
        valid[b:e] = ok
        samples = np.zeros(bins, dtype=np.uint64)
        samples[b:e] = rng.integers(0, 1000, e - b) * (rng.random(e - b) > 0.05)
        sq = np.zeros(bins)
        sq[b:e] = rng.exponential(4.0, e - b) * samples[b:e]
        frags.append(dict(window=(b, e), log_g=lg, valid=valid, sq=sq, samples=samples,
                          round_trips=int(rng.integers(0, 12)), walkers=per_window))
    n = 6
    r_count = int(rng.integers(0, 30))
    energies = rng.uniform(grid[0] - 10.0, grid[1] + 10.0, r_count)
    if r_count > 4:
        energies[1] = energies[3]  # a tie between different configurations
        energies[2] = grid[1]      # exactly the grid maximum
    reps = [(bool(rng.random() > 0.2), float(energies[r]),
             rng.choice(np.array([-1, 1], dtype=np.int8), n)) for r in range(r_count)]
    return dict(bins=bins, count=count, overlap=overlap, grid=grid, per_window=per_window, n=n,
                params=params, frags=frags, reps=reps)


def ours(case, tmp_path):
    lines = [f"{case['bins']} {case['count']} {case['overlap']!r} "
             + " ".join(repr(v) for v in case["grid"]) + f" {case['per_window']} {case['n']}",
             " ".join(repr(case["params"][k]) for k in ORDER)]
    for f in case["frags"]:
        lines.append(f"{f['window'][0]} {f['window'][1]} {f['round_trips']} {f['walkers']}")
        lines.append(" ".join(repr(float(v)) for v in f["sq"]))
        lines.append(" ".join(str(int(v)) for v in f["samples"]))
        lines.append(" ".join(repr(float(v)) for v in f["log_g"]))
        lines.append(" ".join(str(int(v)) for v in f["valid"]))
    lines.append(str(len(case["reps"])))
    for present, energy, spins in case["reps"]:
        lines.append(f"{int(present)} {energy!r} " + " ".join(str(int(s)) for s in spins))
    path = tmp_path / "pilot.txt"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([str(SELFTEST), "adapt", str(path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def theirs(ref, case):
    bins, count = case["bins"], case["count"]
    be = np.array([v for f in case["frags"] for v in f["window"]], dtype=np.uint64)
    cat = lambda key, dt: np.ascontiguousarray(np.concatenate([f[key] for f in case["frags"]]),
                                               dtype=dt)
    log_g, valid = cat("log_g", np.float64), cat("valid", np.uint8)
    sq, samples = cat("sq", np.float64), cat("samples", np.uint64)
    trips = np.array([f["round_trips"] for f in case["frags"]], dtype=np.uint64)
    walkers = np.array([f["walkers"] for f in case["frags"]], dtype=np.uint64)
    params = np.array([case["params"][k] for k in ORDER])
    out = np.zeros(2 * count, dtype=np.uint64)
    d = ctypes.c_double
    st = ref.lib.ref_adapt_windows(d(case["grid"][0]), d(case["grid"][1]), d(case["grid"][2]),
                                   ctypes.c_uint64(bins), ctypes.c_uint64(count),
                                   d(case["overlap"]), _ptr(params, d), _ptr(be, ctypes.c_uint64),
                                   _ptr(log_g, d), _ptr(valid, ctypes.c_uint8), _ptr(sq, d),
                                   _ptr(samples, ctypes.c_uint64), _ptr(trips, ctypes.c_uint64),
                                   _ptr(walkers, ctypes.c_uint64), _ptr(out, ctypes.c_uint64))
    if st != 0:
        return [f"error {ref.error()}"]
    lines = ["windows " + " ".join(str(int(v)) for v in out)]
    n, per = case["n"], case["per_window"]
    reps = case["reps"]
    energies = np.array([r[1] for r in reps] or [0.0])
    present = np.array([r[0] for r in reps] or [0], dtype=np.uint8)
    spins = np.ascontiguousarray(np.concatenate([r[2] for r in reps]) if reps
                                 else np.zeros(n, dtype=np.int8), dtype=np.int8)
    out_spins = np.zeros(count * per * n, dtype=np.int8)
    out_present = np.zeros(count * per, dtype=np.uint8)
    missing, external = ctypes.c_uint64(), ctypes.c_uint64()
    st = ref.lib.ref_select_initial(d(case["grid"][0]), d(case["grid"][1]), d(case["grid"][2]),
                                    ctypes.c_uint64(count), _ptr(out, ctypes.c_uint64),
                                    ctypes.c_uint64(len(reps)), _ptr(energies, d),
                                    _ptr(present, ctypes.c_uint8), ctypes.c_uint64(n),
                                    _ptr(spins, ctypes.c_int8), ctypes.c_uint64(per),
                                    _ptr(out_spins, ctypes.c_int8),
                                    _ptr(out_present, ctypes.c_uint8), ctypes.byref(missing),
                                    ctypes.byref(external))
    if st != 0:
        return lines + [f"error {ref.error()}"]
    lines.append(f"missing {missing.value} external {external.value}")
    for k in range(count * per):
        lines.append(" ".join(str(int(v)) for v in out_spins[k * n:(k + 1) * n])
                     if out_present[k] else "-")
    return lines


@pytest.mark.parametrize("seed", range(40))
def test_planner_matches_reference(ref, tmp_path, seed):
    case = scenario(seed)
    assert ours(case, tmp_path) == theirs(ref, case)


def test_planner_moves_windows_towards_slow_diffusion(ref, tmp_path):
    """A pilot whose upper half diffuses 100x slower: both planners shrink the upper windows
    (more windows where walkers move slowly) and agree on every boundary."""
    case = scenario(1)
    bins, count = 200, 4
    case.update(bins=bins, count=count, overlap=0.5, grid=(-50.0, 50.0, 0.5))
    case["params"] = dict(DEFAULTS)
    case["frags"] = []
    for b, e in partition(bins, count, 0.5):
        samples = np.zeros(bins, dtype=np.uint64)
        samples[b:e] = 100
        sq = np.zeros(bins)
        sq[b:e] = np.where(np.arange(b, e) < bins // 2, 100.0, 1.0) * 100
        lg = np.full(bins, np.nan)
        lg[b:e] = 0.0
        valid = np.zeros(bins, dtype=np.uint8)
        valid[b:e] = 1
        case["frags"].append(dict(window=(b, e), log_g=lg, valid=valid, sq=sq, samples=samples,
                                  round_trips=4, walkers=2))
    a = ours(case, tmp_path)
    assert a == theirs(ref, case)
    edges = [int(v) for v in a[0].split()[1:]]
    widths = [edges[2 * k + 1] - edges[2 * k] for k in range(count)]
    assert widths[-1] < widths[0]


@pytest.mark.parametrize("bad", [dict(diffusivity_floor_fraction=0.0),
                                 dict(curvature_weight=-1.0),
                                 dict(maximum_round_trip_penalty=0.5),
                                 dict(smoothing_width=float("nan")),
                                 dict(minimum_width=1e6)])
def test_planner_errors_match_reference(ref, tmp_path, bad):
    case = scenario(3)
    case["params"].update(bad)
    a = ours(case, tmp_path)
    assert a[0].startswith("error ")
    assert a == theirs(ref, case)
