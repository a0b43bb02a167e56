"""Host-side DOS assembly from device walkers: summarize_walkers and stitch_dos.

Mirrors the host shell (host/analysis.cpp), which follows the reference:
- summarize_impl (parallel.cpp:23-62);
- stitch_dos (analysis.cpp:45-130).

The arithmetic is done in the same order, so the results match the reference's values
bit for bit. Also provides the accuracy measure used for time-to-accuracy: max|Δ log g| after
mean-offset alignment over the bins valid in both estimates (test_main.cpp:377-390).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Fragment:
    """DosFragment of one window (parallel.hpp:14-21): arrays span the full grid."""
    begin: int
    end: int
    log_g: np.ndarray
    histogram: np.ndarray
    standard_error: np.ndarray
    valid: np.ndarray


def summarize_window(snapshots, bins: int, allow_empty: bool = False) -> Fragment:
    """summarize_impl (parallel.cpp:23-62) over the walkers of one window.

    A bin is valid when every walker has visited it. log g is the walker mean of
    log_g[i] - log_g[first valid bin], summed in walker order.
    """
    if not snapshots:
        raise ValueError("No walkers to summarize")
    begin, end = int(snapshots[0].window_begin), int(snapshots[0].window_end)
    log_g = np.full(bins, np.nan)
    err = np.full(bins, np.nan)
    hist = np.zeros(bins, dtype=np.uint64)
    valid = np.zeros(bins, dtype=np.uint8)
    act = np.ones(end - begin, dtype=bool)
    for s in snapshots:
        act &= s.active[begin:end] != 0
        hist[begin:end] += s.histogram[begin:end]
    valid[begin:end] = act
    idx = np.nonzero(act)[0]
    if idx.size == 0:
        if allow_empty:
            return Fragment(begin, end, log_g, hist, err, valid)
        raise RuntimeError("Walkers in an energy window have no common active bin")
    ref = begin + int(idx[0])
    count = float(len(snapshots))
    mean = np.zeros(end - begin)
    for s in snapshots:  # sequential sum in walker order, as the reference
        mean += s.log_g[begin:end] - s.log_g[ref]
    mean /= count
    log_g[begin:end] = np.where(act, mean, np.nan)
    if len(snapshots) > 1:
        squares = np.zeros(end - begin)
        for s in snapshots:
            d = (s.log_g[begin:end] - s.log_g[ref]) - mean
            squares += d * d
        err[begin:end] = np.where(act, np.sqrt(squares / (count * (count - 1.0))), np.nan)
    return Fragment(begin, end, log_g, hist, err, valid)


def _local_slope(values, valid, center, begin, end, width):
    """Least-squares slope over up to five valid bins around center (analysis.cpp:22-41)."""
    lo = max(begin, center - 2) if center > 2 else begin
    hi = min(end, center + 3)
    xs = [i for i in range(lo, hi) if valid[i] and math.isfinite(values[i])]
    if len(xs) < 2:
        return math.nan
    mx = 0.0
    my = 0.0
    for i in xs:
        mx += float(i)
        my += float(values[i])
    mx /= len(xs)
    my /= len(xs)
    num = 0.0
    den = 0.0
    for i in xs:
        dx = float(i) - mx
        num += dx * (float(values[i]) - my)
        den += dx * dx
    return num / den / width if den > 0.0 else math.nan


def _log_sum_exp(values) -> float:
    """max-shifted log(sum exp) over the finite entries (analysis.cpp:12-20)."""
    finite = [float(v) for v in values if math.isfinite(v)]
    if not finite:
        return -math.inf
    top = max(finite)
    total = 0.0
    for v in finite:
        total += math.exp(v - top)
    return top + math.log(total)


def stitch(bins: int, width: float, fragments, complete_range: bool = False,
           spins: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """stitch_dos (analysis.cpp:45-130): windows in order of their begin bin, each joined at
    the overlap bin of best-matching local slope (or the valid bin nearest the overlap
    midpoint) and shifted onto the running estimate. Then either normalised to
    N ln 2 - log_sum_exp (complete_range; every bin must be valid) or shifted so that the
    largest valid log g is 0. Returns (log_g, valid) over the full grid."""
    frags = sorted(fragments, key=lambda f: f.begin)
    log_g = np.full(bins, np.nan)
    valid = np.zeros(bins, dtype=np.uint8)
    first = frags[0]
    valid[first.begin:first.end] = first.valid[first.begin:first.end]
    sel = np.arange(first.begin, first.end)
    ok = first.valid[sel] != 0
    log_g[sel[ok]] = first.log_g[sel[ok]]
    covered_end = first.end
    for right in frags[1:]:
        if right.begin >= covered_end:
            raise ValueError("DOS fragments must overlap and match the global grid")
        ob, oe = right.begin, min(covered_end, right.end)
        mid = ob + (oe - ob) // 2
        join, best_dist = oe, None
        for i in range(ob, oe):
            if (valid[i] and right.valid[i] and math.isfinite(log_g[i])
                    and math.isfinite(right.log_g[i])):
                dist = abs(i - mid)
                if best_dist is None or dist < best_dist:
                    best_dist, join = dist, i
        if join == oe:
            raise RuntimeError("Adjacent DOS fragments have no common valid overlap bin")
        best = math.inf
        for i in range(ob + 2, oe - 2):
            if not (valid[i] and right.valid[i]):
                continue
            lb = _local_slope(log_g, valid, i, ob, oe, width)
            rb = _local_slope(right.log_g, right.valid, i, ob, oe, width)
            diff = abs(lb - rb)
            if math.isfinite(diff) and diff < best:
                best, join = diff, i
        shift = log_g[join] - right.log_g[join]
        for i in range(join, right.end):
            valid[i] = right.valid[i]
            log_g[i] = right.log_g[i] + shift if right.valid[i] else np.nan
        covered_end = max(covered_end, right.end)
    if complete_range:
        if not valid.all():
            raise RuntimeError("Complete-range normalization requires every DOS bin to be valid")
        log_g = log_g + (float(spins) * math.log(2.0) - _log_sum_exp(log_g))
    else:
        if not math.isfinite(_log_sum_exp(log_g)):
            raise RuntimeError("Stitched DOS has no valid bins")
        peak = max(float(v) for v, ok in zip(log_g, valid) if ok)
        log_g = np.where(valid != 0, log_g - peak, log_g)
    return log_g, valid


def engine_dos(engine) -> tuple[np.ndarray, np.ndarray]:
    """Stitched global log g of an engine's walkers (summarize per window, then stitch)."""
    per = engine.config.walkers_per_window
    windows = engine.walkers // per
    frags = []
    for w in range(windows):
        snaps = [engine.read_walker(w * per + p) for p in range(per)]
        frags.append(summarize_window(snaps, engine.bins, allow_empty=True))
    usable = [f for f in frags if f.valid.any()]
    return stitch(engine.bins, engine.config.bin_width, usable)


def max_deviation(log_g, valid, ref_log_g, ref_valid) -> float:
    """max |log g - ref| after removing the mean offset over the bins valid in both
    (test_main.cpp:377-390 aligns the same way). inf when no bin is shared."""
    both = (valid != 0) & (ref_valid != 0) & np.isfinite(log_g) & np.isfinite(ref_log_g)
    if not both.any():
        return math.inf
    d = log_g[both] - ref_log_g[both]
    return float(np.max(np.abs(d - d.mean())))


def average_dos(curves) -> tuple[np.ndarray, np.ndarray]:
    """Mean of several stitched log g curves of one system (independent replicas): each curve
    is shifted to zero mean over the bins valid in every curve, the shifted curves are
    averaged there, and the result is returned with its peak at 0 and that common mask. No
    reference counterpart: the reference runs one REWL job per system (parallel.cpp:456-676)."""
    curves = [(np.asarray(lg, dtype=np.float64), np.asarray(v)) for lg, v in curves]
    if not curves:
        raise ValueError("No curves to average")
    bins = len(curves[0][0])
    if any(len(lg) != bins or len(v) != bins for lg, v in curves):
        raise ValueError("Curves of different grids")
    common = np.ones(bins, dtype=bool)
    for lg, v in curves:
        common &= (v != 0) & np.isfinite(lg)
    if not common.any():
        raise RuntimeError("Replica curves share no valid bin")
    total = np.zeros(int(common.sum()))
    for lg, _ in curves:
        total += lg[common] - lg[common].mean()
    out = np.full(bins, -np.inf)
    out[common] = total / len(curves)
    out[common] -= out[common].max()
    return out, common.astype(np.uint8)
