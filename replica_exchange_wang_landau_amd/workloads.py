"""Synthetic workloads of the batched configuration: uneven dilute dipole systems sampled
together in one launch per epoch.

Each system: N sites placed uniformly at random in a periodic cube of side (N / 0.3)^(1/3)
(number density 0.3), easy axes uniform on the unit sphere, dipolar couplings D = 1 within a
CSR cutoff of 3 (about 34 neighbours per site), stored as FP32. Sizes are drawn uniformly
from {64, ..., 4096}; each system gets clamp(round(4 + 12 (N - 64) / 4032), 4, 16) energy
windows of 8 walkers at overlap 0.75, bins of width 1.0 over its pilot energy range padded
by 5 %, exchanges every MCS and the Philox proposal stream.

Uniform placement puts some pairs very close (r ~ 0.04 for N ~ 4000), whose couplings
(~1/r^3) stretch the energy range to ~10^5 bins of width 1; windows that do not fit in LDS
keep their bins in HBM (the engine decides this per system). The same close pairs make the
energy landscape coarse and trap the target-Metropolis initialisation (wang_landau.cpp:120-241)
in the default band around each window's centre, so the walkers may start anywhere in their
window (target fraction 1.0), raise the temperature after 20 stalled attempts per spin instead
of 1000, and try up to 4 x 10^6 attempts (all WlParameters of the reference).

A pair much closer than the rest splits the density of states into separate lumps (its two
relative orientations, about +-J apart): the pilot range then spans empty energy gaps, and a
window inside a gap holds no state, which no initialisation can reach. The windows of such a
system are therefore laid over the occupied energy intervals only: energies of random
configurations locate the gaps (spacings above a tenth of the pilot range), each occupied
interval gets a share of the system's windows in proportion to its width, and each share is
partitioned like partition_windows (wang_landau.cpp:40-61). A system without gaps gets
exactly partition_windows over its whole grid.
"""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .engine import RunConfig, System, WlParameters

DENSITY = 0.3
CUTOFF = 3.0
MIN_SPINS, MAX_SPINS = 64, 4096


def sizes(count: int = 64, seed: int = 1234) -> np.ndarray:
    """System sizes, uniform in {64, ..., 4096}, from default_rng(seed)."""
    return np.random.default_rng(seed).integers(MIN_SPINS, MAX_SPINS + 1, size=count)


def geometry(n: int, seed: int):
    """(positions, axes, box, periodic) of one system from default_rng(seed): positions
    uniform in the periodic cube, axes normalised Gaussian vectors (uniform on the sphere)."""
    rng = np.random.default_rng(seed)
    side = (n / DENSITY) ** (1 / 3)
    positions = rng.uniform(0.0, side, size=(n, 3))
    axes = rng.normal(size=(n, 3))
    axes /= np.sqrt((axes * axes).sum(axis=1))[:, None]
    return positions, axes, np.array([side, side, side]), True


def windows_for(n: int) -> int:
    return min(16, max(4, round(4 + 12 * (n - 64) / 4032)))


def systems(count: int = 64, seed: int = 1234, backend=None, n_list=None):
    """The batch's coupling systems (FP32 CSR couplings, cutoff 3); system s uses
    default_rng(seed + s) for its geometry."""
    ns = sizes(count, seed) if n_list is None else n_list
    return [System(*geometry(int(n), seed + s), cutoff=CUTOFF, backend=backend,
                   precision=abi.COUPLINGS_FP32) for s, n in enumerate(ns)]


def grid(lo: float, hi: float, bin_width: float = 1.0, padding: float = 0.05):
    """Pilot range padded by `padding` of its span, snapped outward to the bin width."""
    pad = padding * (hi - lo)
    return (bin_width * math.floor((lo - pad) / bin_width),
            bin_width * math.ceil((hi + pad) / bin_width))


def partition(bins: int, count: int, overlap: float):
    """partition_windows (wang_landau.cpp:40-61): `count` windows of equal width over `bins`
    bins with the given overlap, as (begin, end) pairs."""
    if count == 1:
        return [(0, bins)]
    width = math.ceil(bins / (1.0 + (count - 1) * (1.0 - overlap)))
    width = min(max(width, 2), bins)
    out = []
    for w in range(count):
        begin = int(math.floor(w * (bins - width) / (count - 1) + 0.5))  # llround
        out.append((begin, bins if w == count - 1 else begin + width))
    return out


def sample_energies(system: System, samples: int = 2048, seed: int = 0) -> np.ndarray:
    """Energies of `samples` uniformly random configurations (local_fields/total_energy)."""
    spins = np.random.default_rng(seed).choice(np.array([-1, 1], dtype=np.int8),
                                               size=(samples, system.n))
    return system.fields(spins)[1]


def occupied_windows(emin: float, emax: float, count: int, energies, overlap: float = 0.75,
                     bin_width: float = 1.0):
    """Windows over the occupied energy intervals of [emin, emax): gaps are spacings of the
    sorted sample energies (those inside the grid) wider than a tenth of the range; interior
    interval edges reach 10 % into their gaps. Each interval gets windows in proportion to its
    width (at most one per two bins). Returns (begin, end) bin pairs."""
    bins = int(round((emax - emin) / bin_width))
    e = np.sort(np.asarray(energies, dtype=np.float64))
    e = e[(e >= emin) & (e <= emax)]
    gap = np.nonzero(np.diff(e) > 0.1 * (emax - emin))[0]
    if len(gap) == 0:
        return partition(bins, count, overlap)
    edges = [emin]
    for g in gap:
        a, b = e[g], e[g + 1]
        edges += [a + 0.1 * (b - a), b - 0.1 * (b - a)]
    edges.append(emax)
    intervals = []
    for k in range(0, len(edges), 2):
        lo_bin = max(0, int(math.floor((edges[k] - emin) / bin_width)))
        hi_bin = min(bins, int(math.ceil((edges[k + 1] - emin) / bin_width)))
        if hi_bin - lo_bin >= 2:
            intervals.append((lo_bin, hi_bin))
    widths = np.array([b - a for a, b in intervals], dtype=np.float64)
    cap = np.maximum(1, widths.astype(int) // 2)
    if cap.sum() < count:
        return partition(bins, count, overlap)
    share = np.minimum(cap, np.maximum(1, np.floor(count * widths / widths.sum()))).astype(int)
    while share.sum() < count:  # the rest to the widest intervals per window, within caps
        room = np.where(share < cap, widths / share, -1.0)
        share[int(np.argmax(room))] += 1
    while share.sum() > count:
        share[int(np.argmax(np.where(share > 1, share, 0)))] -= 1
    out = []
    for (a, b), k in zip(intervals, share):
        out += [(a + x, a + y) for x, y in partition(b - a, int(k), overlap)]
    return out


def config(system: System, lo: float, hi: float, seed: int, walkers_per_window: int = 8,
           final_factor: float = 1e-7, max_attempts: int = 0, exchange_mcs: int = 1,
           energies=None) -> RunConfig:
    """The run configuration of one system of the batch; `energies` (sample_energies) lays
    the windows over the occupied intervals."""
    n = system.n
    emin, emax = grid(lo, hi)
    count = windows_for(n)
    explicit = None
    if energies is not None:
        explicit = occupied_windows(emin, emax, count, energies)
        if explicit == partition(int(round(emax - emin)), count, 0.75):
            explicit = None  # no gap: the reference's own partition
    return RunConfig(emin=float(emin), emax=float(emax), bin_width=1.0,
                     windows=count, walkers_per_window=walkers_per_window,
                     explicit_windows=explicit,
                     overlap=0.75, seed=seed, exchange_interval_attempts=exchange_mcs * n,
                     max_attempts=max_attempts, rng=abi.RNG_PHILOX,
                     wl=WlParameters(final_factor=final_factor,
                                     check_interval_attempts=100 * n,
                                     initialization_max_attempts=4_000_000,
                                     initialization_target_fraction=1.0,
                                     initialization_stall_attempts_per_spin=20))
