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
keep their bins in HBM (the engine decides this per system).
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


def config(system: System, lo: float, hi: float, seed: int, walkers_per_window: int = 8,
           final_factor: float = 1e-7, max_attempts: int = 0,
           exchange_mcs: int = 1) -> RunConfig:
    """The run configuration of one system of the batch."""
    n = system.n
    emin, emax = grid(lo, hi)
    return RunConfig(emin=float(emin), emax=float(emax), bin_width=1.0,
                     windows=windows_for(n), walkers_per_window=walkers_per_window,
                     overlap=0.75, seed=seed, exchange_interval_attempts=exchange_mcs * n,
                     max_attempts=max_attempts, rng=abi.RNG_PHILOX,
                     wl=WlParameters(final_factor=final_factor,
                                     check_interval_attempts=100 * n))
