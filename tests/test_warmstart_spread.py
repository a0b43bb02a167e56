"""The Gaussian part of the warm-start target (warmstart.energy_spread, gaussian_log_g), pinned
by brute force rather than by its formula: over all 2^N spin configurations of a small system,
E = sum_{i<j} J_ij s_i s_j has mean 0 and variance s N."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from replica_exchange_wang_landau_amd import System
from replica_exchange_wang_landau_amd import warmstart as ws
from replica_exchange_wang_landau_amd.engine import simple_cubic


def brute_force_moments(J: np.ndarray) -> tuple[float, float]:
    n = J.shape[0]
    s = np.array(list(itertools.product((-1.0, 1.0), repeat=n)))  # 2^n x n
    energy = 0.5 * np.einsum("ci,ij,cj->c", s, J, s)  # sum over i < j (J symmetric, J_ii = 0)
    return float(energy.mean()), float(energy.var())


@pytest.mark.parametrize("system_kind", ["csr slab", "dense cube"])
def test_energy_spread_is_the_infinite_temperature_variance(orc, system_kind):
    if system_kind == "csr slab":
        slab = ws.Slab(3, 2, 0.2, 1, 11)  # 3 x 3 x 2 diluted, in-plane axes, cutoff 3
        system = System(*slab.geometry(), backend=orc, cutoff=3.0)
    else:
        system = System(*simple_cubic(2, 2, 3, 1.0, (0.0, 0.6, 0.8), True), backend=orc)
    J = system.dense()
    assert np.array_equal(J, J.T) and not np.diag(J).any()
    mean, var = brute_force_moments(J)
    spread = ws.energy_spread(system)
    assert abs(mean) < 1e-9 * np.sqrt(var)
    assert var == pytest.approx(spread * system.n, rel=1e-12)


def test_gaussian_part_peaks_at_zero_with_the_spread_as_curvature():
    e = np.array([-0.5, -0.1, 0.0, 0.1, 0.5])
    g = ws.gaussian_log_g(e, 0.25)
    assert g[2] == 0.0 and np.all(g[[0, 1, 3, 4]] < 0.0)
    assert g == pytest.approx(-(e * e) / 0.5)  # -(E/N)^2 / (2 s)
    # the second difference of g over an energy step h is -h^2 / s
    h = 0.1
    assert (ws.gaussian_log_g(h, 0.25) - 2 * ws.gaussian_log_g(0.0, 0.25)
            + ws.gaussian_log_g(-h, 0.25)) == pytest.approx(-h * h / 0.25)
