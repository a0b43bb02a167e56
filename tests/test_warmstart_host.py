"""Host side of the learned warm start: the slab family's geometry and the network's input
scaling (no GPU)."""
from __future__ import annotations

import numpy as np

from replica_exchange_wang_landau_amd.warmstart import FEATURES, Slab, network_input, slab_family


def test_slab_geometry():
    s = Slab(8, 3, 0.25, 1, 42)
    pos, axes, box, periodic = s.geometry()
    assert periodic and list(box) == [8.0, 8.0, 0.0]  # z open: box length 0
    assert 0.6 * 192 < len(pos) < 0.9 * 192
    assert np.all(pos >= 0) and np.all(pos[:, :2] < 8) and np.all(pos[:, 2] < 3)
    assert np.allclose(np.linalg.norm(axes, axis=1), 1.0) and not axes[:, 2].any()
    pos2, axes2, _, _ = s.geometry()
    assert np.array_equal(pos, pos2) and np.array_equal(axes, axes2)
    full = Slab(6, 2, 0.0, 0, 1).geometry()
    assert len(full[0]) == 72 and np.all(full[1] == [0.0, 0.0, 1.0])
    n, density, dilution, thickness, mode = s.features(len(pos))
    assert density == len(pos) / 192 and thickness == 3.0 and mode == 1.0


def test_family_ranges_and_input_rows():
    fam = slab_family(200, 7)
    assert {s.lx for s in fam} <= set(range(6, 13)) and {s.lz for s in fam} <= set(range(2, 7))
    assert all(0.0 <= s.dilution <= 0.4 for s in fam) and {s.axis_mode for s in fam} == {0, 1}
    x = network_input(np.array([-2.0, 0.0, 1.0]), 500.0, 0.8, 0.2, 4.0, 1.0)
    assert x.shape == (3, FEATURES) and x.dtype == np.float32
    assert np.allclose(x[:, 0], [-1.0, 0.0, 0.5]) and np.allclose(x[0, 1:], [0.0, 0.8, 0.5, 0.0, 1.0])
