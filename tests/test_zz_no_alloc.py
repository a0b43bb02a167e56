"""wlx_system_fields_noalloc: local_fields + total_energy staged through a fixed 4 MiB
device scratch array instead of per-call device allocations. Its results must equal
wlx_system_fields bit for bit, on consecutive calls of decreasing size (the scratch keeps
the previous call's contents), and oversize requests must fail with a clear error."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from replica_exchange_wang_landau_amd import abi
from replica_exchange_wang_landau_amd.engine import InvalidArgument, System, _ptr, simple_cubic

pytestmark = pytest.mark.gpu


def noalloc(system, spins, with_fields=True):
    lib = abi.load_library()
    count = spins.shape[0]
    fields = np.full((count, system.n), np.nan) if with_fields else None
    energies = np.full(count, np.nan)
    status = lib.wlx_system_fields_noalloc(
        system.handle, _ptr(np.ascontiguousarray(spins), ctypes.c_int8), ctypes.c_uint64(count),
        _ptr(fields, ctypes.c_double), _ptr(energies, ctypes.c_double))
    system.backend.check(status)
    return fields, energies


def test_noalloc_fields_equal_allocating_entry_point(gpu):
    rng = np.random.default_rng(7)
    c2 = System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu)
    csr = System(rng.uniform(0, 9.0, (700, 3)), np.tile([0.0, 0.6, 0.8], (700, 1)),
                 np.array([9.0, 9.0, 9.0]), True, cutoff=3.0, backend=gpu)
    for system, counts in ((c2, (700, 33)), (csr, (200, 5))):
        for count in counts:  # the second call is smaller: stale scratch must not leak in
            spins = rng.choice(np.array([-1, 1], dtype=np.int8), size=(count, system.n))
            want_f, want_e = system.fields(spins)
            got_f, got_e = noalloc(system, spins)
            assert got_f.tobytes() == want_f.tobytes()
            assert got_e.tobytes() == want_e.tobytes()
            _, only_e = noalloc(system, spins, with_fields=False)
            assert only_e.tobytes() == want_e.tobytes()


def test_noalloc_fields_reject_oversize_and_bad_spins(gpu):
    c2 = System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=gpu)
    too_many = np.ones((1000, c2.n), dtype=np.int8)  # 1000 * 512 * 9 B > 4 MiB
    with pytest.raises(InvalidArgument, match="4 MiB scratch"):
        noalloc(c2, too_many)
    bad = np.ones((2, c2.n), dtype=np.int8)
    bad[1, 3] = 0
    with pytest.raises(InvalidArgument, match="Spins must be"):
        noalloc(c2, bad)
    ok = np.ones((3, c2.n), dtype=np.int8)
    assert noalloc(c2, ok)[1].tobytes() == c2.fields(ok)[1].tobytes()
