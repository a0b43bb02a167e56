"""Parity of the HIP engine (kernels a and b) with the C oracle and the golden fixtures.

Bar: bit-exact for couplings, fields, energies, per-attempt traces, walker states, estimators
and run counters (Xoshiro mode = the reference's own arithmetic and stream). Every call goes
through the C ABI (librewl.so).
"""
from __future__ import annotations

import json

import numpy as np
import pytest

import test_oracle_pinning as op
from harness import (GOLDEN, assert_walkers_equal, c1_config, c1_geometry, c2_config)
from replica_exchange_wang_landau_amd import abi
from replica_exchange_wang_landau_amd.engine import (Engine, InvalidArgument, LogicError, RunConfig,
                                                     System, WlParameters, simple_cubic)

pytestmark = pytest.mark.gpu


def pair_systems(gpu, orc, *geometry, **kw):
    return System(*geometry, backend=gpu, **kw), System(*geometry, backend=orc, **kw)


def pair_engines(gpu, orc, geometry, cfg, initialize=True, **kw):
    sg, so = pair_systems(gpu, orc, *geometry, **kw)
    eg, eo = Engine(sg, cfg), Engine(so, cfg)
    if initialize:
        eg.initialize()
        eo.initialize()
    return eg, eo


# ---- kernel (a) ---------------------------------------------------------------------------
def test_dense_couplings_bit_exact(gpu, orc):
    fx = json.loads((GOLDEN / "couplings.json").read_text())
    g, o = pair_systems(gpu, orc, *c1_geometry())
    assert g.dense().tobytes() == o.dense().tobytes() == op.fromhex(fx["c1_dense"]).tobytes()
    g, o = pair_systems(gpu, orc, *simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True))
    assert op.digest(g.dense()) == fx["c2_dense_sha256"]
    g, o = pair_systems(gpu, orc, *simple_cubic(16, 16, 16, 1.0, (0, 0, 1), True))
    assert g.dense().tobytes() == o.dense().tobytes()


def test_random_geometry_dense_and_csr_bit_exact(gpu, orc):
    fx = json.loads((GOLDEN / "couplings.json").read_text())["random_slab"]
    pos = np.array([[float.fromhex(v) for v in p] for p in fx["positions"]])
    ax = np.array([[float.fromhex(v) for v in a] for a in fx["axes"]])
    box = np.array(fx["box"])
    g = System(pos, ax, box, True, backend=gpu)
    assert op.digest(g.dense()) == fx["dense_sha256"]
    gc, oc = pair_systems(gpu, orc, pos, ax, box, True, cutoff=2.0)
    assert op.digest(gc.dense()) == fx["csr2_dense_view_sha256"]
    for a, b in zip(gc.csr(), oc.csr()):
        assert a.tobytes() == b.tobytes()
    # uneven N, tilted axes, anisotropic box, both backends
    rng = np.random.default_rng(11)
    for n, cutoff in [(97, 0.0), (333, 3.0), (1000, 2.5)]:
        L = (n / 0.3) ** (1 / 3)
        p = rng.uniform(0, L, size=(n, 3))
        a = rng.normal(size=(n, 3))
        a /= np.sqrt((a * a).sum(axis=1))[:, None]
        g, o = pair_systems(gpu, orc, p, a, np.array([L, L * 1.1, L * 0.9]), True, cutoff=cutoff)
        assert g.dense().tobytes() == o.dense().tobytes(), (n, cutoff)
        if cutoff > 0:
            for x, y in zip(g.csr(), o.csr()):
                assert x.tobytes() == y.tobytes()


def test_geometry_errors(gpu):
    axes = np.array([[0, 0, 1.0], [0, 0, 1.0]])
    with pytest.raises(InvalidArgument, match="Coincident dipole sites"):
        System(np.zeros((2, 3)), axes, [2, 2, 2], True, backend=gpu)
    with pytest.raises(InvalidArgument, match="Ising axes must be normalized"):
        System([[0, 0, 0], [1, 0, 0]], [[0, 0, 2.0], [0, 0, 1]], [2, 2, 2], True, backend=gpu)
    with pytest.raises(InvalidArgument, match="Geometry contains no sites"):
        System(np.zeros((0, 3)), np.zeros((0, 3)), [2, 2, 2], True, backend=gpu)


def test_fields_and_energies_bit_exact(gpu, orc):
    rng = np.random.default_rng(3)
    for geometry in [c1_geometry(), simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True),
                     simple_cubic(5, 7, 3, 1.0, (0.6, 0, 0.8), True)]:
        g, o = pair_systems(gpu, orc, *geometry)
        spins = rng.choice(np.array([-1, 1], dtype=np.int8), size=(9, g.n))
        fg, eg = g.fields(spins)
        fo, eo = o.fields(spins)
        assert fg.tobytes() == fo.tobytes() and eg.tobytes() == eo.tobytes()
        # E = 1/2 sum s_i h_i against an independent FP64 recompute
        J = g.dense()
        e64 = 0.5 * np.einsum("ci,ij,cj->c", spins.astype(float), J, spins.astype(float))
        np.testing.assert_allclose(eg, e64, rtol=1e-12, atol=1e-12)


# ---- kernel (b): walkers -------------------------------------------------------------------
def test_c1_walker_trace_bit_exact(gpu, orc):
    fx = json.loads((GOLDEN / "walkers.json").read_text())["c1_walker0"]
    eg, eo = pair_engines(gpu, orc, c1_geometry(), c1_config())
    for e in (eg, eo):
        e.trace(fx["attempts"])
        e.sample(fx["attempts"])
    for k in range(2):
        sg, bg, fg = eg.read_trace(k)
        so, bo, fo = eo.read_trace(k)
        assert np.array_equal(sg, so) and np.array_equal(bg, bo) and np.array_equal(fg, fo)
        assert_walkers_equal(eg, eo, k)
    _, bins, flags = eg.read_trace(0)
    assert np.packbits((flags & 1).astype(np.uint8)).tobytes().hex() == fx["accepted_bits"]
    op.check_state(eg.read_walker(0), fx["state"])


def test_forced_nalivaiko_traditional_bit_exact(gpu, orc):
    fx = json.loads((GOLDEN / "walkers.json").read_text())["c2_walker5_forced_nalivaiko"]
    eg, eo = op.forced_nalivaiko_engine(gpu), op.forced_nalivaiko_engine(orc)
    for e in (eg, eo):
        e.trace(fx["attempts"])
        e.sample(fx["attempts"])
        e.check()
    _, bins, _ = eg.read_trace(5)
    assert op.digest(bins.astype(np.int64)) == fx["bins_sha256"]
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
    assert eg.read_walker(5).forced_accepted == fx["state"]["forced_accepted"] > 0


def test_initialization_bit_exact_and_failure_text(gpu, orc):
    pos, axes, box, per = simple_cubic(2, 1, 1, 1.0, (1, 0, 0), False)
    cfg = RunConfig(emin=-2.5, emax=3.5, bin_width=1.0, windows=1, walkers_per_window=3,
                    seed=123, wl=WlParameters(flatness=0.8, minimum_visits=1,
                                              check_interval_attempts=10,
                                              initialization_max_attempts=100))
    eg, eo = pair_engines(gpu, orc, (pos, axes, box, per), cfg)
    for k in range(3):
        assert_walkers_equal(eg, eo, k)
    g, o = op.init_failure_engine(gpu), op.init_failure_engine(orc)
    with pytest.raises(RuntimeError) as gi:
        g.initialize()
    with pytest.raises(RuntimeError) as oi:
        o.initialize()
    assert str(gi.value) == str(oi.value)
    for text in op.INIT_FAILURE_FRAGMENTS:
        assert text in str(gi.value)


def test_c2_initialization_and_supplied_starts_bit_exact(gpu, orc):
    cfg = c2_config()
    geo = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    eg, eo = pair_engines(gpu, orc, geo, cfg)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
    # supplied starts: in-window ones are kept, the rest run the target search
    rng = np.random.default_rng(5)
    spins = rng.choice(np.array([-1, 1], dtype=np.int8), size=(eg.walkers, 512))
    supplied = (np.arange(eg.walkers) % 3 != 0).astype(np.uint8)
    eg2, eo2 = pair_engines(gpu, orc, geo, cfg, initialize=False)
    eg2.initialize(spins, supplied)
    eo2.initialize(spins, supplied)
    for k in range(eg.walkers):
        assert_walkers_equal(eg2, eo2, k)


@pytest.mark.parametrize("name", ["c1_prefix", "exact_regression_2x2x1",
                                  "rewl_2x2x2_two_windows", "c2_prefix_20mcs"])
def test_run_rewl_golden_on_gpu(gpu, name):
    fx = next(r for r in json.loads((GOLDEN / "runs.json").read_text()) if r["name"] == name)
    nx, ny, nz, periodic = fx["geometry"]
    s = System(*simple_cubic(nx, ny, nz, 1.0, (0, 0, 1), periodic), backend=gpu)
    e = Engine(s, op.config_from_fixture(fx))
    e.initialize()
    st = e.run()
    op.check_run_against_fixture(e, fx, st)


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_c2_epochs_bit_exact(gpu, orc, rng_mode):
    """C2 at full size: 128 walkers, 40 exchange epochs, every walker bit-identical."""
    cfg = c2_config(max_attempts=512 * 40, rng=rng_mode)
    eg, eo = pair_engines(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    sg, so = eg.run(), eo.run()
    for key in ("epoch", "done", "attempted", "accepted", "exchange_attempted",
                "exchange_accepted"):
        assert getattr(sg, key) == getattr(so, key), key
    assert sg.attempted == 128 * 512 * 40 and sg.exchange_accepted > 0
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)
    # energies against an FP64 recompute from the spins (1e-10 relative)
    spins = np.stack([eg.read_walker(k).spins for k in range(eg.walkers)])
    _, exact = eg.system.fields(spins)
    energies = np.array([eg.read_walker(k).energy for k in range(eg.walkers)])
    np.testing.assert_allclose(energies, exact, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_c2_many_windows_bit_exact(gpu, orc, rng_mode):
    """The C2 system split into 48 windows of 16 walkers (about 100 bins each, 768 walkers),
    the layout of the fastest time to accuracy: 20 exchange epochs, every walker
    bit-identical."""
    cfg = c2_config(max_attempts=512 * 20, rng=rng_mode, windows=48)
    eg, eo = pair_engines(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    sg, so = eg.run(), eo.run()
    for key in ("epoch", "done", "attempted", "accepted", "exchange_attempted",
                "exchange_accepted"):
        assert getattr(sg, key) == getattr(so, key), key
    assert eg.walkers == 768 and sg.attempted == 768 * 512 * 20 and sg.exchange_accepted > 0
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def test_exchange_schedule_bit_exact(gpu, orc):
    cfg = c2_config(max_attempts=0)
    eg, eo = pair_engines(gpu, orc, simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), cfg)
    for epoch in range(6):
        for e in (eg, eo):
            e.sample(512)
        assert eg.exchange(epoch) == eo.exchange(epoch)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def test_statistics_iteration_and_restore_on_gpu(gpu, orc):
    pos, axes, box, per = simple_cubic(1, 1, 1, 1.0, (0, 0, 1), False)
    cfg = RunConfig(emin=-1, emax=1, bin_width=0.5, seed=11,
                    wl=WlParameters(flatness=0.8, minimum_visits=1, final_factor=1e-8,
                                    check_interval_attempts=1))
    e = Engine(System(pos, axes, box, per, backend=gpu), cfg)
    e.initialize()
    snap = e.read_walker(0)
    snap.active[:] = 0
    snap.histogram[:] = 0
    snap.active[0] = snap.active[1] = 1
    snap.histogram[0], snap.histogram[1] = 80, 120
    e.write_walker(0, snap)
    st = e.statistics(0)
    assert (st.active_bins, st.covered_bins, st.minimum) == (2, 2, 80)
    assert st.mean == 100.0 and abs(st.min_over_mean - 0.8) <= 1e-14
    snap = e.read_walker(0)
    snap.active[2] = 1
    e.write_walker(0, snap)
    assert not e.ready(0)
    with pytest.raises(LogicError, match="WL iteration is not complete"):
        e.next_iteration(0)
    snap = e.read_walker(0)
    snap.info.attempted = 1000
    snap.active[2] = 0
    e.write_walker(0, snap)
    e.next_iteration(0)
    assert e.read_walker(0).factor == 0.5
    snap = e.read_walker(0)
    snap.info.stage = abi.STAGE_INVERSE_TIME
    snap.info.factor = 0.002
    e.write_walker(0, snap)
    e.sample(1)
    assert abs(e.read_walker(0).factor - 2.0 / 1001.0) <= 1e-14
    e.sample(1)
    assert abs(e.read_walker(0).factor - 3.0 / 1002.0) <= 1e-14
    bad = e.read_walker(0)
    bad.info.energy += 1.0
    with pytest.raises(RuntimeError, match="Checkpoint energy does not match"):
        e.write_walker(0, bad)


def test_extended_quotient_means_bit_exact(gpu, orc):
    """Histogram means use the reference's long double division; check awkward ratios."""
    pos, axes, box, per = simple_cubic(1, 1, 1, 1.0, (0, 0, 1), False)
    cfg = RunConfig(emin=-1, emax=1, bin_width=0.25, seed=3,
                    wl=WlParameters(check_interval_attempts=1))
    eg, eo = pair_engines(gpu, orc, (pos, axes, box, per), cfg)
    rng = np.random.default_rng(1)
    for _ in range(40):
        snap = eg.read_walker(0)
        snap.active[:] = 1
        snap.histogram[:] = rng.integers(0, 2**40, size=len(snap.histogram), dtype=np.uint64)
        snap.histogram[rng.integers(0, 8)] = rng.integers(1, 7)
        eg.write_walker(0, snap)
        eo.write_walker(0, snap)
        a, b = eg.statistics(0), eo.statistics(0)
        assert a.mean.hex() == b.mean.hex() and a.min_over_mean.hex() == b.min_over_mean.hex()


def test_c1_converges_to_exact_dos(gpu, orc):
    """C1 plain WL to ln f = 1e-6: GPU run bit-identical to the oracle's, DOS within 0.02 of
    exact enumeration after mean-offset alignment over the occupied bins."""
    cfg = c1_config()
    eg, eo = pair_engines(gpu, orc, c1_geometry(), cfg)
    sg = eg.run()
    so = eo.run()
    assert sg.converged and so.converged
    assert (sg.attempted, sg.accepted, sg.epoch) == (so.attempted, so.accepted, so.epoch)
    for k in range(2):
        assert_walkers_equal(eg, eo, k)
    lg, _, _, valid = op.fragment(eg, 0)
    exact = json.loads((GOLDEN / "exact_c1.json").read_text())["counts"]
    occupied = sorted(int(b) for b in exact)
    diff = np.array([lg[b] - np.log(exact[str(b)]) for b in occupied])
    assert np.max(np.abs(diff - diff.mean())) < 0.02


# ---- lookahead sampling (WLX_RNG_PHILOX) against the sequential oracle ----------------------
@pytest.mark.parametrize("chunks", [[1000], [1, 63, 64, 129, 7, 250], [17] * 9])
def test_lookahead_trace_bit_exact(gpu, orc, chunks):
    """Partial batches, odd starting attempts and proposal blocks crossing calls."""
    cfg = c1_config(rng=abi.RNG_PHILOX)
    eg, eo = pair_engines(gpu, orc, c1_geometry(), cfg)
    total = sum(chunks)
    for e in (eg, eo):
        e.trace(total)
        for c in chunks:
            e.sample(c)
    for k in range(2):
        assert all(np.array_equal(x, y) for x, y in zip(eg.read_trace(k), eo.read_trace(k)))
        assert_walkers_equal(eg, eo, k)


def test_lookahead_forced_nalivaiko_bit_exact(gpu, orc):
    pos, axes, box, per = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    cfg = c2_config(windows=2, walkers_per_window=3, explicit_windows=[(0, 600), (570, 1027)],
                    rng=abi.RNG_PHILOX)
    cfg.wl = WlParameters(flatness=0.6, minimum_visits=5, final_factor=1e-6,
                          check_interval_attempts=512, force_accept_after_attempts=4,
                          inverse_time_enabled=False, nalivaiko_mod=True)
    eg, eo = pair_engines(gpu, orc, (pos, axes, box, per), cfg)
    for e in (eg, eo):
        e.trace(20000)
        e.sample(20000)
        e.check()
    for k in range(eg.walkers):
        assert all(np.array_equal(x, y) for x, y in zip(eg.read_trace(k), eo.read_trace(k)))
        assert_walkers_equal(eg, eo, k)
    assert sum(eg.read_walker(k).forced_accepted for k in range(eg.walkers)) > 0


@pytest.mark.parametrize("geometry,grid", [
    ((5, 7, 3, 1.0, (0.6, 0.0, 0.8), True), (-120.0, 120.0, 0.3)),   # N=105, non-pow2 bins
    ((16, 16, 4, 1.0, (0.0, 0.0, 1.0), True), (-700.0, 900.0, 2.0)),  # N=1024
    ((16, 16, 8, 1.0, (0.0, 0.0, 1.0), True), (-1400.0, 1800.0, 4.0)),  # N=2048
    ((15, 15, 15, 1.0, (0.0, 0.6, 0.8), True), (-2600.0, 3200.0, 5.0)),  # N=3375, 53 sites/lane
])
def test_lookahead_uneven_sizes_bit_exact(gpu, orc, geometry, grid):
    cfg = RunConfig(emin=grid[0], emax=grid[1], bin_width=grid[2], windows=2,
                    walkers_per_window=2, seed=31, exchange_interval_attempts=300,
                    max_attempts=300 * 12, rng=abi.RNG_PHILOX,
                    wl=WlParameters(check_interval_attempts=600))
    eg, eo = pair_engines(gpu, orc, simple_cubic(*geometry), cfg)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def _stages_config(max_epochs):
    return RunConfig(emin=-60.0, emax=80.0, bin_width=0.5, windows=2, walkers_per_window=2,
                     seed=5, exchange_interval_attempts=54, max_attempts=54 * max_epochs,
                     rng=abi.RNG_PHILOX,
                     wl=WlParameters(final_factor=1e-3, check_interval_attempts=540))


def test_lookahead_stages_through_freeze_bit_exact(gpu, orc):
    """3x3x3 periodic REWL on the Philox stream through WL -> 1/t -> frozen (the common-
    configuration kernel), compared with the oracle after every 500 epochs."""
    eg, eo = pair_engines(gpu, orc, simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True),
                          _stages_config(4000))
    stages = set()
    for _ in range(8):
        sg, so = eg.run(500), eo.run(500)
        assert (sg.attempted, sg.accepted, sg.exchange_accepted, sg.epoch) == (
            so.attempted, so.accepted, so.exchange_accepted, so.epoch)
        for k in range(eg.walkers):
            assert_walkers_equal(eg, eo, k)
            stages.add(int(eg.read_walker(k).stage))
    assert stages == {abi.STAGE_WANG_LANDAU, abi.STAGE_INVERSE_TIME, abi.STAGE_FROZEN}


def test_lookahead_first_visits_in_inverse_time_bit_exact(gpu, orc):
    """Walkers restored into the 1/t stage with some visited bins cleared rediscover them:
    the active count, and with it every later 1/t factor of the batch, changes mid-batch."""
    eg, eo = pair_engines(gpu, orc, simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True),
                          _stages_config(0))
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
    assert any(eg.read_walker(k).active_bins > 0 for k in range(eg.walkers))


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_multi_walker_checkpoint_resume_bit_exact(gpu, orc, tmp_path, rng_mode):
    """All walkers saved after 150 C2 epochs, restored into a fresh device engine and run
    150 more: identical to the oracle's uninterrupted 300 epochs."""
    cfg = c2_config(rng=rng_mode)
    geo = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    eg, eo = pair_engines(gpu, orc, geo, cfg)
    eg.run(150)
    eg.save_checkpoint(tmp_path / "c2.npz")
    resumed = Engine(System(*geo, backend=gpu), cfg)
    resumed.initialize()
    resumed.load_checkpoint(tmp_path / "c2.npz")
    sg = resumed.run(150)
    so = eo.run(300)
    assert (sg.epoch, sg.attempted, sg.accepted, sg.exchange_attempted, sg.exchange_accepted) == (
        so.epoch, so.attempted, so.accepted, so.exchange_attempted, so.exchange_accepted)
    for k in range(eo.walkers):
        assert_walkers_equal(resumed, eo, k)


@pytest.mark.parametrize("geometry,grid", [
    ((5, 7, 3, 1.0, (0.6, 0.0, 0.8), True), (-120.0, 120.0, 0.3)),   # N=105
    ((16, 16, 4, 1.0, (0.0, 0.0, 1.0), True), (-700.0, 900.0, 2.0)),  # N=1024
    ((15, 15, 15, 1.0, (0.0, 0.6, 0.8), True), (-2600.0, 3200.0, 5.0)),  # N=3375
])
def test_xoshiro_lookahead_uneven_sizes_bit_exact(gpu, orc, geometry, grid):
    """The lookahead kernel on the reference's Xoshiro stream (words consumed one by one as
    site draws and uniforms) against the sequential oracle."""
    cfg = RunConfig(emin=grid[0], emax=grid[1], bin_width=grid[2], windows=2,
                    walkers_per_window=2, seed=31, exchange_interval_attempts=300,
                    max_attempts=300 * 12, rng=abi.RNG_XOSHIRO,
                    wl=WlParameters(check_interval_attempts=600))
    eg, eo = pair_engines(gpu, orc, simple_cubic(*geometry), cfg)
    sg, so = eg.run(), eo.run()
    assert (sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.attempted, so.accepted, so.exchange_accepted)
    for k in range(eg.walkers):
        assert_walkers_equal(eg, eo, k)


def test_xoshiro_lookahead_rejected_site_draw_bit_exact(gpu, orc):
    """bounded(N) redraws a word below (2^64 - N) mod N (rng.hpp:29-36). A state with s1 = 0
    makes the next Xoshiro256** output 0, so the walker's first site draw is rejected."""
    geo = simple_cubic(5, 7, 3, 1.0, (0.6, 0.0, 0.8), True)  # N = 105
    assert ((1 << 64) - 105) % 105 > 0
    cfg = RunConfig(emin=-120.0, emax=120.0, bin_width=0.3, windows=1, walkers_per_window=2,
                    seed=5, exchange_interval_attempts=1000, rng=abi.RNG_XOSHIRO,
                    wl=WlParameters(check_interval_attempts=10**9))
    eg, eo = pair_engines(gpu, orc, geo, cfg)
    for k in range(2):
        snap = eo.read_walker(k)
        snap.info.rng_state[1] = 0
        eg.write_walker(k, snap)
        eo.write_walker(k, snap)
    for chunk in (1, 200, 1000):
        eg.sample(chunk)
        eo.sample(chunk)
        for k in range(2):
            assert_walkers_equal(eg, eo, k)


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_warm_started_checkpoint_resume_bit_exact(gpu, orc, tmp_path, rng_mode):
    """A warm-started C2 run (per-walker 1/t clock offset) saved after 400 epochs, restored
    into a fresh engine and run 400 more equals the oracle's uninterrupted 800 epochs: the
    clock offset travels with the checkpoint, so the switch to 1/t and its factors agree."""
    cfg = c2_config(rng=rng_mode)
    cfg.wl.check_interval_attempts = 512 * 4
    geo = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    eg, eo = pair_engines(gpu, orc, geo, cfg)
    ref = json.loads((GOLDEN / "c2_reference_dos.json").read_text())
    start = np.array([float.fromhex(v) for v in ref["log_g"]])
    valid = np.array(ref["valid"], dtype=bool)
    start = np.interp(np.arange(eg.bins), np.nonzero(valid)[0], start[valid])  # finite everywhere
    for e in (eg, eo):
        e.warm_start(start, 1e-3)
    eg.run(400)
    offsets = [eg.read_walker(k).info.clock_offset for k in range(eg.walkers)]
    assert all(o > 0 for o in offsets)
    eg.save_checkpoint(tmp_path / "warm.npz")
    resumed = Engine(System(*geo, backend=gpu), cfg)
    resumed.initialize()
    resumed.load_checkpoint(tmp_path / "warm.npz")
    assert [resumed.read_walker(k).info.clock_offset for k in range(eg.walkers)] == offsets
    sg, so = resumed.run(400), eo.run(800)
    assert (sg.epoch, sg.attempted, sg.accepted, sg.exchange_accepted) == (
        so.epoch, so.attempted, so.accepted, so.exchange_accepted)
    stages = {resumed.read_walker(k).info.stage for k in range(eg.walkers)}
    assert abi.STAGE_INVERSE_TIME in stages, stages  # the 1/t switch happened after the resume
    for k in range(eo.walkers):
        assert_walkers_equal(resumed, eo, k)
