"""Host DOS assembly (replica_exchange_wang_landau_amd.analysis) and the multi-walker
checkpoint, on the CPU:
- summarize_window must reproduce the reference's own fragments (golden runs);
- stitch must reproduce the compiled reference's stitch_dos bit for bit;
- a checkpointed and resumed run must equal an uninterrupted one.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import test_oracle_pinning as op
from harness import REF_LIB, c2_config
from replica_exchange_wang_landau_amd import abi, analysis
from replica_exchange_wang_landau_amd.engine import Engine, System, simple_cubic, _ptr

needs_ref = pytest.mark.skipif(not REF_LIB.exists(), reason="compiled reference not built")


@pytest.mark.parametrize("name", ["c1_prefix", "rewl_2x2x2_two_windows", "c2_prefix_20mcs"])
def test_summarize_matches_reference_fragments(orc, name):
    fx = next(r for r in op.load("runs.json") if r["name"] == name)
    nx, ny, nz, periodic = fx["geometry"]
    e = Engine(System(*simple_cubic(nx, ny, nz, 1.0, (0, 0, 1), periodic), backend=orc),
               op.config_from_fixture(fx))
    e.initialize()
    e.run()
    per = e.config.walkers_per_window
    for w in range(e.config.windows):
        f = analysis.summarize_window([e.read_walker(w * per + p) for p in range(per)], e.bins)
        assert f.log_g.tobytes() == op.fromhex(fx["fragment_log_g"][w]).tobytes()
        assert f.standard_error.tobytes() == op.fromhex(fx["fragment_error"][w]).tobytes()
        assert f.histogram.tolist() == fx["fragment_histogram"][w]
        assert f.valid.tolist() == fx["fragment_valid"][w]


def _ref_stitch(grid, frags, complete=False, spins=0):
    lib = ctypes.CDLL(str(REF_LIB))
    bins = len(frags[0].log_g)
    be = np.array([[f.begin, f.end] for f in frags], dtype=np.uint64).ravel()
    lg = np.concatenate([f.log_g for f in frags])
    err = np.concatenate([f.standard_error for f in frags])
    hist = np.concatenate([f.histogram for f in frags]).astype(np.uint64)
    valid = np.concatenate([f.valid for f in frags]).astype(np.uint8)
    out = np.empty(bins)
    out_valid = np.empty(bins, dtype=np.uint8)
    status = lib.ref_stitch(ctypes.c_double(grid[0]), ctypes.c_double(grid[1]),
                            ctypes.c_double(grid[2]), ctypes.c_uint64(bins),
                            ctypes.c_uint64(len(frags)), _ptr(be, ctypes.c_uint64),
                            _ptr(lg, ctypes.c_double), _ptr(err, ctypes.c_double),
                            _ptr(hist, ctypes.c_uint64), _ptr(valid, ctypes.c_uint8),
                            ctypes.c_int(complete), ctypes.c_uint64(spins),
                            _ptr(out, ctypes.c_double), _ptr(out_valid, ctypes.c_uint8))
    assert status == 0
    return out, out_valid


@needs_ref
def test_stitch_matches_reference_on_engine_fragments(orc):
    """C2 geometry, 8 windows, a short REWL run: stitched log g equals stitch_dos's."""
    cfg = c2_config(max_attempts=512 * 300)
    e = Engine(System(*simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True), backend=orc), cfg)
    e.initialize()
    e.run()
    per = cfg.walkers_per_window
    frags = [analysis.summarize_window([e.read_walker(w * per + p) for p in range(per)], e.bins,
                                       allow_empty=True) for w in range(cfg.windows)]
    frags = [f for f in frags if f.valid.any()]
    ours, ours_valid = analysis.stitch(e.bins, cfg.bin_width, frags)
    ref, ref_valid = _ref_stitch((cfg.emin, cfg.emax, cfg.bin_width), frags)
    assert ours_valid.tolist() == ref_valid.tolist()
    assert ours.tobytes() == ref.tobytes()
    both = analysis.engine_dos(e)
    assert both[0].tobytes() == ours.tobytes()


@needs_ref
def test_stitch_matches_reference_on_synthetic_fragments():
    """Random smooth curves on overlapping windows with holes: same joins and shifts."""
    rng = np.random.default_rng(2)
    bins = 200
    grid = (-50.0, 50.0, 0.5)
    for trial in range(20):
        starts = sorted(rng.choice(np.arange(0, 150), size=4, replace=False))
        frags = []
        base = np.cumsum(rng.normal(0.0, 0.3, size=bins))
        for k, b in enumerate(starts):
            e = min(bins, (starts[k + 1] + 30) if k + 1 < len(starts) else bins)
            valid = np.zeros(bins, dtype=np.uint8)
            valid[b:e] = rng.random(e - b) > 0.08
            lg = np.full(bins, np.nan)
            lg[b:e] = base[b:e] + rng.normal(0.0, 0.02, size=e - b) + rng.normal(0.0, 5.0)
            lg[valid == 0] = np.nan
            frags.append(analysis.Fragment(int(b), int(e), lg, np.zeros(bins, dtype=np.uint64),
                                           np.full(bins, np.nan), valid))
        # the windows must overlap the running coverage
        if any(frags[k + 1].begin >= max(f.end for f in frags[:k + 1])
               for k in range(len(frags) - 1)):
            continue
        try:
            ours = analysis.stitch(bins, grid[2], frags)
        except RuntimeError:
            with pytest.raises(AssertionError):
                _ref_stitch(grid, frags)
            continue
        ref = _ref_stitch(grid, frags)
        assert ours[1].tolist() == ref[1].tolist(), trial
        assert ours[0].tobytes() == ref[0].tobytes(), trial


def test_max_deviation_aligns_mean_offset():
    ref = np.array([0.0, 1.0, 2.0, np.nan, 4.0])
    est = ref + 7.5
    est[2] += 0.03
    valid = np.array([1, 1, 1, 0, 1], dtype=np.uint8)
    # offset 7.5 + 0.03/4 removed; the largest residual sits at bin 2
    assert abs(analysis.max_deviation(est, valid, ref, valid) - 0.0225) < 1e-12
    assert analysis.max_deviation(est, np.zeros(5, dtype=np.uint8), ref, valid) == float("inf")


@pytest.mark.parametrize("rng_mode", [abi.RNG_XOSHIRO, abi.RNG_PHILOX])
def test_checkpoint_resume_equals_uninterrupted_run(orc, tmp_path, rng_mode):
    """save_checkpoint after 40 epochs + load into a fresh engine + 60 epochs == 100 epochs."""
    geo = simple_cubic(4, 4, 2, 1.0, (0, 0, 1), True)
    from replica_exchange_wang_landau_amd import RunConfig, WlParameters
    cfg = RunConfig(emin=-60.0, emax=60.0, bin_width=1.0, windows=3, walkers_per_window=2,
                    seed=21, exchange_interval_attempts=32, rng=rng_mode,
                    wl=WlParameters(check_interval_attempts=320))
    a = Engine(System(*geo, backend=orc), cfg)
    a.initialize()
    a.run(100)
    b = Engine(System(*geo, backend=orc), cfg)
    b.initialize()
    b.run(40)
    b.save_checkpoint(tmp_path / "rewl.npz")
    c = Engine(System(*geo, backend=orc), cfg)
    c.initialize()
    c.load_checkpoint(tmp_path / "rewl.npz")
    c.run(60)
    sa, sc = a.status(), c.status()
    assert (sa.epoch, sa.attempted, sa.accepted, sa.exchange_attempted, sa.exchange_accepted) == (
        sc.epoch, sc.attempted, sc.accepted, sc.exchange_attempted, sc.exchange_accepted)
    from harness import assert_walkers_equal
    for k in range(a.walkers):
        assert_walkers_equal(a, c, k)


def test_average_dos_removes_offsets_and_averages_noise():
    from replica_exchange_wang_landau_amd.analysis import average_dos, max_deviation

    rng = np.random.default_rng(5)
    bins = 40
    truth = -0.01 * (np.arange(bins) - 17.0) ** 2
    curves = []
    for r in range(16):
        valid = np.ones(bins, dtype=np.uint8)
        valid[: r % 3] = 0                   # ragged low edges
        lg = truth + 3.0 * r + rng.normal(0.0, 0.1, bins)
        lg[valid == 0] = -np.inf
        curves.append((lg, valid))
    avg, common = average_dos(curves)
    assert common[:2].sum() == 0 and common[2:].all()
    assert avg[common != 0].max() == 0.0
    # offsets cancel exactly without noise; with it the error falls like 1/sqrt(16)
    exact, _ = average_dos([(truth + 7.0, np.ones(bins, np.uint8)), (truth - 2.0,
                                                                       np.ones(bins, np.uint8))])
    assert max_deviation(exact, np.ones(bins), truth, np.ones(bins)) < 1e-12
    single = max_deviation(curves[5][0], curves[5][1], truth, np.ones(bins))
    assert max_deviation(avg, common, truth, np.ones(bins)) < 0.5 * single
    with pytest.raises(RuntimeError, match="share no valid bin"):
        average_dos([(truth, np.r_[np.ones(20), np.zeros(20)]),
                     (truth, np.r_[np.zeros(20), np.ones(20)])])
