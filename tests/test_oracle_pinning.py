"""Pins the C oracle before it is trusted as the checker of the HIP engine.

Checks, per oracle function: the reference's own closed-form tests (test_main.cpp), golden
vectors written by the compiled reference (tests/golden/), published known answers
(splitmix64, numpy's Philox4x64-10), independent recomputation (numpy J), exact enumeration,
and live bit equality with the compiled reference where it is built (oracle/_ref).
"""
from __future__ import annotations

import ctypes
import math
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from harness import (GOLDEN, Reference, assert_walkers_equal, c1_config, c1_geometry, c2_config,
                     info_dict)
from replica_exchange_wang_landau_amd.engine import (Engine, InvalidArgument, LogicError, RunConfig,
                                                     System, WlParameters, _ptr, simple_cubic)


def load(name):
    return json.loads((GOLDEN / name).read_text())


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fromhex(values):
    return np.array([float.fromhex(v) for v in values])


# ---- RNG ----------------------------------------------------------------------------------
def test_splitmix_known_answer(orc):
    # splitmix64 seeded with 1234567 (Vigna's reference implementation) yields these words;
    # the stream-0 seeding XORs the seed with 0xd2b74407b1ce6e93 first (rng.hpp:52-55).
    state = (ctypes.c_uint64 * 4)()
    seed = 1234567 ^ 0xD2B74407B1CE6E93
    orc.lib.orc_xoshiro_seed(ctypes.c_uint64(seed), ctypes.c_uint64(0), state)
    assert list(state) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                           4593380528125082431]


def test_xoshiro_golden(orc):
    fx = load("rng.json")
    for case in fx["words"]:
        state = (ctypes.c_uint64 * 4)()
        orc.lib.orc_xoshiro_seed(ctypes.c_uint64(case["seed"]), ctypes.c_uint64(case["stream"]),
                                 state)
        words = [orc.lib.orc_xoshiro_next(state) for _ in range(8)]
        assert words == case["words"]
        assert list(state) == case["state_after"]
    for case in fx["bounded"]:
        state = (ctypes.c_uint64 * 4)()
        orc.lib.orc_xoshiro_seed(ctypes.c_uint64(case["seed"]), ctypes.c_uint64(case["stream"]),
                                 state)
        vals = [orc.lib.orc_xoshiro_bounded(state, ctypes.c_uint64(case["bound"]))
                for _ in range(32)]
        assert vals == case["values"]


def test_xoshiro_matches_compiled_reference(orc, ref):
    for seed, stream in [(3, 0), (2**64 - 1, 99), (123456789, 5)]:
        words, state = ref.xoshiro(seed, stream, 16)
        st = (ctypes.c_uint64 * 4)()
        orc.lib.orc_xoshiro_seed(ctypes.c_uint64(seed), ctypes.c_uint64(stream), st)
        assert [orc.lib.orc_xoshiro_next(st) for _ in range(16)] == [int(w) for w in words]


def test_philox_matches_numpy(orc):
    for key, counter in [((1, 2), (0, 0, 0, 0)), ((2**64 - 1, 5), (7, 1, 2**64 - 1, 3)),
                         ((1234, 0x57414C4B), (99, 0x57414C4B, 0, 0))]:
        g = np.random.Philox(key=np.array(key, dtype=np.uint64),
                             counter=np.array(counter, dtype=np.uint64))
        expected = [int(x) for x in g.random_raw(4)]
        # numpy increments the counter before its first block
        ctr = list(counter)
        ctr[0] = (ctr[0] + 1) % 2**64
        if ctr[0] == 0:
            ctr[1] = (ctr[1] + 1) % 2**64
        c = (ctypes.c_uint64 * 4)(*ctr)
        k = (ctypes.c_uint64 * 2)(*key)
        out = (ctypes.c_uint64 * 4)()
        orc.lib.orc_philox4x64(c, k, out)
        assert list(out) == expected


# ---- couplings ----------------------------------------------------------------------------
def pair(orc, d, ai, aj, scale=1.0):
    out = ctypes.c_double()
    arr = [np.array(v, dtype=np.float64) for v in (d, ai, aj)]
    st = orc.lib.orc_dipolar_pair(*[_ptr(a, ctypes.c_double) for a in arr],
                                  ctypes.c_double(scale), ctypes.byref(out))
    return st, out.value


def test_pair_formula_closed_forms(orc):
    # test_main.cpp:24-28
    assert abs(pair(orc, (1, 0, 0), (1, 0, 0), (1, 0, 0))[1] - (-2.0)) <= 1e-14
    assert abs(pair(orc, (1, 0, 0), (0, 0, 1), (0, 0, 1))[1] - 1.0) <= 1e-14
    assert abs(pair(orc, (2, 0, 0), (0, 0, 1), (0, 0, 1))[1] - 0.125) <= 1e-14
    assert pair(orc, (0, 0, 0), (0, 0, 1), (0, 0, 1))[0] == 1  # coincident sites


def test_minimum_image_cases(orc):
    # test_main.cpp:30-38
    def mi(d, box):
        out = np.empty(3)
        orc.lib.orc_minimum_image(_ptr(np.array(d, float), ctypes.c_double),
                                  _ptr(np.array(box, float), ctypes.c_double), 1,
                                  _ptr(out, ctypes.c_double))
        return out

    np.testing.assert_allclose(mi((6, -5, 3.1), (10, 8, 6)), (-4, 3, -2.9), atol=1e-14)
    np.testing.assert_allclose(mi((6, -5, 31), (10, 8, 0)), (-4, 3, 31), atol=1e-14)
    # round-half-even of nearbyint: 5/10 = 0.5 rounds to 0, 15/10 = 1.5 rounds to 2
    np.testing.assert_array_equal(mi((5, 15, -5), (10, 10, 10)), (5, -5, -5))


def numpy_dense(positions, axes, box, periodic, scale=1.0):
    """Independent FP64 recompute of J_ij = D[e_i.e_j - 3(e_i.r)(e_j.r)]/r^3 (README.md:10)."""
    d = positions[None, :, :] - positions[:, None, :]
    if periodic:
        for k in range(3):
            if box[k] > 0:
                d[..., k] -= box[k] * np.rint(d[..., k] / box[k])
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, 1.0)
    rhat = d / r[..., None]
    ij = axes @ axes.T
    ir = np.einsum("ik,ijk->ij", axes, rhat)
    jr = np.einsum("jk,ijk->ij", axes, rhat)
    J = scale * (ij - 3 * ir * jr) / r**3
    np.fill_diagonal(J, 0.0)
    return J


def test_dense_couplings_golden_and_recompute(orc):
    fx = load("couplings.json")
    s = System(*c1_geometry(), backend=orc)
    J = s.dense()
    assert J.tobytes() == fromhex(fx["c1_dense"]).tobytes()
    pos, axes, box, per = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    s2 = System(pos, axes, box, per, backend=orc)
    J2 = s2.dense()
    assert digest(J2) == fx["c2_dense_sha256"]
    assert np.array_equal(J2, J2.T)  # bitwise symmetric
    np.testing.assert_allclose(J2, numpy_dense(pos, axes, box, per), rtol=1e-13, atol=1e-15)


def test_csr_couplings_golden(orc):
    fx = load("couplings.json")["random_slab"]
    pos = np.array([[float.fromhex(v) for v in p] for p in fx["positions"]])
    ax = np.array([[float.fromhex(v) for v in a] for a in fx["axes"]])
    box = np.array(fx["box"])
    dense = System(pos, ax, box, True, backend=orc)
    csr = System(pos, ax, box, True, cutoff=2.0, backend=orc)
    assert digest(dense.dense()) == fx["dense_sha256"]
    assert digest(csr.dense()) == fx["csr2_dense_view_sha256"]
    offsets, cols, vals = csr.csr()
    assert csr.backend_name == "csr" and offsets[-1] == len(cols)
    for i in range(csr.n):  # rows ascending, no diagonal
        row = cols[offsets[i]:offsets[i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0) and i not in row
    np.testing.assert_allclose(dense.dense(), numpy_dense(pos, ax, box, True), rtol=1e-12,
                               atol=1e-13)


def test_dense_equals_csr_at_large_cutoff(orc):
    # test_main.cpp:139-143
    pos, axes, box, per = simple_cubic(3, 2, 1, 1.0, (0, 0, 1), False)
    d = System(pos, axes, box, per, coupling_scale=0.7, backend=orc)
    c = System(pos, axes, box, per, coupling_scale=0.7, cutoff=100.0, backend=orc)
    assert d.dense().tobytes() == c.dense().tobytes()


def test_incremental_fields_match_recompute(orc):
    # test_main.cpp:144-153: dE and fields after each flip equal a full recompute
    pos, axes, box, per = simple_cubic(3, 2, 1, 1.0, (0, 0, 1), False)
    s = System(pos, axes, box, per, coupling_scale=0.7, backend=orc)
    J = s.dense()
    spins = np.array([1, -1, 1, 1, -1, -1], dtype=np.int8)
    fields, e0 = s.fields(spins)
    for i in range(6):
        delta = -2.0 * spins[i] * fields[i]
        flipped = spins.copy()
        flipped[i] = -spins[i]
        f1, e1 = s.fields(flipped)
        assert abs(e0 + delta - e1) <= 1e-12 * max(1, abs(e1))
        np.testing.assert_allclose(fields + (-2.0 * spins[i]) * J[:, i], f1, atol=1e-12)


def test_geometry_validation_messages(orc):
    pos = np.zeros((2, 3))
    axes = np.array([[0, 0, 1.0], [0, 0, 1.0]])
    with pytest.raises(InvalidArgument, match="Coincident dipole sites"):
        System(pos, axes, [2, 2, 2], True, backend=orc)
    with pytest.raises(InvalidArgument, match="Ising axes must be normalized"):
        System([[0, 0, 0], [1, 0, 0]], [[0, 0, 2.0], [0, 0, 1]], [2, 2, 2], True, backend=orc)
    with pytest.raises(InvalidArgument, match="At least one periodic box length"):
        System([[0, 0, 0], [1, 0, 0]], axes, [0, 0, 0], True, backend=orc)


# ---- grid and windows -----------------------------------------------------------------------
def test_grid_and_windows(orc):
    # test_main.cpp:156-162
    idx = lambda e: orc.lib.orc_grid_index(ctypes.c_double(-2), ctypes.c_double(2),
                                           ctypes.c_double(0.5), ctypes.c_double(e))
    assert idx(-2) == 0 and idx(2) == 7 and idx(2.1) == -1 and idx(float("nan")) == -1
    be = np.empty(8, dtype=np.uint64)
    assert orc.lib.orc_partition_windows(ctypes.c_uint64(100), ctypes.c_uint64(4),
                                         ctypes.c_double(0.75), _ptr(be, ctypes.c_uint64)) == 0
    assert be[0] == 0 and be[-1] == 100
    assert all(be[2 * w] < be[2 * w - 1] for w in range(1, 4))
    # the C2 layout: 1255 bins, 8 windows at overlap 0.75
    be = np.empty(16, dtype=np.uint64)
    orc.lib.orc_partition_windows(ctypes.c_uint64(1255), ctypes.c_uint64(8), ctypes.c_double(0.75),
                                  _ptr(be, ctypes.c_uint64))
    assert list(be[0::2]) == [0, 114, 228, 342, 456, 570, 684, 798]
    assert be[1] - be[0] == 457


# ---- walkers --------------------------------------------------------------------------------
def test_walker_trajectory_golden(orc):
    fx = load("walkers.json")["c1_walker0"]
    cfg = c1_config()
    s = System(*c1_geometry(), backend=orc)
    e = Engine(s, cfg)
    e.initialize()
    e.trace(fx["attempts"])
    e.sample(fx["attempts"])
    sites, bins, flags = e.read_trace(0)
    assert np.packbits((flags & 1).astype(np.uint8)).tobytes().hex() == fx["accepted_bits"]
    assert digest(bins.astype(np.int64)) == fx["bins_sha256"]
    check_state(e.read_walker(0), fx["state"])


def check_state(w, st):
    assert w.energy.hex() == st["energy"] and w.factor.hex() == st["factor"]
    for key in ("attempted", "accepted", "forced_accepted", "last_accepted_attempt", "stage"):
        assert getattr(w, key) == st[key], key
    assert list(w.rng_state) == st["rng_state"]
    assert digest(w.spins) == st["spins_sha256"]
    assert digest(w.fields) == st["fields_sha256"]
    assert digest(w.log_g) == st["log_g_sha256"]
    assert digest(w.histogram) == st["histogram_sha256"]
    assert digest(w.active) == st["active_sha256"]


def forced_nalivaiko_engine(backend):
    """Walker id 5 in window [570, 1027) of the C2 lattice: forced acceptance after 4
    rejected attempts, traditional WL (no 1/t), Nalivaiko refinement mask."""
    pos, axes, box, per = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    s = System(pos, axes, box, per, backend=backend)
    cfg = c2_config(windows=2, walkers_per_window=3, explicit_windows=[(0, 600), (570, 1027)])
    cfg.wl = WlParameters(flatness=0.6, minimum_visits=5, final_factor=1e-6,
                          check_interval_attempts=512, force_accept_after_attempts=4,
                          inverse_time_enabled=False, nalivaiko_mod=True)
    e = Engine(s, cfg)
    e.initialize()
    return e


def test_forced_nalivaiko_traditional_walker_golden(orc):
    fx = load("walkers.json")["c2_walker5_forced_nalivaiko"]
    e = forced_nalivaiko_engine(orc)
    e.trace(fx["attempts"])
    e.sample(fx["attempts"])
    _, bins, flags = e.read_trace(5)
    assert digest(bins.astype(np.int64)) == fx["bins_sha256"]
    assert int((flags & 1).sum()) == fx["accepted"]
    w = e.read_walker(5)
    assert w.forced_accepted > 0
    check_state(w, fx["state"])


def test_exact_enumeration_c1(orc):
    """The C1 DOS by brute force over 2^8 states (analysis.cpp:162-181) = golden counts."""
    fx = load("exact_c1.json")
    s = System(*c1_geometry(), backend=orc)
    states = ((np.arange(256)[:, None] >> np.arange(8)) & 1).astype(np.int8) * 2 - 1
    _, energies = s.fields(states)
    counts = {}
    for e in energies:
        b = orc.lib.orc_grid_index(ctypes.c_double(-20), ctypes.c_double(20),
                                   ctypes.c_double(0.25), ctypes.c_double(e))
        counts[str(b)] = counts.get(str(b), 0) + 1
    assert counts == fx["counts"]


def fragment(engine: Engine, window: int, allow_empty=False):
    """summarize_impl (parallel.cpp:23-62) over one window's walkers, in the reference order."""
    per = engine.config.walkers_per_window
    snaps = [engine.read_walker(window * per + p) for p in range(per)]
    begin, end = snaps[0].window_begin, snaps[0].window_end
    bins = engine.bins
    log_g = np.full(bins, np.nan)
    err = np.full(bins, np.nan)
    hist = np.zeros(bins, dtype=np.uint64)
    valid = np.zeros(bins, dtype=np.uint8)
    for i in range(begin, end):
        v = True
        for s in snaps:
            v = v and s.active[i] != 0
            hist[i] += s.histogram[i]
        valid[i] = 1 if v else 0
    ref_bins = [i for i in range(begin, end) if valid[i]]
    if not ref_bins:
        if allow_empty:
            return log_g, err, hist, valid
        raise RuntimeError("Walkers in an energy window have no common active bin")
    r = ref_bins[0]
    for i in range(begin, end):
        if not valid[i]:
            continue
        acc = 0.0
        for s in snaps:
            acc += float(s.log_g[i]) - float(s.log_g[r])
        acc /= float(len(snaps))
        log_g[i] = acc
        if len(snaps) > 1:
            sq = 0.0
            for s in snaps:
                dlt = (float(s.log_g[i]) - float(s.log_g[r])) - acc
                sq += dlt * dlt
            err[i] = math.sqrt(sq / (float(len(snaps)) * float(len(snaps) - 1)))
    return log_g, err, hist, valid


def config_from_fixture(fx) -> RunConfig:
    c = fx["config"]
    return RunConfig(emin=c["emin"], emax=c["emax"], bin_width=c["bin_width"], windows=c["windows"],
                     walkers_per_window=c["walkers_per_window"], seed=c["seed"],
                     exchange_interval_attempts=c["exchange_interval_attempts"],
                     max_attempts=c["max_attempts"], wl=WlParameters(**c["wl"]))


def check_run_against_fixture(engine: Engine, fx, status):
    att, acc, forced, xatt, xacc = fx["counters"]
    assert (status.attempted, status.accepted, status.forced_accepted) == (att, acc, forced)
    assert (status.exchange_attempted, status.exchange_accepted) == (xatt, xacc)
    assert bool(status.converged) == fx["converged"]
    for k in range(engine.walkers):
        w = engine.read_walker(k)
        wi, wd = fx["walker_integers"][k], fx["walker_doubles"][k]
        assert w.attempted == wi[2] and w.accepted == wi[3] and w.forced_accepted == wi[4]
        assert w.attempted - w.last_accepted_attempt == wi[5]
        assert w.energy.hex() == wd[0] and w.factor.hex() == wd[1], k
    for window in range(engine.config.windows):
        lg, err, hist, valid = fragment(engine, window)
        assert lg.tobytes() == fromhex(fx["fragment_log_g"][window]).tobytes()
        assert err.tobytes() == fromhex(fx["fragment_error"][window]).tobytes()
        assert hist.tolist() == fx["fragment_histogram"][window]
        assert valid.tolist() == fx["fragment_valid"][window]


@pytest.mark.parametrize("name", ["c1_prefix", "exact_regression_2x2x1",
                                  "rewl_2x2x2_two_windows", "c2_prefix_20mcs"])
def test_run_rewl_golden(orc, name):
    fx = next(r for r in load("runs.json") if r["name"] == name)
    nx, ny, nz, periodic = fx["geometry"]
    s = System(*simple_cubic(nx, ny, nz, 1.0, (0, 0, 1), periodic), backend=orc)
    e = Engine(s, config_from_fixture(fx))
    e.initialize()
    st = e.run()
    check_run_against_fixture(e, fx, st)


def test_exact_regression_rmse(orc):
    """test_main.cpp:364-391: REWL on 2x2x1 open converges to the exact DOS (RMSE <= 0.05)."""
    fx = next(r for r in load("runs.json") if r["name"] == "exact_regression_2x2x1")
    s = System(*simple_cubic(2, 2, 1, 1.0, (0, 0, 1), False), backend=orc)
    e = Engine(s, config_from_fixture(fx))
    e.initialize()
    assert e.run().converged
    lg, _, _, valid = fragment(e, 0)
    states = ((np.arange(16)[:, None] >> np.arange(4)) & 1).astype(np.int8) * 2 - 1
    _, energies = s.fields(states)
    counts = {}
    for en in energies:
        b = int(np.floor((en + 20) / 0.25))
        counts[b] = counts.get(b, 0) + 1
    occupied = sorted(counts)
    assert all(valid[b] for b in occupied)
    diff = np.array([lg[b] - np.log(counts[b]) for b in occupied])
    assert np.sqrt(np.mean((diff - diff.mean()) ** 2)) <= 0.05


def test_walker_matches_compiled_reference_live(orc, ref):
    """Per-attempt parity with the compiled reference walker on a window of the C2 lattice."""
    pos, axes, box, per = simple_cubic(8, 8, 8, 1.0, (0, 0, 1), True)
    rs = ref.system(pos, axes, box, per)
    cfg = c2_config(windows=1, walkers_per_window=1, explicit_windows=None)
    cfg.wl.check_interval_attempts = 512
    rw = rs.walker(0, cfg, 0, cfg.bins)
    acc, bins = rw.attempts(30000)
    s = System(pos, axes, box, per, backend=orc)
    e = Engine(s, cfg)
    e.initialize()
    e.trace(30000)
    e.sample(30000)
    _, b, f = e.read_trace(0)
    assert np.array_equal(acc, f & 1) and np.array_equal(bins, b.astype(np.int64))
    r = rw.read()
    w = e.read_walker(0)
    assert r["log_g"].tobytes() == w.log_g.tobytes()
    assert r["fields"].tobytes() == w.fields.tobytes()
    assert list(r["info"].rng_state) == list(w.rng_state)


def test_walker_statistics_and_iteration(orc):
    """test_main.cpp:534-625 (histogram statistics, 1/t factors 2/1001 and 3/1002)."""
    pos, axes, box, per = simple_cubic(1, 1, 1, 1.0, (0, 0, 1), False)
    s = System(pos, axes, box, per, backend=orc)
    cfg = RunConfig(emin=-1, emax=1, bin_width=0.5, seed=11,
                    wl=WlParameters(flatness=0.8, minimum_visits=1, final_factor=1e-8,
                                    check_interval_attempts=1))
    e = Engine(s, cfg)
    e.initialize()
    snap = e.read_walker(0)
    snap.active[:] = 0
    snap.histogram[:] = 0
    snap.active[0] = snap.active[1] = 1
    snap.histogram[0], snap.histogram[1] = 80, 120
    e.write_walker(0, snap)
    st = e.statistics(0)
    assert (st.active_bins, st.covered_bins, st.minimum) == (2, 2, 80)
    assert st.mean == 100.0 and abs(st.min_over_mean - 0.8) <= 1e-14 and st.coverage == 1.0
    snap = e.read_walker(0)
    snap.active[2] = 1
    e.write_walker(0, snap)
    st = e.statistics(0)
    assert st.active_bins == 3 and st.minimum == 0 and st.covered_bins == 2
    assert abs(st.mean - 200.0 / 3.0) <= 1e-14
    assert not e.ready(0)
    with pytest.raises(LogicError, match="WL iteration is not complete"):
        e.next_iteration(0)
    snap = e.read_walker(0)
    snap.info.attempted = 1000
    snap.active[2] = 0
    e.write_walker(0, snap)
    assert e.ready(0)
    e.next_iteration(0)
    w = e.read_walker(0)
    assert w.factor == 0.5 and e.statistics(0).minimum == 0
    snap = e.read_walker(0)
    snap.info.stage = 1
    snap.info.factor = 0.002
    e.write_walker(0, snap)
    e.sample(1)
    assert abs(e.read_walker(0).factor - 2.0 / 1001.0) <= 1e-14
    e.sample(1)
    assert abs(e.read_walker(0).factor - 3.0 / 1002.0) <= 1e-14


def test_restore_validation_messages(orc):
    s = System(*simple_cubic(2, 2, 1), backend=orc)
    e = Engine(s, RunConfig(emin=-20, emax=20, bin_width=0.25,
                            wl=WlParameters(flatness=0.1, minimum_visits=1, final_factor=0.1,
                                            check_interval_attempts=10)))
    e.initialize()
    snap = e.read_walker(0)
    bad = e.read_walker(0)
    bad.info.last_accepted_attempt = bad.info.attempted + 1
    with pytest.raises(InvalidArgument, match="invalid last accepted attempt"):
        e.write_walker(0, bad)
    bad = e.read_walker(0)
    bad.info.energy += 1.0
    with pytest.raises(RuntimeError, match="Checkpoint energy does not match"):
        e.write_walker(0, bad)
    e.write_walker(0, snap)


def init_failure_engine(backend):
    pos, axes, box, per = simple_cubic(2, 1, 1, 1.0, (1, 0, 0), False)
    s = System(pos, axes, box, per, backend=backend)
    cfg = RunConfig(emin=-2.0, emax=2.0, bin_width=0.1, windows=1, walkers_per_window=1,
                    seed=123, explicit_windows=[(20, 30)],
                    wl=WlParameters(flatness=0.8, minimum_visits=1, final_factor=1e-8,
                                    check_interval_attempts=10, initialization_max_attempts=10,
                                    initialization_stall_attempts_per_spin=1))
    return Engine(s, cfg)


INIT_FAILURE_FRAGMENTS = [
    "Walker initialization failed: window 0, local walker 0: ", "target band",
    "energy window bins [20, 30)", "energy bounds [0, 1)", "after 10 attempts",
    "reached energy range", "closest energy", "adaptive search used 3 temperature increases",
    "and 1 random restarts"]


def test_initialization_failure_message(orc):
    """test_main.cpp:195-216: the failure text names the window, bounds and search activity."""
    e = init_failure_engine(orc)
    with pytest.raises(RuntimeError) as info:
        e.initialize()
    for fragment_text in INIT_FAILURE_FRAGMENTS:
        assert fragment_text in str(info.value)


def test_initialization_failure_matches_reference(orc, ref):
    pos, axes, box, per = simple_cubic(2, 1, 1, 1.0, (1, 0, 0), False)
    rs = ref.system(pos, axes, box, per)
    e = init_failure_engine(orc)
    with pytest.raises(RuntimeError) as info:
        e.initialize()
    with pytest.raises(RuntimeError) as ref_info:
        rs.walker(0, e.config, 20, 30)
    prefix = "Walker initialization failed: window 0, local walker 0: "
    assert str(info.value) == prefix + str(ref_info.value)


def test_target_metropolis_initialization(orc):
    """test_main.cpp:164-193: reproducible start inside the central band, one H entry."""
    pos, axes, box, per = simple_cubic(2, 1, 1, 1.0, (1, 0, 0), False)
    cfg = RunConfig(emin=-2.5, emax=3.5, bin_width=1.0, windows=1, walkers_per_window=2,
                    seed=123, wl=WlParameters(flatness=0.8, minimum_visits=1, final_factor=1e-8,
                                              check_interval_attempts=10,
                                              initialization_max_attempts=100))
    a = Engine(System(pos, axes, box, per, backend=orc), cfg)
    b = Engine(System(pos, axes, box, per, backend=orc), cfg)
    a.initialize()
    b.initialize()
    target, half = 0.5, 0.5 * 0.5 * 6.0
    for k in range(2):
        w = a.read_walker(k)
        assert abs(w.energy - target) <= half
        assert int(w.histogram.sum()) == 1 and w.attempted == 0 and w.accepted == 0
        assert_walkers_equal(a, b, k)


def assert_sampling_equal(a, b):
    assert a["squared_displacement"].tobytes() == b["squared_displacement"].tobytes()
    assert np.array_equal(a["displacement_samples"], b["displacement_samples"])
    assert a["round_trips"] == b["round_trips"]
    for x, y in zip(a["representatives"], b["representatives"]):
        assert (x is None) == (y is None)
        if x is not None:
            assert x[0].hex() == y[0].hex() and np.array_equal(x[1], y[1])


@pytest.mark.parametrize("window", [(0, 160), (40, 100)])
def test_window_statistics_match_compiled_reference(orc, ref, window):
    """collect_window_statistics (wang_landau.cpp:104-118, 272-282, 317-342): squared
    displacements, samples, round trips and representatives after initialisation and 40000
    attempts, against the compiled reference walker. A narrow window makes round trips
    frequent."""
    pos, axes, box, per = simple_cubic(2, 2, 2, 1.0, (0, 0, 1), False)
    cfg = c1_config(explicit_windows=[window])
    cfg.wl.collect_window_statistics = True
    cfg.wl.round_trip_margin_fraction = 0.15
    rw = ref.system(pos, axes, box, per).walker(0, cfg, *window)
    e = Engine(System(pos, axes, box, per, backend=orc), cfg)
    e.initialize()
    assert_sampling_equal(e.read_sampling(0), rw.sampling())
    rw.attempts(40000)
    e.sample(40000)
    a, b = e.read_sampling(0), rw.sampling()
    assert_sampling_equal(a, b)
    assert sum(r is not None for r in a["representatives"]) >= 5
    if window == (40, 100):  # C1 reaches both margins of this window, not of the full grid
        assert a["round_trips"] > 0


def test_rewl_window_statistics_match_compiled_reference(orc, ref):
    """run_rewl with collect_window_statistics on 3x3x3 periodic, 3 windows x 2 walkers,
    exchanges every 2 MCS: an accepted swap_configuration updates both walkers' round-trip
    state and representatives (wang_landau.cpp:483-495). Per-window sums of squared
    displacements and samples (summarize_sampling_impl, parallel.cpp:64-79), round trips per
    window and walker, and the representatives in result order, against the compiled
    reference."""
    pos, axes, box, per = simple_cubic(3, 3, 3, 1.0, (0, 0, 1), True)
    cfg = RunConfig(emin=-60.0, emax=80.0, bin_width=0.5, windows=3, walkers_per_window=2,
                    seed=21, exchange_interval_attempts=54, max_attempts=54 * 100,
                    wl=WlParameters(check_interval_attempts=2700, collect_window_statistics=True,
                                    round_trip_margin_fraction=0.1))
    theirs = ref.run_rewl_sampling(ref.system(pos, axes, box, per), cfg)
    e = Engine(System(pos, axes, box, per, backend=orc), cfg)
    e.initialize()
    st = e.run()
    assert st.exchange_accepted > 0
    per_window = cfg.walkers_per_window
    reps = []
    for w in range(cfg.windows):
        sq = np.zeros(e.bins)
        samples = np.zeros(e.bins, dtype=np.uint64)
        trips = 0
        for k in range(w * per_window, (w + 1) * per_window):
            s = e.read_sampling(k)
            sq += s["squared_displacement"]
            samples += s["displacement_samples"]
            trips += s["round_trips"]
            assert s["round_trips"] == theirs["walker_round_trips"][k]
            reps += [r for r in s["representatives"] if r is not None]
        assert sq.tobytes() == theirs["squared"][w].tobytes()
        assert np.array_equal(samples, theirs["samples"][w])
        assert trips == theirs["round_trips"][w] and theirs["walkers"][w] == per_window
    assert sum(int(t) for t in theirs["round_trips"]) > 0
    assert len(reps) == len(theirs["representatives"])
    for (ea, sa), (eb, sb) in zip(reps, theirs["representatives"]):
        assert ea.hex() == eb.hex() and np.array_equal(sa, sb)


def test_round_trip_margin_validation(orc):
    pos, axes, box, per = simple_cubic(2, 2, 2, 1.0, (0, 0, 1), False)
    for bad in (0.5, -0.1, float("nan")):
        cfg = c1_config()
        cfg.wl.round_trip_margin_fraction = bad
        with pytest.raises(InvalidArgument, match="Invalid WL parameters"):
            Engine(System(pos, axes, box, per, backend=orc), cfg)


# ---- warm start (no reference counterpart) --------------------------------------------------
def c1_exact_log_g(orc):
    """log of the exact C1 state counts per bin (2^8 enumeration), 0 on empty bins."""
    s = System(*c1_geometry(), backend=orc)
    states = ((np.arange(256)[:, None] >> np.arange(8)) & 1).astype(np.int8) * 2 - 1
    _, energies = s.fields(states)
    counts = np.zeros(160)
    for en in energies:
        counts[orc.lib.orc_grid_index(ctypes.c_double(-20), ctypes.c_double(20),
                                      ctypes.c_double(0.25), ctypes.c_double(en))] += 1
    return np.where(counts > 0, np.log(np.maximum(counts, 1)), 0.0), counts


def test_warm_start_closed_forms(orc):
    """orc_engine_warm_start, pinned by what its definition fixes rather than by its formula:
    every window's log g keeps the input's bin-to-bin differences and starts at 0 at the
    window minimum; histograms and visited masks are empty; ln f is the given factor in the WL
    stage; the 1/t clock offset is ceil(window_bins / factor) (ws / f = 107 / 0.3 is not an
    integer, so ceil is told apart from floor and rounding). Through the run: in the 1/t stage
    the factor is active / (attempted + offset) exactly, and never exceeds the warm-start ln f
    (active <= window_bins, so active / (attempted + window_bins / f) <= f): a dropped offset
    gives active / attempted, far above f right after the switch."""
    s = System(*c1_geometry(), backend=orc)
    cfg = c1_config(windows=2, walkers_per_window=2, overlap=0.5)
    cfg.wl.final_factor = 1e-5
    e = Engine(s, cfg)
    e.initialize()
    e.run(50)  # histograms, masks and counters are no longer empty
    exact, counts = c1_exact_log_g(orc)
    rng = np.random.default_rng(5)
    start = exact + rng.normal(0.0, 0.05, exact.size) * (counts > 0) + 3.0
    f0 = 0.3
    e.warm_start(start, f0)
    offsets = {}
    for k in range(e.walkers):
        w = e.read_walker(k)
        b, t = w.info.window_begin, w.info.window_end
        assert t - b == 107  # ceil(160 / 1.5): two windows over the 160-bin grid at overlap 0.5
        part = w.log_g[b:t]
        assert part.min() == 0.0 and part[np.argmin(start[b:t])] == 0.0
        assert np.array_equal(np.diff(part), np.diff(start[b:t] - start[b:t].min()))
        assert np.abs(np.diff(part) - np.diff(start[b:t])).max() < 1e-12
        assert not w.log_g[:b].any() and not w.log_g[t:].any()
        assert not w.histogram.any() and not w.active.any() and not w.refinement_active.any()
        assert w.info.active_bins == 0 and w.info.stage == 0 and w.factor == f0
        assert w.info.clock_offset == 357 == math.ceil(107 / f0)
        offsets[k] = w.info.clock_offset
    switched = set()
    while True:
        st = e.run(20)
        for k in range(e.walkers):
            info = e.read_walker(k).info
            assert info.clock_offset == offsets[k]
            if info.stage == 1:
                switched.add(k)
                assert info.factor == info.active_bins / (info.attempted + offsets[k]), k
            if info.stage >= 1:
                assert info.factor <= f0
        if st.done:
            break
    assert switched == set(range(e.walkers))
    assert st.converged


def shifted_c1(orc, seed=1234):
    """C1 on its grid shifted by half a bin ([-20.125, 19.875], no state on a bin edge), 1/t
    to ln f = 1e-6, Philox stream; and the exact counts per bin."""
    s = System(*c1_geometry(), backend=orc)
    states = ((np.arange(256)[:, None] >> np.arange(8)) & 1).astype(np.int8) * 2 - 1
    _, energies = s.fields(states)
    exact = np.bincount(np.floor((energies + 20.125) / 0.25).astype(int), minlength=160)
    cfg = RunConfig(emin=-20.125, emax=19.875, bin_width=0.25, windows=1, walkers_per_window=2,
                    seed=seed, exchange_interval_attempts=80, max_attempts=0, rng=1,
                    wl=WlParameters(flatness=0.8, final_factor=1e-6, check_interval_attempts=800))
    e = Engine(s, cfg)
    e.initialize()
    return e, exact


def max_deviation_from(e, exact):
    lg, _, _, valid = fragment(e, 0, allow_empty=True)
    occupied = np.nonzero(exact)[0]
    if not all(valid[b] for b in occupied):
        return np.inf
    d = np.array([lg[b] - np.log(exact[b]) for b in occupied])
    return float(np.abs(d - d.mean()).max())


def test_warm_start_from_exact_dos_stays_and_converges(orc):
    """From the exact log g at ln f = 1e-5 the walkers hold the DOS within 0.01 (a flat start
    is still far off after the same attempts) and the run freezes (ln f = 1e-6) within 0.02 of
    the enumeration: the converged bound holds whatever the start. The factor must be small:
    WL at a fixed ln f keeps log g noisy by about sqrt(ln f) (0.03-0.04 at 1e-3 here), so a
    start already within the bound only stays there at ln f of order bound^2 or below."""
    warm, exact = shifted_c1(orc)
    flat, _ = shifted_c1(orc)
    warm.warm_start(np.where(exact > 0, np.log(np.maximum(exact, 1)), 0.0), 1e-5)
    for e in (warm, flat):
        e.run(2000)  # 320 000 attempts
    assert max_deviation_from(warm, exact) < 0.01 < max_deviation_from(flat, exact)
    st = warm.run()
    assert st.converged and max_deviation_from(warm, exact) < 0.02


@pytest.mark.parametrize("seed", [1234, 1, 2, 3])
def test_philox_walker_converges_to_exact_enumeration(orc, seed):
    """The oracle's Philox4x64-10 walker (proposal site from the high word of a 64 x 64-bit
    product, a uniform from an independent word of the same block) run to ln f = 1e-6 on the
    C1 lattice reproduces the exact 2^8-state DOS within 0.02 (max |dlog g| after mean-offset
    alignment over the 17 occupied bins) and with the reference's RMSE bound of 0.05
    (test_main.cpp:364-391). A biased site draw or a uniform correlated with the site moves
    the converged DOS away from the enumeration. The grid is C1's shifted by half a bin: on
    C1's own grid the exact energies sit on bin edges, and rounding in the running energy
    books a few states into neighbouring empty bins; a walker that visited such a bin in one
    stage can wait indefinitely to see it again in the next (the 1/t coverage rule of
    wang_landau.cpp:369-382), on either stream."""
    e, exact = shifted_c1(orc, seed)
    occupied = np.nonzero(exact)[0]
    assert len(occupied) == 17
    st = e.run(2_000_000)
    assert st.converged
    lg, _, _, valid = fragment(e, 0)
    assert [b for b in np.nonzero(valid)[0]] == list(occupied)
    d = np.array([lg[b] - np.log(exact[b]) for b in occupied])
    d -= d.mean()
    assert np.abs(d).max() < 0.02 and np.sqrt(np.mean(d * d)) <= 0.05


def test_fp32_couplings_are_the_rounded_fp64_build(orc):
    """precision = FP32 (no reference counterpart): every coupling of the dense and the CSR
    build is the FP64 coupling rounded to the nearest binary32 value; the matrix stays
    symmetric and the CSR pattern is unchanged; an unknown precision is refused."""
    rng = np.random.default_rng(3)
    n, side = 150, (150 / 0.3) ** (1 / 3)
    pos = rng.uniform(0.0, side, size=(n, 3))
    axes = rng.normal(size=(n, 3))
    axes /= np.linalg.norm(axes, axis=1)[:, None]
    box = np.array([side] * 3)
    for cutoff in (0.0, 3.0):
        s64 = System(pos, axes, box, True, cutoff=cutoff, backend=orc)
        s32 = System(pos, axes, box, True, cutoff=cutoff, backend=orc, precision=1)
        d64, d32 = s64.dense(), s32.dense()
        assert d32.tobytes() == d64.astype(np.float32).astype(np.float64).tobytes()
        assert np.array_equal(d32, d32.T) and np.count_nonzero(d32 != d64) > n
        if cutoff:
            (o64, c64, v64), (o32, c32, v32) = s64.csr(), s32.csr()
            assert o64.tolist() == o32.tolist() and c64.tolist() == c32.tolist()
            assert v32.tobytes() == v64.astype(np.float32).astype(np.float64).tobytes()
    with pytest.raises(InvalidArgument, match="Unknown coupling precision"):
        System(pos, axes, box, True, backend=orc, precision=2)
