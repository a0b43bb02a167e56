"""Benchmark: attempted flips/s of the REWL walker-step hot path on one MI355X.

Workload (the reference's GPU-scale plain REWL case): simple-cubic 8x8x8 periodic dipoles
(N = 512, D = 1, all axes +z), dense FP64 J (512 x 512), energy grid [-273, 354.5] with bin
0.5 (1255 bins; the seed-1234 10^4-MCS pilot range padded by 5 % and snapped to the bin),
8 windows at overlap 0.75, 16 walkers per window (128 walkers), replica exchange every
1 MCS (512 attempts), 1/t stage down to ln f = 1e-7.

A step is one exchange epoch: every walker runs one 512-attempt batch (sampling kernel),
then the flatness cadence and the even/odd exchange run on the device. `value` is the
whole-job attempted-flip rate over K timed epochs, inputs resident in HBM, with the
counter-based (Philox4x64-10) proposal stream and the lookahead sampling kernel. The same
workload on the Xoshiro stream of the reference (sequential kernel, trajectories identical to
the reference's) is timed beside it and reported under config.xoshiro.

The roofline object prices the sampling kernel's dense row stream: each accepted flip
streams one row J[i, :] of N doubles (N * 8 bytes); achieved = accepted flips per launch *
N * 8 bytes / mean launch duration (HIP events on the engine stream around every 8th launch of
the timed region).

cpu_baseline runs the reference's own run_rewl (compiled from its sources into oracle/_ref)
on the same workload with OpenMP on the host cores, bounded to a fixed attempt budget.

time_to_accuracy.plain runs fresh C2 engines (4 seeds) and reports, per seed, the device time
to the first snapshot after which the stitched DOS stays within max|dlog g| < 0.01 of a
converged C2 reference DOS (mean offset removed; tests/golden/c2_reference_dos.json, made by
tests/golden/make_c2_reference_dos.py, whose own error estimate is in
c2_reference_dos_error.json), and the median over the seeds.
time_to_accuracy.windows does the same with the system, grid and schedule split into 48
windows of 16 walkers (768 walkers in one launch per epoch).
time_to_accuracy.replicas runs 8 independent C2 runs (seeds 1234..1241) batched into one
lookahead-walker launch per epoch and measures the average of their stitched DOS: the
128-walker workload fills an eighth of the chip's wave slots, the batch fills them with
replicas.

`--c4` runs only the C4 precision study instead: a 16x16x16 dense system with FP32 Kahan
fields, the FP64 matrix-core resync every 10 MCS, and the drift it measures.

`--c5` runs only the learned warm start: converged DOS curves of a slab family (one batch),
the BF16 network trained on them by the project's kernels, and time to accuracy from flat and
warm starts on the held-out 12x12x4 slab at dilution 0.2 (tests/golden/c5_reference_dos.json).

`--c2-warm` runs only the warm-start requirement study on C2: time to accuracy from the
converged reference log g plus smooth errors of known size, at several initial ln f, against
the flat start on the same seeds.

`--c3` runs only the batched study: 64 uneven dilute CSR systems (N from 64 to 4096) whose
walkers all sample in one launch per epoch (the LDS walker kernel), replayed from HIP graphs.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK_GBS = 8000.0
CLOCK_GHZ = 2.4  # MI355X peak engine clock, for cycles per attempt
# HIP events bracket every TIMING_EVERY-th sampling launch of the timed region: the mean
# launch duration for the roofline, without an event pair around every launch of the rate
TIMING_EVERY = 8
# epochs per HIP graph replay in the untimed long runs (time to accuracy)
GRAPH_EPOCHS = 16
C2_GRID = (-273.0, 354.5, 0.5)
# HBM bytes per launch of the sampling kernels on this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a copy-kernel calibration; the summaries (with their commands) are
# committed under profiles/ and are read, not re-measured, by this script
TRAFFIC_SUMMARIES = ("lookahead_staged_kernel_c2_8x8x8.md", "lookahead_fast_kernel_c2_8x8x8.md")


def traffic_bytes(kernel: str):
    """HBM bytes per launch recorded for `kernel` in the committed counter summaries (the first
    that names it), or None."""
    for name in TRAFFIC_SUMMARIES:
        path = ROOT / "profiles" / name
        if not path.exists():
            continue
        for line in path.read_text().splitlines():
            words = line.split()
            if len(words) == 3 and words[0] == "traffic_bytes_per_launch" and words[2] == kernel:
                return float(words[1])
    return None


def c2_config(max_attempts=0, rng=0, seed=1234, windows=8):
    from replica_exchange_wang_landau_amd import RunConfig, WlParameters

    return RunConfig(emin=C2_GRID[0], emax=C2_GRID[1], bin_width=C2_GRID[2], windows=windows,
                     walkers_per_window=16, overlap=0.75, seed=seed,
                     exchange_interval_attempts=512, max_attempts=max_attempts, rng=rng,
                     wl=WlParameters(final_factor=1e-7, check_interval_attempts=512 * 100))


def gpu_leg(steps: int, warmup: int, rng: int):
    """Time `steps` exchange epochs after `warmup` untimed ones."""
    from replica_exchange_wang_landau_amd import Engine, System

    system = System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True)
    engine = Engine(system, c2_config(rng=rng))
    engine.initialize()
    engine.run(warmup)
    # the event pool is created here, outside the timed region; run() polls the device every
    # 64 epochs, so a short timed run is one chunk of launches and one synchronisation
    engine.set_timing(True, every=TIMING_EVERY)
    before = engine.status()
    t0 = time.perf_counter()
    engine.run(steps, status=False)  # returns after the device has finished the epochs
    t1 = time.perf_counter()
    after = engine.status()
    kernel_ms, launches = engine.kernel_time()
    attempted = after.attempted - before.attempted
    accepted = after.accepted - before.accepted
    epochs = after.epoch - before.epoch
    seconds = t1 - t0
    # sampling kernel: accepted flips per launch * row bytes / launch time
    row_bytes = system.n * 8
    per_launch_bytes = (accepted / max(1, epochs)) * row_bytes
    achieved = per_launch_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    return {
        "attempted": attempted, "accepted": accepted, "epochs": epochs, "seconds": seconds,
        "rate": attempted / seconds, "kernel_ms": kernel_ms, "launches": launches,
        "kernel": engine.kernel()[1],
        "achieved_gbs": achieved, "per_launch_bytes": per_launch_bytes,
        "ambiguous": after.ambiguous_decisions,
    }


REFERENCE_DOS = ROOT / "tests" / "golden" / "c2_reference_dos.json"


def reference_dos():
    """(log g, valid mask, record) of the converged C2 reference, or None."""
    if not REFERENCE_DOS.exists():
        return None
    import numpy as np

    ref = json.loads(REFERENCE_DOS.read_text())
    return (np.array([float.fromhex(v) for v in ref["log_g"]]),
            np.array(ref["valid"], dtype=np.uint8), ref)


def stays_below(times, devs, threshold: float):
    """The first snapshot time from which every later snapshot lies below `threshold`, or
    None when the last snapshot does not."""
    reached = None
    for t, d in zip(times, devs):
        if d < threshold:
            reached = t if reached is None else reached
        else:
            reached = None
    return reached


def accuracy_series(run_epochs, read_dos, ref, budget_s: float, interval: int, label: str,
                    threshold: float = 0.01, stay: int = 3):
    """Device seconds (inside run_epochs only) and max|dlog g| at every snapshot: runs
    `interval` epochs between snapshots until the deviation has stayed below the threshold
    for `stay` snapshots, the run is done or the budget is spent."""
    import numpy as np

    from replica_exchange_wang_landau_amd import analysis

    ref_lg, ref_valid = ref[0], ref[1]
    device, times, devs, below = 0.0, [], [], 0
    while device < budget_s:
        t0 = time.perf_counter()
        done = run_epochs(interval)
        device += time.perf_counter() - t0
        try:
            lg, valid = read_dos()
        except (RuntimeError, ValueError):
            continue  # a window without a common valid bin yet
        if np.any((ref_valid != 0) & (valid == 0)):
            dev = float("inf")  # reference bins not all found yet
        else:
            dev = analysis.max_deviation(lg, valid, ref_lg, ref_valid)
        times.append(device)
        devs.append(dev)
        below = below + 1 if dev < threshold else 0
        print(f"{label}: {device:.2f} s, max|dlog g| {dev:.4f}", file=sys.stderr, flush=True)
        if below >= stay or done:
            break
    return times, devs


def tta_seeds(make_engine, seeds, budget_s: float, interval: int, label: str,
              threshold: float = 0.01, ref=None):
    """Time to accuracy of one run per seed (each engine alone on the device, HIP-graph
    replays): per seed the first snapshot after which max|dlog g| against `ref` (log g, valid
    mask; the C2 reference by default) stays below the threshold (stays_below), and the
    median over the seeds (None where a seed did not get there)."""
    import statistics

    from replica_exchange_wang_landau_amd import analysis

    ref = reference_dos() if ref is None else ref
    if ref is None:
        return None
    per_seed, attempted, walkers, kernel = [], [], 0, None
    for seed in seeds:
        engine = make_engine(seed)
        engine.set_graph(GRAPH_EPOCHS)
        walkers, kernel = engine.walkers, engine.kernel()[1]

        def run(n, e=engine):
            return bool(e.run(n).done)

        times, devs = accuracy_series(run, lambda e=engine: analysis.engine_dos(e), ref,
                                      budget_s, interval, f"{label} seed {seed}", threshold)
        reached = stays_below(times, devs, threshold)
        per_seed.append(reached)
        attempted.append(int(engine.status().attempted))
        engine.close()
    done = [t for t in per_seed if t is not None]
    median = (statistics.median(per_seed) if len(done) == len(per_seed) else
              (statistics.median(sorted(done) + [float("inf")] * (len(per_seed) - len(done)))
               if done else None))
    return {"seconds_median": median if median != float("inf") else None,
            "seconds": per_seed, "seeds": list(seeds), "attempted": attempted,
            "walkers": walkers, "kernel": kernel, "interval_epochs": interval,
            "budget_seconds": budget_s}


def cpu_leg(mcs_per_walker: int):
    """The compiled reference (oracle/_ref) if present, else the C oracle port."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    cfg = c2_config(max_attempts=512 * mcs_per_walker)
    from replica_exchange_wang_landau_amd.engine import simple_cubic

    pos, axes, box, per = simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True)
    from harness import REF_LIB, Reference, oracle

    if REF_LIB.exists():
        ref = Reference()
        system = ref.system(pos, axes, box, per)
        t0 = time.perf_counter()
        out = ref.run_rewl(system, cfg)
        seconds = time.perf_counter() - t0
        attempted = int(out["counters"][0])
        kind = "reference"
    else:
        from replica_exchange_wang_landau_amd import Engine, System

        s = System(pos, axes, box, per, backend=oracle())
        e = Engine(s, cfg)
        t0 = time.perf_counter()
        e.initialize()
        st = e.run()
        seconds = time.perf_counter() - t0
        attempted = st.attempted
        kind = "port"
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"value": attempted / seconds, "unit": "attempted flips/s", "cores": threads,
            "host_cpus": os.cpu_count(), "usable_cpus": usable,
            "kind": kind,
            "sample": f"run_rewl on the same workload, {mcs_per_walker} MCS per walker "
                      f"({attempted} attempts incl. walker init) in {seconds:.2f} s, OpenMP on "
                      f"{threads} of the host's {os.cpu_count()} CPUs ({usable} usable by "
                      f"this process; at most 16 threads, the box's share)"}


def c4_system():
    """C4 geometry: simple-cubic 16x16x16 (N = 4096), spacing 1, periodic box 16, each site's
    axis drawn uniformly from {+x, -x, +y, -y, +z, -z} by numpy's default_rng(1234)."""
    import numpy as np

    from replica_exchange_wang_landau_amd import System
    from replica_exchange_wang_landau_amd.engine import simple_cubic

    pos, _, box, per = simple_cubic(16, 16, 16, 1.0, (0.0, 0.0, 1.0), True)
    choice = np.random.default_rng(1234).integers(0, 6, size=len(pos))
    unit = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    return System(pos, unit[choice], box, per)


def c4_leg(epochs: int, warmup: int, resyncs: int):
    """C4 precision study: dense J (FP32 copy streamed by the walkers, FP64 kept for the
    resync), FP32 Kahan fields, FP64 H = J.S resync every 10 MCS, 32 windows x 8 walkers,
    bin 2.0 over the seed-1234 10^3-MCS pilot range padded by 5 %, Philox stream. (A pilot of
    10^4 MCS streams 41 M rows of the 128 MiB J on one wave and takes minutes.)"""
    import math

    from replica_exchange_wang_landau_amd import Engine, RunConfig, WlParameters, abi

    system = c4_system()
    n = system.n
    print(f"c4: N={n}, pilot of 1000 MCS", file=sys.stderr, flush=True)
    lo, hi = system.pilot(1234, 1_000 * n)
    print(f"c4: pilot range [{lo:.1f}, {hi:.1f}]", file=sys.stderr, flush=True)
    pad = 0.05 * (hi - lo)
    emin, emax = 2.0 * math.floor((lo - pad) / 2.0), 2.0 * math.ceil((hi + pad) / 2.0)
    cfg = RunConfig(emin=emin, emax=emax, bin_width=2.0, windows=32, walkers_per_window=8,
                    overlap=0.75, seed=1234, exchange_interval_attempts=n, rng=abi.RNG_PHILOX,
                    fields=abi.FIELDS_FP32_KAHAN, resync_interval_attempts=10 * n,
                    wl=WlParameters(final_factor=1e-7, check_interval_attempts=100 * n))
    engine = Engine(system, cfg)
    engine.initialize()
    print(f"c4: {engine.walkers} walkers initialised", file=sys.stderr, flush=True)
    engine.run(warmup)
    before = engine.status()
    t0 = time.perf_counter()
    after = engine.run(epochs)
    seconds = time.perf_counter() - t0
    print(f"c4: {epochs} epochs in {seconds:.2f} s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    for _ in range(resyncs):
        engine.resync()
    resync_s = (time.perf_counter() - t1) / max(1, resyncs)
    f, e, count = engine.drift()
    flops = 2.0 * n * n * engine.walkers
    return {
        "workload": f"C4: simple-cubic 16x16x16 periodic, random +-x/y/z axes (seed 1234), "
                    f"N={n}, grid [{emin},{emax}] bin 2.0 (10^3-MCS pilot), 32 windows x 8 "
                    f"walkers, FP32 Kahan fields, FP64 resync every 10 MCS",
        "attempted_flips_per_s": (after.attempted - before.attempted) / seconds,
        "acceptance": (after.accepted - before.accepted) / max(1, after.attempted - before.attempted),
        # algorithmic bytes: every accepted flip streams its FP32 row of J (N x 4 B); the 64 MiB
        # FP32 J exceeds the 4 MiB L2 of an XCD, so the rows come from MALL or HBM
        "row_bytes_gbs": (after.accepted - before.accepted) * n * 4.0 / seconds / 1e9,
        "epochs_timed": epochs,
        "resync": {"ms": resync_s * 1e3, "flop": flops, "achieved_tflops": flops / resync_s / 1e12,
                   "peak_tflops": 78.6, "frac": flops / resync_s / 1e12 / 78.6,
                   "note": "three launches (spin matrix, MFMA J.S, finish) timed together"},
        "drift": {"max_field_abs": f, "max_energy_bins": e, "resyncs": count},
    }


def c3_leg(epochs: int, warmup: int, graph: int, pilot_mcs: int = 1000,
           cadence: str = "equalized"):
    """C3 batched study (replica_exchange_wang_landau_amd.workloads): 64 systems, N uniform
    in {64..4096}, uniform random placement at density 0.3, axes uniform on the sphere, FP32
    CSR couplings within cutoff 3; per system the seed-1234 pilot range (pilot_mcs MCS, all
    systems in one launch) padded by 5 % on bins of 1.0, clamp(round(4 + 12 (N - 64) / 4032),
    4, 16) windows over the occupied energy intervals at overlap 0.75, 8 walkers per window,
    1/t to ln f = 1e-7, seed 1234 + s, Philox stream. The exchange cadence is picked for the
    batch: "equalized" exchanges every round(4096 / N) MCS, so that every walker makes about
    4096 attempts per epoch and no wave idles while the largest systems finish theirs; "1"
    exchanges every MCS. Every walker of every
    system runs in one sampling launch per epoch (LDS walker; window bins in HBM where a
    window does not fit in LDS), epochs replayed from HIP graphs of `graph` epochs."""
    from replica_exchange_wang_landau_amd import Batch, Engine, System, abi, workloads

    t_setup = time.perf_counter()
    systems = workloads.systems(64, 1234)
    lo, hi = System.pilot_many(systems, 1234, [pilot_mcs * s.n for s in systems])
    pilot_s = time.perf_counter() - t_setup
    print(f"c3: pilots of {len(systems)} systems in {pilot_s:.1f} s", file=sys.stderr, flush=True)
    engines = []
    for s, (system, a, b) in enumerate(zip(systems, lo, hi)):
        mcs = max(1, round(workloads.MAX_SPINS / system.n)) if cadence == "equalized" else 1
        cfg = workloads.config(system, a, b, seed=1234 + s, exchange_mcs=mcs,
                               energies=workloads.sample_energies(system, 2048, s))
        e = Engine(system, cfg)
        e.initialize()
        e.set_kernel(abi.KERNEL_LDS)
        engines.append(e)
        if s % 16 == 15:
            print(f"c3: {s + 1} engines initialised", file=sys.stderr, flush=True)
    setup_s = time.perf_counter() - t_setup
    batch = Batch(engines)
    batch.set_graph(graph)
    # the LDS budget (which windows keep their bins in HBM), picked on the live run
    budget, budget_seconds = batch.autotune(graph if graph > 0 else 16)  # plain launches at 0
    layout = batch.layout()
    print(f"c3: LDS budget {budget} B ({budget_seconds}), layout {layout}", file=sys.stderr,
          flush=True)
    batch.run(warmup)
    before = [e.status() for e in engines]
    t0 = time.perf_counter()
    ran, done = batch.run(epochs)
    seconds = time.perf_counter() - t0
    after = [e.status() for e in engines]
    attempted = sum(x.attempted - y.attempted for x, y in zip(after, before))
    accepted = sum(x.accepted - y.accepted for x, y in zip(after, before))
    # algorithmic bytes: an accepted flip reads its CSR row, z (FP32 value 4 B + column 4 B)
    row_bytes = sum((x.accepted - y.accepted) * (e.system.nnz / e.system.n) * 8.0
                    for x, y, e in zip(after, before, engines))
    walkers = sum(e.walkers for e in engines)
    return {
        "workload": "C3: 64 dilute dipole systems, N uniform in {64..4096} (seed 1234), "
                    "uniform random placement at density 0.3 in a periodic cube, axes uniform "
                    "on the sphere, CSR cutoff 3 with FP32 couplings, 4-16 windows (over the "
                    "occupied energy intervals) x 8 walkers, bin 1.0, Philox stream",
        "exchange_cadence": cadence,
        "systems": len(engines), "walkers": walkers,
        "spins": [int(e.system.n) for e in engines],
        "gapped_systems": sum(e.config.explicit_windows is not None for e in engines),
        "lds_budget": budget,
        "autotune_seconds": dict(zip(("160KiB", "64KiB", "32KiB"), budget_seconds)),
        "launch_groups": layout["groups"], "bins_in_hbm_systems": layout["bins_in_hbm"],
        "max_lds_bytes": layout["max_lds_bytes"],
        "attempted_flips_per_s": attempted / seconds, "epochs_timed": ran, "engines_done": done,
        "acceptance": accepted / max(1, attempted),
        "ms_per_epoch": seconds * 1e3 / max(1, ran),
        "graph_epochs": graph,
        "roofline": {"bound": "hbm", "achieved": row_bytes / seconds / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": row_bytes / seconds / 1e9 / HBM_PEAK_GBS,
                     "note": "CSR row bytes (8 per entry) of accepted flips over the epoch time"},
        "setup_seconds": {"pilot": pilot_s, "total": setup_s},
    }


def tta_replicas_leg(budget_seconds: float, interval_epochs: int, replicas: int,
                     graph: int = 16):
    """Time to accuracy of `replicas` independent C2 runs (Philox, seeds 1234 + r) run as one
    batch (one lookahead-walker launch over all their walkers per epoch, HIP-graph replays of
    `graph` epochs), measured on the average of their stitched DOS (analysis.average_dos), with
    the stays-below criterion of tta_seeds."""
    from replica_exchange_wang_landau_amd import Batch, Engine, System, analysis

    ref = reference_dos()
    if ref is None or replicas < 1:
        return None
    engines = []
    for r in range(replicas):
        e = Engine(System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True),
                   c2_config(rng=1, seed=1234 + r))
        e.initialize()
        engines.append(e)
    batch = Batch(engines)
    batch.set_graph(graph)
    kernel = batch.kernel()

    def run(n):
        return batch.run(n)[1] == len(engines)

    times, devs = accuracy_series(
        run, lambda: analysis.average_dos([analysis.engine_dos(e) for e in engines]), ref,
        budget_seconds, interval_epochs, "tta replicas")
    attempted = sum(e.status().attempted for e in engines)
    device = times[-1] if times else 0.0
    batch.close()
    return {"replicas": replicas, "kernel": kernel, "threshold": 0.01,
            "seconds": stays_below(times, devs, 0.01), "attempted": int(attempted),
            "rate": attempted / device if device else None,
            "interval_epochs": interval_epochs, "graph_epochs": graph,
            "budget_seconds": budget_seconds}


def train_warm_start(curves: int, steps: int, final: float = 1e-4, tag: str = "c5",
                     gaussian: bool = False):
    """The warm-start network: `curves` slabs (Lx = Ly in 6..12, Lz in 2..6, dilution in
    [0, 0.4], +z or random in-plane axes; seed 1234) run to ln f = `final` together in one batch,
    and the 4 x 128 BF16 network trained on (E/N, N, density, dilution, Lz, axis mode) ->
    log g / N by the project's kernels (Adam 1e-3, then 3e-4 and 1e-4 for the last quarter
    each, batch 1024). With `gaussian` the target is log g / N minus each realization's
    Gaussian part -(E/N)^2 / (2 s), s = sum_{i<j} J_ij^2 / N (warmstart.energy_spread), so the
    network learns the non-Gaussian shape. The last tenth of the curves validates it. Returns
    (network, per-spin
    validation RMS after removing each curve's mean offset, training record)."""
    import numpy as np

    from replica_exchange_wang_landau_amd import warmstart as ws

    t0 = time.perf_counter()
    slabs = ws.slab_family(curves, 1234)
    data = ws.converged_curves(
        slabs, final_factor=final,
        report=lambda d, n: print(f"{tag}: {d}/{n} training curves converged", file=sys.stderr,
                                  flush=True))
    keep = [(s, c) for s, c in zip(slabs, data) if (c[2] != 0).sum() > 4]
    split = len(keep) * 9 // 10
    x, y = ws.training_set([s for s, _ in keep[:split]], [c for _, c in keep[:split]], gaussian)
    vx, vy = ws.training_set([s for s, _ in keep[split:]], [c for _, c in keep[split:]], gaussian)
    curve_s = time.perf_counter() - t0
    mlp = ws.Mlp(1234)
    t1 = time.perf_counter()
    losses = np.concatenate([
        mlp.train(x, y, steps=steps // 2, batch=1024, learning_rate=1e-3, seed=5),
        mlp.train(x, y, steps=steps // 4, batch=1024, learning_rate=3e-4, seed=6),
        mlp.train(x, y, steps=steps - steps // 2 - steps // 4, batch=1024, learning_rate=1e-4,
                  seed=7)])
    train_s = time.perf_counter() - t1
    pred = mlp.predict(vx)
    offsets, at = [], 0
    validation = []  # per unseen curve: E/N of its bins and the network's error there
    for _, c in keep[split:]:
        m = int((c[2] != 0).sum())
        d = pred[at:at + m] - vy[at:at + m]
        offsets.append(d - d.mean())
        validation.append((2.0 * vx[at:at + m, 0].astype(np.float64), d.astype(np.float64)))
        at += m
    val_rms = float(np.sqrt(np.mean(np.concatenate(offsets) ** 2)))
    print(f"{tag}: {len(keep)} curves, {len(x)} samples in {curve_s:.0f} s; trained {steps} "
          f"steps in {train_s:.1f} s, loss {losses[0]:.3g} -> {losses[-1]:.3g}", file=sys.stderr,
          flush=True)
    record = {"curves": split, "validation_curves": len(keep) - split, "final_ln_f": final,
              "validation_rms_per_spin": val_rms, "samples": int(len(x)), "steps": steps,
              "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
              "target": "log g / N minus the Gaussian part" if gaussian else "log g / N",
              "curve_seconds": curve_s, "train_seconds": train_s}
    return mlp, val_rms, record, validation


C5_REFERENCE = ROOT / "tests" / "golden" / "c5_reference_dos.json"


def c5_leg(curves: int, final: float, steps: int, budget_s: float, seeds: int,
           factors=(1e-2, 1e-4), interval: int = 2000, gaussian: bool = False):
    """Learned warm start on the held-out slab (train_warm_start for the network).

    Reference: tests/golden/c5_reference_dos.json (make_c5_reference_dos.py), the average of
    independent runs of the held-out 12x12x4 slab (dilution 0.2) frozen at ln f <= 1e-6, on
    the grid of warmstart.slab_config from its recorded pilot range. Time to accuracy as in
    tta_seeds (first snapshot after which max|dlog g| stays below 0.01, median over `seeds`
    seeds) from a flat start and from the network's log g at each initial ln f in `factors`.
    Also reported: the network start's own max|dlog g| to the reference (start_error), the
    gain that error allows when the run's ln f must fall below f_0.01 anyway,
    1 / (1 - 0.01 / start_error), and the spread between dilution realizations recorded with
    the reference (a floor under any network that sees only the slab's parameters)."""
    import numpy as np

    from replica_exchange_wang_landau_amd import Engine, analysis
    from replica_exchange_wang_landau_amd import warmstart as ws

    if not C5_REFERENCE.exists():
        return None
    rec = json.loads(C5_REFERENCE.read_text())
    ref = (np.array([float.fromhex(v) for v in rec["log_g"]]),
           np.array(rec["valid"], dtype=np.uint8))
    mlp, val_rms, training, _ = train_warm_start(curves, steps, final, gaussian=gaussian)
    slab = ws.Slab(*rec["slab"])
    system = slab.system()
    if system.n != rec["spins"]:
        raise RuntimeError("held-out slab differs from the recorded one")
    lo, hi = rec["pilot"]

    def engine(seed, factor=None, guess=None):
        e = Engine(system, ws.slab_config(system, lo, hi, seed, 1e-6))
        if [e.config.emin, e.config.emax, e.config.bin_width] != rec["grid"]:
            raise RuntimeError("held-out grid differs from the recorded one")
        e.initialize()
        if factor is not None:
            e.warm_start(guess, factor)
        return e

    probe = engine(0)
    guess = ws.predicted_log_g(mlp, probe, slab.features(system.n),
                               ws.energy_spread(system) if gaussian else None)
    probe.close()
    start = guess - guess[ref[1] != 0].mean()
    d = (start - ref[0])[ref[1] != 0]
    d -= d.mean()
    start_error = float(np.max(np.abs(d)))
    seed_list = list(range(77, 77 + seeds))
    flat = tta_seeds(lambda s: engine(s), seed_list, budget_s, interval, "c5 flat", ref=ref)
    warm = {}
    for f0 in factors:
        warm[f"{f0:g}"] = tta_seeds(lambda s, f0=f0: engine(s, f0, guess), seed_list, budget_s,
                                    interval, f"c5 warm ln f {f0:g}", ref=ref)
    speedup = {k: flat["seconds_median"] / v["seconds_median"] for k, v in warm.items()
               if flat["seconds_median"] and v["seconds_median"]}
    return {
        "workload": "C5: learned warm start (4 x 128 BF16 MLP, FP32 accumulate) on the "
                    "held-out 12x12x4 slab, dilution 0.2, +z axes",
        "training": training, "spins": int(system.n),
        "reference": {"runs": rec["runs"], "final_ln_f": rec["final_factor"],
                      "max_dev_between_runs": rec["max_pairwise_max_dev"]},
        "realization_spread_max_dev": [r["max_dev"] for r in rec["siblings"]],
        "start_error_max_dev": start_error,
        "start_error_rms": float(np.sqrt(np.mean(d * d))),
        "gain_allowed_by_start_error": (1.0 / (1.0 - 0.01 / start_error)
                                        if start_error > 0.01 else None),
        "flat": flat, "warm": warm, "speedup_median": speedup,
    }


def c2_warm_leg(errors, factors, seeds: int, budget_s: float, interval: int):
    """How close a warm start must be to pay on C2 (8x8x8 periodic dense, 8 windows x 16
    walkers, Philox stream). Each start is the converged C2 reference log g plus a smooth error
    of known amplitude, eps * sin(6 pi x) with x running from 0 to 1 over the reference's valid
    bins (three periods, so every window sees a slope); each is run at each initial ln f in
    `factors`. Time to accuracy as in tta_seeds, against the same reference, next to the flat
    start on the same seeds. start_error is the start's own max|dlog g| to the reference."""
    import numpy as np

    from replica_exchange_wang_landau_amd import Engine, System, analysis

    ref = reference_dos()
    if ref is None:
        return None
    ref_lg, ref_valid = ref[0], ref[1]
    idx = np.flatnonzero(ref_valid != 0)
    x = (idx - idx[0]) / max(1, idx[-1] - idx[0])
    shape = np.zeros_like(ref_lg)
    shape[idx] = np.sin(6.0 * np.pi * x)
    seed_list = list(range(1234, 1234 + seeds))

    def engine(seed, start=None, factor=None):
        e = Engine(System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True),
                   c2_config(rng=1, seed=seed))
        e.initialize()
        if start is not None:
            e.warm_start(start, factor)
        return e

    flat = tta_seeds(lambda s: engine(s), seed_list, 40.0, interval, "c2 flat")
    print("c2 warm: flat leg done", file=sys.stderr, flush=True)
    rows = []
    for eps in errors:
        start = np.where(ref_valid != 0, ref_lg + eps * shape, 0.0)
        row = {"eps": eps,
               "start_error": analysis.max_deviation(start, ref_valid, ref_lg, ref_valid),
               "factors": {}}
        for f0 in factors:
            r = tta_seeds(lambda s, st=start, f=f0: engine(s, st, f), seed_list, budget_s,
                          interval, f"c2 warm eps {eps:g} ln f {f0:g}")
            med = r["seconds_median"]
            row["factors"][f"{f0:g}"] = {
                "seconds": r["seconds"], "seconds_median": med,
                "speedup_median": (flat["seconds_median"] / med
                                   if med and flat["seconds_median"] else None)}
        best = [v["speedup_median"] for v in row["factors"].values() if v["speedup_median"]]
        row["best_speedup"] = max(best) if best else None
        rows.append(row)
        print(f"c2 warm: eps {eps:g} done, best speedup {row['best_speedup']}",
              file=sys.stderr, flush=True)
    paying = [r["start_error"] for r in rows if (r["best_speedup"] or 0.0) >= 5.0]
    return {
        "workload": "C2 warm-start requirement: 8x8x8 periodic dense, 8 windows x 16 walkers, "
                    "Philox stream, started from the converged reference log g plus "
                    "eps * sin(6 pi x) over its valid bins",
        "flat": flat, "rows": rows, "budget_seconds_per_warm_run": budget_s,
        "interval_epochs": interval,
        "largest_start_error_paying_5x": max(paying) if paying else None,
    }


def calibrate_copy(mib: int):
    """A plain device copy of known size (read mib MiB, write mib MiB), five times: the
    reference point that scales FETCH_SIZE / WRITE_SIZE counter readings to bytes."""
    import torch

    src = torch.ones(mib * 2**20 // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4,
                      "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--rng", choices=["xoshiro", "philox"], default="philox")
    ap.add_argument("--cpu-mcs", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tta-seconds", type=float, default=40.0,
                    help="device-time budget of the time-to-accuracy leg (0 skips it)")
    ap.add_argument("--tta-interval", type=int, default=2000, help="epochs between snapshots")
    ap.add_argument("--tta-seeds", type=int, default=4,
                    help="independent runs per time-to-accuracy figure (median reported)")
    ap.add_argument("--tta-replicas", type=int, default=8,
                    help="independent C2 replicas batched for the averaged time to accuracy "
                         "(0 skips it)")
    ap.add_argument("--tta-replicas-seconds", type=float, default=60.0,
                    help="device-time budget of the replica leg")
    ap.add_argument("--tta-windows", type=int, default=48,
                    help="windows of the many-window time to accuracy (0 skips it)")
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only run a plain copy of MIB MiB (counter calibration)")
    ap.add_argument("--c5", action="store_true",
                    help="only run the learned warm-start study")
    ap.add_argument("--c5-curves", type=int, default=512, help="training curves (--c5)")
    ap.add_argument("--c5-final", type=float, default=1e-4,
                    help="ln f the training curves are converged to (--c5)")
    ap.add_argument("--c5-steps", type=int, default=60_000, help="training steps (--c5)")
    ap.add_argument("--c5-gaussian", action="store_true",
                    help="train the network on the non-Gaussian part of log g / N (--c5)")
    ap.add_argument("--c5-seconds", type=float, default=60.0,
                    help="device-time budget per held-out run (--c5)")
    ap.add_argument("--c2-warm", action="store_true",
                    help="only run the C2 warm-start requirement study (reference log g plus "
                         "errors of known size)")
    ap.add_argument("--c2-warm-errors", type=float, nargs="+",
                    default=[0.005, 0.01, 0.03, 0.1, 0.3, 1.0],
                    help="start errors (amplitudes) of the warm-start study (--c2-warm)")
    ap.add_argument("--c2-warm-factors", type=float, nargs="+",
                    default=[1e-3, 1e-4, 1e-5, 1e-6],
                    help="initial ln f of the warm-start study (--c2-warm)")
    ap.add_argument("--c3", action="store_true",
                    help="only run the C3 batched study (64 uneven CSR systems in one launch)")
    ap.add_argument("--graph", type=int, default=16, help="epochs per HIP graph (--c3)")
    ap.add_argument("--c3-cadence", choices=["equalized", "1"], default="equalized",
                    help="exchange cadence of the C3 batch (--c3)")
    ap.add_argument("--c4", action="store_true",
                    help="only run the C4 FP32-field precision study (resync kernel, drift)")
    ap.add_argument("--line-c4-epochs", type=int, default=40,
                    help="timed C4 epochs in the default line's compact c4 object (0: leave out)")
    ap.add_argument("--line-c3-epochs", type=int, default=32,
                    help="timed C3 epochs in the default line's compact c3 object (0: leave out)")
    args = ap.parse_args()
    if args.c5:
        print(json.dumps(c5_leg(args.c5_curves, args.c5_final, args.c5_steps, args.c5_seconds,
                                args.tta_seeds, gaussian=args.c5_gaussian)))
        return
    if args.c2_warm:
        print(json.dumps(c2_warm_leg(args.c2_warm_errors, args.c2_warm_factors,
                                     args.tta_seeds, 2.5, 256)))
        return
    if args.c3:
        print(json.dumps(c3_leg(args.steps if args.steps != 3000 else 200, 20, args.graph,
                                cadence=args.c3_cadence)))
        return
    if args.c4:
        print(json.dumps(c4_leg(args.steps if args.steps != 3000 else 200, 20, 10)))
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU only; there is no multi-GPU path "
              "(the walkers would shard by energy window across GPUs)", file=sys.stderr)
        sys.exit(2)
    rng = 0 if args.rng == "xoshiro" else 1
    # N = 512 sites without trace, forcing or the Nalivaiko mask: the Philox stream runs the
    # staged lookahead walker (128 walkers, one per CU), the Xoshiro stream its lookahead kernel
    g = gpu_leg(args.steps, args.warmup, rng)
    kernel = g["kernel"]
    print(f"bench: {args.rng} leg done", file=sys.stderr, flush=True)
    other = gpu_leg(args.steps, args.warmup, 1 - rng)
    print("bench: second-stream leg done", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_leg(args.cpu_mcs)
    print("bench: cpu leg done", file=sys.stderr, flush=True)
    tta = None
    if args.tta_seconds > 0 and reference_dos() is not None:
        from replica_exchange_wang_landau_amd import Engine, System

        def c2_engine(windows):
            def make(seed):
                e = Engine(System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True),
                           c2_config(rng=1, seed=seed, windows=windows))
                e.initialize()
                return e
            return make
        seeds = list(range(1234, 1234 + args.tta_seeds))
        ref = reference_dos()[2]
        tta = {"threshold": 0.01,
               "criterion": "device seconds to the first snapshot after which max|dlog g| "
                            "(mean offset removed) stays below the threshold, median over "
                            f"{len(seeds)} seeds",
               "reference": f"converged plain REWL to ln f 1e-7, seed {ref['seed']}, "
                            f"{ref['attempted']} attempts (tests/golden/c2_reference_dos.json)"}
        error = ROOT / "tests" / "golden" / "c2_reference_dos_error.json"
        if error.exists():
            e = json.loads(error.read_text())
            tta["reference_error"] = {
                "independent_runs": e["runs"], "their_ln_f": e["final_factor"],
                "max_dev_between_them": e["max_pairwise_max_dev"],
                "max_dev_to_reference": e["max_vs_reference"]}
        tta["plain"] = tta_seeds(c2_engine(8), seeds, args.tta_seconds, args.tta_interval,
                                 "tta plain")
        print("bench: plain time-to-accuracy leg done", file=sys.stderr, flush=True)
        if cpu and tta["plain"]["seconds_median"] is not None:
            # the reference CPU path at its measured rate over the median run's attempts
            att = sorted(tta["plain"]["attempted"])[len(seeds) // 2]
            tta["plain"]["cpu_estimate_seconds"] = att / cpu["value"]
        if args.tta_windows > 0:
            tta["windows"] = tta_seeds(c2_engine(args.tta_windows), seeds, 60.0, 500,
                                       f"tta {args.tta_windows} windows")
            tta["windows"]["windows"] = args.tta_windows
            print("bench: many-window leg done", file=sys.stderr, flush=True)
        if args.tta_replicas > 0:
            tta["replicas"] = tta_replicas_leg(args.tta_replicas_seconds, args.tta_interval,
                                               args.tta_replicas)
            print("bench: replica leg done", file=sys.stderr, flush=True)
        plain = tta["plain"]["seconds_median"]
        if plain:
            tta["speedup_vs_plain"] = {
                k: plain / v for k, v in (
                    ("windows", (tta.get("windows") or {}).get("seconds_median")),
                    ("replicas", (tta.get("replicas") or {}).get("seconds"))) if v}
    line = {
        "metric": "attempted flips/s (all walkers)",
        "value": g["rate"],
        "unit": "flips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["seconds"] * 1e3 / max(1, g["epochs"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain REWL, simple-cubic 8x8x8 periodic dense FP64 J (N=512), "
                        "grid [-273,354.5] bin 0.5, 8 windows x 16 walkers, exchange every 1 MCS",
            "walkers": 128, "spins": 512, "rng": args.rng, "fields": "fp64",
            "epochs_timed": g["epochs"], "attempted": g["attempted"],
            "acceptance": g["accepted"] / max(1, g["attempted"]),
            "ambiguous_log_decisions": g["ambiguous"],
            ("philox" if rng == 0 else "xoshiro"): {
                "value": other["rate"], "kernel": other["kernel"], "kernel_ms": other["kernel_ms"],
                "ms_per_step": other["seconds"] * 1e3 / max(1, other["epochs"]),
                "acceptance": other["accepted"] / max(1, other["attempted"])},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": g["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": g["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic_bytes(kernel),
            "kernel": kernel,
            "kernel_ms": g["kernel_ms"],
            "bytes_per_launch": g["per_launch_bytes"],
            # `achieved` prices the algorithmic row stream (a 2 MiB J that stays in L2); the
            # counter-measured HBM bytes over the same launch time are the HBM use proper.
            # What bounds the kernel is each walker's serial decision chain: cycles per attempt
            # of one walker over the launch (512 attempts per walker per launch)
            "hbm_measured_gbs": (traffic_bytes(kernel) / (g["kernel_ms"] * 1e-3) / 1e9
                                 if traffic_bytes(kernel) and g["kernel_ms"] > 0 else None),
            "row_bytes_note": "algorithmic row bytes, served from L2 (J is 2 MiB)",
            "cycles_per_attempt": g["kernel_ms"] * 1e-3 * CLOCK_GHZ * 1e9 / 512.0,
            "clock_ghz": CLOCK_GHZ,
        },
        "cpu_baseline": cpu,
        "time_to_accuracy": tta,
    }
    # the other GPU-scale configurations, compact, so that the default line measures them too
    if args.line_c4_epochs > 0:
        c4 = c4_leg(args.line_c4_epochs, 5, 10)
        line["c4"] = {
            "workload": c4["workload"], "attempted_flips_per_s": c4["attempted_flips_per_s"],
            "acceptance": c4["acceptance"], "row_bytes_gbs": c4["row_bytes_gbs"],
            "row_bytes_frac_hbm_peak": c4["row_bytes_gbs"] / HBM_PEAK_GBS,
            "epochs_timed": c4["epochs_timed"],
            "resync_ms": c4["resync"]["ms"], "resync_tflops": c4["resync"]["achieved_tflops"],
            "resync_frac_fp64_mfma_peak": c4["resync"]["frac"],
            "drift": c4["drift"]}
        print("bench: c4 leg done", file=sys.stderr, flush=True)
    if args.line_c3_epochs > 0:
        graph = 16 if args.line_c3_epochs % 16 == 0 else 0
        c3 = c3_leg(args.line_c3_epochs, 16, graph)
        line["c3"] = {k: c3[k] for k in (
            "workload", "exchange_cadence", "systems", "walkers", "attempted_flips_per_s",
            "epochs_timed", "ms_per_epoch", "graph_epochs", "lds_budget", "acceptance")}
        line["c3"]["row_bytes_frac_hbm_peak"] = c3["roofline"]["frac"]
        print("bench: c3 leg done", file=sys.stderr, flush=True)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
