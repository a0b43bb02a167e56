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
N * 8 bytes / mean launch duration (HIP events on the engine stream).

cpu_baseline runs the reference's own run_rewl (compiled from its sources into oracle/_ref)
on the same workload with OpenMP on the host cores, bounded to a fixed attempt budget.
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
C2_GRID = (-273.0, 354.5, 0.5)
# HBM bytes per launch of the sampling kernels on this workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a copy-kernel calibration; the summary (with its commands) is
# committed here and is read, not re-measured, by this script
TRAFFIC_SUMMARY = ROOT / "profiles" / "lookahead_fast_kernel_c2_8x8x8.md"


def traffic_bytes(kernel: str):
    """HBM bytes per launch recorded for `kernel` in the committed counter summary, or None."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    for line in TRAFFIC_SUMMARY.read_text().splitlines():
        words = line.split()
        if len(words) == 3 and words[0] == "traffic_bytes_per_launch" and words[2] == kernel:
            return float(words[1])
    return None


def c2_config(max_attempts=0, rng=0):
    from replica_exchange_wang_landau_amd import RunConfig, WlParameters

    return RunConfig(emin=C2_GRID[0], emax=C2_GRID[1], bin_width=C2_GRID[2], windows=8,
                     walkers_per_window=16, overlap=0.75, seed=1234,
                     exchange_interval_attempts=512, max_attempts=max_attempts, rng=rng,
                     wl=WlParameters(final_factor=1e-7, check_interval_attempts=512 * 100))


def gpu_leg(steps: int, warmup: int, rng: int):
    """Time `steps` exchange epochs after `warmup` untimed ones."""
    from replica_exchange_wang_landau_amd import Engine, System

    system = System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True)
    engine = Engine(system, c2_config(rng=rng))
    engine.initialize()
    engine.run(warmup)
    before = engine.status()
    engine.set_timing(True)
    t0 = time.perf_counter()
    after = engine.run(steps)  # returns after the device has finished (status read)
    t1 = time.perf_counter()
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
        "achieved_gbs": achieved, "per_launch_bytes": per_launch_bytes,
        "ambiguous": after.ambiguous_decisions,
    }


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
    return {"value": attempted / seconds, "unit": "attempted flips/s", "cores": threads,
            "kind": kind,
            "sample": f"run_rewl on the same workload, {mcs_per_walker} MCS per walker "
                      f"({attempted} attempts incl. walker init) in {seconds:.2f} s"}


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
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only run a plain copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU only; there is no multi-GPU path "
              "(the walkers would shard by energy window across GPUs)", file=sys.stderr)
        sys.exit(2)
    rng = 0 if args.rng == "xoshiro" else 1
    # N = 512 sites: 8 register slots per lane; Philox without trace, forcing or the
    # Nalivaiko mask runs the common-configuration lookahead kernel
    kernel = "sample_kernel<8,false>" if rng == 0 else "lookahead_fast_kernel<8>"
    g = gpu_leg(args.steps, args.warmup, rng)
    other = gpu_leg(args.steps, args.warmup, 1 - rng)
    cpu = None if args.no_cpu else cpu_leg(args.cpu_mcs)
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
                "value": other["rate"], "kernel_ms": other["kernel_ms"],
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
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
