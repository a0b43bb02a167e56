"""Estimates the error of tests/golden/c2_reference_dos.json and writes
tests/golden/c2_reference_dos_error.json.

The C2 reference DOS is one long run (ln f = 1e-7, seed 20261203); its exact DOS is out of
reach. This script runs `--runs` further independent REWL runs of the same system and grid
(Philox stream, seeds 1..runs), split into 48 windows of 16 walkers (the layout that converges
fastest per attempt), all walkers frozen at ln f <= `--final` (1e-6 by default), batched into
one launch per epoch. It reports the largest pairwise max|dlog g| between the independent runs
(each with the mean offset removed over the bins valid in both) and each run's max|dlog g|
against the reference. The reference, run 10x further down in ln f, is closer to the truth
than any of these runs, so their spread bounds its error from above.

    python tests/golden/estimate_c2_reference_error.py --runs 4 --final 1e-6
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

from replica_exchange_wang_landau_amd import (Batch, Engine, RunConfig, System,  # noqa: E402
                                              WlParameters, abi, analysis)

REFERENCE = Path(__file__).resolve().parent / "c2_reference_dos.json"
OUT = Path(__file__).resolve().parent / "c2_reference_dos_error.json"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=4)
    ap.add_argument("--final", type=float, default=1e-6)
    ap.add_argument("--windows", type=int, default=48)
    ap.add_argument("--seconds", type=float, default=900.0)
    ap.add_argument("--out", default=str(OUT))
    args = ap.parse_args()
    ref = json.loads(REFERENCE.read_text())
    ref_lg = np.array([float.fromhex(v) for v in ref["log_g"]])
    ref_valid = np.array(ref["valid"], dtype=np.uint8)
    emin, emax, width = ref["grid"]
    system = System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True)
    engines = []
    for seed in range(1, args.runs + 1):
        cfg = RunConfig(emin=emin, emax=emax, bin_width=width, windows=args.windows,
                        walkers_per_window=16, overlap=0.75, seed=seed,
                        exchange_interval_attempts=512, max_attempts=0, rng=abi.RNG_PHILOX,
                        wl=WlParameters(final_factor=args.final,
                                        check_interval_attempts=512 * 100))
        e = Engine(system, cfg)
        e.initialize()
        engines.append(e)
    batch = Batch(engines)
    batch.set_graph(16)
    t0 = time.perf_counter()
    done, last = 0, t0
    while done < len(engines) and time.perf_counter() - t0 < args.seconds:
        _, done = batch.run(2048)
        if time.perf_counter() - last > 20.0:
            last = time.perf_counter()
            factors = [max(e.read_walker(k).factor for k in range(0, e.walkers, 97))
                       for e in engines]
            print(f"{last - t0:.0f} s: {done}/{len(engines)} runs frozen, largest ln f "
                  f"{['%.2g' % f for f in factors]}", flush=True)
    seconds = time.perf_counter() - t0
    curves = [analysis.engine_dos(e) for e in engines]
    pairwise = [analysis.max_deviation(*curves[i], *curves[j])
                for i in range(len(curves)) for j in range(i + 1, len(curves))]
    vs_ref = [analysis.max_deviation(lg, v, ref_lg, ref_valid) for lg, v in curves]
    result = {
        "description": "spread of independent C2 runs (48 windows x 16 walkers, frozen at "
                       "ln f <= final) and their distance to the ln f = 1e-7 reference, as "
                       "max|dlog g| after mean-offset alignment",
        "runs": args.runs, "final_factor": args.final, "windows": args.windows,
        "frozen_runs": int(done), "seconds": seconds,
        "attempted": [int(e.status().attempted) for e in engines],
        "pairwise_max_dev": pairwise, "max_pairwise_max_dev": max(pairwise),
        "vs_reference_max_dev": vs_ref, "max_vs_reference": max(vs_ref),
    }
    Path(args.out).write_text(json.dumps(result, indent=1) + "\n")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
