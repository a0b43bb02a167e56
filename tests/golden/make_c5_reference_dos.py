"""Writes tests/golden/c5_reference_dos.json: the converged density of states of the held-out
slab of the learned warm start (12 x 12 x 4 sites, periodic in x and y, open in z, site
dilution 0.2, +z axes, dipolar D = 1, CSR cutoff 3; warmstart.Slab(12, 4, 0.2, 0, 999001)),
and how far other realizations of the same slab parameters lie from it.

The slab has about 460 spins, so its exact DOS is out of reach. The reference is the average
(analysis.average_dos) of --runs independent REWL runs of this engine (Philox stream, seeds
1..runs) frozen at ln f <= --final (1e-6 by default), on the grid and windows of
warmstart.slab_config from a 2000-MCS pilot (seed 1234). The record holds the grid, the
averaged stitched log g, the largest pairwise max|dlog g| between the runs (the reference's
error is below that spread) and each run's distance to the average.

Three more realizations of the same parameters (dilution seeds 999002..999004) run to
--sibling-final (1e-5) in the same batch. Each one's log g per spin, interpolated linearly in
E/N onto the held-out slab's bins and scaled by the held-out N, is compared with the
reference (mean offset removed): a network that sees only (E/N, N, density, dilution,
thickness, axis mode) cannot tell realizations apart, so this spread is a floor under its
warm-start error on the held-out slab.

    python tests/golden/make_c5_reference_dos.py --runs 4 --final 1e-6
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

from replica_exchange_wang_landau_amd import Batch, Engine, System, analysis  # noqa: E402
from replica_exchange_wang_landau_amd import warmstart as ws  # noqa: E402

OUT = Path(__file__).resolve().parent / "c5_reference_dos.json"
HELD_OUT = ws.Slab(12, 4, 0.2, 0, 999_001)
SIBLINGS = [ws.Slab(12, 4, 0.2, 0, 999_001 + k) for k in (1, 2, 3)]


def per_spin_curve(engine):
    """(E/N at the valid bins, log g / N there) of a converged engine."""
    lg, valid = analysis.engine_dos(engine)
    ok = valid != 0
    centers = engine.config.emin + (np.arange(engine.bins) + 0.5) * engine.config.bin_width
    return centers[ok] / engine.n, lg[ok] / engine.n


def realization_spread(held_e, held_lg, held_valid, n_held, sibling):
    """max and RMS over the shared E/N range of N_held * (s_sibling - s_held), mean offset
    removed, with the sibling's log g per spin interpolated linearly in E/N."""
    e_sib, s_sib = sibling
    centers = held_e[held_valid != 0]
    ref = held_lg[held_valid != 0]
    inside = (centers >= e_sib.min()) & (centers <= e_sib.max())
    d = n_held * np.interp(centers[inside], e_sib, s_sib) - ref[inside]
    d -= d.mean()
    return float(np.max(np.abs(d))), float(np.sqrt(np.mean(d * d))), int(inside.sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=4)
    ap.add_argument("--final", type=float, default=1e-6)
    ap.add_argument("--sibling-final", type=float, default=1e-5)
    ap.add_argument("--seconds", type=float, default=1000.0)
    ap.add_argument("--out", default=str(OUT))
    args = ap.parse_args()
    slabs = [HELD_OUT] + SIBLINGS
    systems = [s.system() for s in slabs]
    lo, hi = System.pilot_many(systems, 1234, [2000 * s.n for s in systems])
    engines = []
    for seed in range(1, args.runs + 1):
        e = Engine(systems[0], ws.slab_config(systems[0], lo[0], hi[0], seed, args.final))
        e.initialize()
        engines.append(e)
    for k in range(1, len(slabs)):
        e = Engine(systems[k], ws.slab_config(systems[k], lo[k], hi[k], 100 + k,
                                              args.sibling_final))
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
            factors = [max(e.read_walker(k).factor for k in range(0, e.walkers, 8))
                       for e in engines]
            print(f"{last - t0:.0f} s: {done}/{len(engines)} runs frozen, largest ln f "
                  f"{['%.2g' % f for f in factors]}", flush=True)
    seconds = time.perf_counter() - t0
    if done < len(engines):
        print(f"not converged within {args.seconds} s", flush=True)
        sys.exit(1)
    runs = engines[:args.runs]
    curves = [analysis.engine_dos(e) for e in runs]
    avg_lg, avg_valid = analysis.average_dos(curves)
    pairwise = [analysis.max_deviation(*curves[i], *curves[j])
                for i in range(len(curves)) for j in range(i + 1, len(curves))]
    vs_avg = [analysis.max_deviation(lg, v, avg_lg, avg_valid) for lg, v in curves]
    cfg = runs[0].config
    centers = cfg.emin + (np.arange(runs[0].bins) + 0.5) * cfg.bin_width
    n_held = systems[0].n
    siblings = []
    for slab, e in zip(SIBLINGS, engines[args.runs:]):
        worst, rms, shared = realization_spread(centers / n_held, avg_lg, avg_valid, n_held,
                                                per_spin_curve(e))
        siblings.append({"dilution_seed": slab.seed, "spins": int(e.n),
                         "max_dev": worst, "rms_dev": rms, "shared_bins": shared})
    result = {
        "description": "held-out slab 12x12x4, dilution 0.2 (seed 999001), +z axes, CSR cutoff "
                       "3: average of independent REWL runs frozen at ln f <= final, stitched "
                       "log g with the peak at 0; siblings: other dilution realizations, "
                       "N_held * per-spin log g difference on the shared E/N range",
        "slab": [HELD_OUT.lx, HELD_OUT.lz, HELD_OUT.dilution, HELD_OUT.axis_mode,
                 HELD_OUT.seed],
        "spins": int(n_held),
        "grid": [cfg.emin, cfg.emax, cfg.bin_width], "windows": cfg.windows,
        "walkers_per_window": cfg.walkers_per_window, "pilot": [float(lo[0]), float(hi[0])],
        "runs": args.runs, "final_factor": args.final, "seconds": seconds,
        "attempted": [int(e.status().attempted) for e in runs],
        "pairwise_max_dev": pairwise, "max_pairwise_max_dev": max(pairwise),
        "vs_average_max_dev": vs_avg,
        "log_g": [v.hex() for v in avg_lg.tolist()],
        "valid": [int(v) for v in avg_valid.tolist()],
        "sibling_final_factor": args.sibling_final,
        "siblings": siblings,
    }
    Path(args.out).write_text(json.dumps(result) + "\n")
    print(json.dumps({k: v for k, v in result.items() if k not in ("log_g", "valid")}),
          flush=True)
    batch.close()


if __name__ == "__main__":
    main()
