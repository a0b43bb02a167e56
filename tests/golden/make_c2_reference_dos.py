"""Writes tests/golden/c2_reference_dos.json: the converged C2 density of states that
time-to-accuracy is measured against.

C2 is the 8x8x8 periodic dense system with 8 windows x 16 walkers, exchanges every MCS and
the 1/t stage down to ln f = 1e-7. Its exact DOS is out of reach (2^512 states), so the
reference is a long plain-REWL run of this engine to convergence: every walker frozen at
ln f <= 1e-7. It uses its own seed, so it is independent of the runs measured against it. The run takes
roughly 6e11 attempts, which needs several GPU sessions. Each invocation runs for at most
--seconds, checkpoints every walker (Engine.save_checkpoint) and resumes from that checkpoint
on the next call, restarting each configuration's energy and fields from an FP64 recompute. When all walkers are frozen, the stitched log g (stitch_dos normalisation:
peak 0) and its valid mask are written.

    python tests/golden/make_c2_reference_dos.py --state STATE.npz --seconds 900
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from replica_exchange_wang_landau_amd import Engine, RunConfig, System, WlParameters, abi  # noqa
from replica_exchange_wang_landau_amd import analysis  # noqa: E402

OUT = Path(__file__).resolve().parent / "c2_reference_dos.json"
SEED = 20261203


def c2_reference_config() -> RunConfig:
    return RunConfig(emin=-273.0, emax=354.5, bin_width=0.5, windows=8, walkers_per_window=16,
                     overlap=0.75, seed=SEED, exchange_interval_attempts=512,
                     max_attempts=0, rng=abi.RNG_PHILOX,
                     wl=WlParameters(final_factor=1e-7, check_interval_attempts=512 * 100))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--state", required=True, help="checkpoint to resume from and write")
    ap.add_argument("--seconds", type=float, default=900.0)
    ap.add_argument("--out", default=str(OUT))
    args = ap.parse_args()
    cfg = c2_reference_config()
    engine = Engine(System.simple_cubic(8, 8, 8, 1.0, (0.0, 0.0, 1.0), True), cfg)
    engine.initialize()
    state = Path(args.state)
    if state.exists():
        engine.load_checkpoint(state, recompute_fields=True)
        print(f"resumed from {state}", flush=True)
    t0 = time.perf_counter()
    last_print = t0
    st = engine.status()
    while not st.done and time.perf_counter() - t0 < args.seconds:
        st = engine.run(4096)
        now = time.perf_counter()
        if now - last_print > 30.0:
            infos = [engine.read_walker(k).info for k in range(0, engine.walkers, 16)]
            stages = [int(i.stage) for i in infos]
            factors = [f"{i.factor:.3g}" for i in infos]
            print(f"epoch={st.epoch} attempted={st.attempted} stages={stages} "
                  f"factors={factors} elapsed={now - t0:.0f}s", flush=True)
            last_print = now
    engine.save_checkpoint(state)
    print(f"saved {state}: epoch={st.epoch} attempted={st.attempted} done={st.done} "
          f"converged={st.converged}", flush=True)
    if st.converged:
        log_g, valid = analysis.engine_dos(engine)
        Path(args.out).write_text(json.dumps({
            "description": "C2 8x8x8 periodic dense, 8 windows x 16 walkers, exchange every MCS, "
                           "1/t to ln f = 1e-7, Philox stream, converged plain REWL; stitched "
                           "log g with the peak at 0",
            "grid": [cfg.emin, cfg.emax, cfg.bin_width], "seed": SEED,
            "attempted": int(st.attempted), "epochs": int(st.epoch),
            "log_g": [v.hex() for v in log_g.tolist()],
            "valid": [int(v) for v in valid.tolist()],
        }) + "\n")
        print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
