"""wl_run on the MI355X engine against the reference's own wl_run (compiled from its sources
into oracle/_ref): identical stdout lines (elapsed time aside), identical output files byte
for byte (_dos, _thermo, _window_*, _workers_stat, the WLCHKP3 checkpoint) and identical
metadata apart from the engine object this build adds and the host thread count."""
from __future__ import annotations

import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
OURS = ROOT / "replica_exchange_wang_landau_amd" / "bin" / "wl_run"
REF = ROOT / "oracle" / "_ref" / "wl_run"

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not REF.exists(), reason="compiled reference wl_run not built")]


def run(exe: Path, args, cwd: Path):
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "4"))
    return subprocess.run([str(exe), *map(str, args)], cwd=cwd, capture_output=True, text=True,
                          env=env, timeout=600)


def normalise_stdout(text: str) -> str:
    text = re.sub(r"adaptive_window_elapsed_seconds=\S+", "", text)
    return re.sub(r" elapsed_seconds=\S+", "", text)


def normalise_metadata(text: str, directory: Path) -> str:
    text = text.replace(str(directory), "DIR")
    text = re.sub(r'"openmp_threads": \d+', '"openmp_threads": X', text)
    return re.sub(r',\n  "engine": \{[^\n]*\}', "", text)


def compare(tmp_path: Path, args, files=None, extra=None):
    a, b = tmp_path / "ref", tmp_path / "ours"
    a.mkdir(exist_ok=True)
    b.mkdir(exist_ok=True)
    for src in extra or []:
        for d in (a, b):
            (d / src.name).write_bytes(src.read_bytes())
    ra, rb = run(REF, args, a), run(OURS, args, b)
    assert rb.returncode == ra.returncode, (ra.stderr, rb.stderr)
    assert normalise_stdout(rb.stdout) == normalise_stdout(ra.stdout)
    assert rb.stderr == ra.stderr
    names = sorted(p.name for p in a.iterdir())
    assert sorted(p.name for p in b.iterdir()) == names
    for name in names:
        x, y = (a / name).read_bytes(), (b / name).read_bytes()
        if name.endswith("_metadata.json"):
            assert normalise_metadata(y.decode(), b) == normalise_metadata(x.decode(), a), name
            assert '"engine": {"implementation": "mi355x-hip"' in y.decode()
        else:
            assert x == y, name
    return ra, rb


LATTICE_2x2x2 = ["--nx", 2, "--ny", 2, "--nz", 2, "--periodic", "false", "--emin", -20, "--emax",
                 20, "--bin-width", 0.25]


def test_c1_plain_wl_converges_identically(tmp_path):
    ra, rb = compare(tmp_path, LATTICE_2x2x2 + ["--walkers", 2, "--seed", 1234, "--final-factor",
                                                "1e-4", "--max-mcs", 0])
    assert "converged=yes" in rb.stdout


def test_two_window_rewl_unconverged_writes_worker_statistics(tmp_path):
    ra, rb = compare(tmp_path, LATTICE_2x2x2 + ["--windows", 2, "--walkers", 3, "--seed", 77,
                                                "--exchange-interval-mcs", 2, "--max-mcs", 3000,
                                                "--output", "two"])
    assert rb.returncode == 2 and (tmp_path / "ours" / "two_workers_stat.csv").exists()


def test_smoke_test_preset(tmp_path):
    compare(tmp_path, ["--smoke-test", "--seed", 5])


def test_pilot_range(tmp_path):
    ra, rb = compare(tmp_path, ["--nx", 8, "--ny", 8, "--nz", 8, "--box-x", 8, "--pilot",
                                "--max-mcs", 10000, "--seed", 1234])
    assert "sampled_min=-244.261 sampled_max=325.76" in rb.stdout


def test_csv_geometry_ini_csr_and_nalivaiko(tmp_path):
    rng = np.random.default_rng(4)
    grid = np.stack(np.meshgrid(np.arange(4.0), np.arange(4.0), np.arange(2.0), indexing="ij"),
                    -1).reshape(-1, 3)
    pos = grid + rng.uniform(-0.2, 0.2, size=grid.shape)
    n = len(pos)
    axes = rng.normal(size=(n, 3))
    axes /= np.sqrt((axes * axes).sum(axis=1))[:, None]
    csv = tmp_path / "system.csv"
    csv.write_text("x,y,z,mx,my,mz\n" + "".join(
        ",".join(repr(float(v)) for v in (*p, *a)) + "\n" for p, a in zip(pos, axes)))
    ini = tmp_path / "run.ini"
    ini.write_text("[geometry]\nfile = system.csv\nperiodic = true\nbox_x = 4.0\nbox_y = 4.0\n"
                   "box_z = 0.0\n[physics]\ncutoff = 2.5\n[energy]\nminimum = -60\n"
                   "maximum = 60\nbin_width = 1.0\n[parallel]\nwindows = 3\nwalkers = 2\n"
                   "exchange_interval_mcs = 1\n[wl]\nnalivaiko_mod = true\ninverse_time = false\n"
                   "flatness = 0.5\nminimum_visits = 5\n[run]\nseed = 11\nmax_mcs = 400\n"
                   "[output]\nprefix = csvrun\n")
    compare(tmp_path, ["--config", "run.ini"], extra=[csv, ini])


def test_checkpoint_write_and_resume(tmp_path):
    args = LATTICE_2x2x2 + ["--seed", 9, "--max-mcs", 500, "--checkpoint", "walker.chk",
                            "--checkpoint-interval-mcs", 100, "--output", "ck"]
    compare(tmp_path, args)
    resumed = [900 if (i > 0 and args[i - 1] == "--max-mcs") else a for i, a in enumerate(args)]
    compare(tmp_path, resumed)


@pytest.mark.parametrize("args", [
    ["--nx", 3, "--ny", 3, "--nz", 3, "--periodic", "true", "--emin", -60, "--emax", 80,
     "--bin-width", 0.5, "--windows", 3, "--walkers", 2, "--seed", 21,
     "--adaptive-iterations", 2, "--adaptive-pilot-mcs", 200, "--max-mcs", 2000,
     "--exchange-interval-mcs", 2],
    ["--nx", 4, "--ny", 4, "--nz", 4, "--periodic", "true", "--emin", -120, "--emax", 160,
     "--bin-width", 1.0, "--windows", 4, "--walkers", 3, "--overlap", 0.5, "--seed", 8,
     "--adaptive-iterations", 3, "--adaptive-pilot-mcs", 300, "--adaptive-round-trip-margin",
     0.2, "--max-mcs", 1500],
])
def test_adaptive_windows_identical(tmp_path, args):
    """--adaptive-windows: pilots whose statistics are collected by the device walkers, the
    window planner and warm-start selection on the host. Identical pilot lines (round trips,
    seeded walkers, warm starts, energy ranges), final windows and outputs."""
    ra, rb = compare(tmp_path, args + ["--adaptive-windows", "true", "--output", "ad"])
    assert rb.stdout.count("adaptive_window_round_trips=") == args[args.index(
        "--adaptive-iterations") + 1]


def test_multi_walker_checkpoint_resume_identical(tmp_path):
    """A 2x3 run checkpointed (WLCHKM2: every walker's WLCHKP3 record, the driver's cadence
    and the exchange counters) at 600 MCS and resumed to 1000 MCS writes the same DOS,
    window, thermodynamics and walker-statistics files as an uninterrupted 1000-MCS run. The
    reference checkpoints 1x1 runs only (parallel.cpp:354,522), so this compares this
    engine's two runs."""
    base = ["--nx", 3, "--ny", 3, "--nz", 3, "--periodic", "true", "--emin", -60, "--emax", 80,
            "--bin-width", 0.5, "--windows", 2, "--walkers", 3, "--seed", 5,
            "--exchange-interval-mcs", 2, "--output", "run"]
    a, b = tmp_path / "resumed", tmp_path / "straight"
    a.mkdir()
    b.mkdir()
    first = run(OURS, base + ["--max-mcs", 600, "--checkpoint", "run.chk",
                              "--checkpoint-interval-mcs", 100], a)
    assert first.returncode == 2, first.stderr
    assert (a / "run.chk").read_bytes()[:8] == b"WLCHKM2\0"
    resumed = run(OURS, base + ["--max-mcs", 1000, "--checkpoint", "run.chk",
                                "--checkpoint-interval-mcs", 100], a)
    straight = run(OURS, base + ["--max-mcs", 1000, "--checkpoint-interval-mcs", 100], b)
    assert "warning: checkpoint was not restored" not in resumed.stderr
    assert normalise_stdout(resumed.stdout) == normalise_stdout(straight.stdout)
    for name in ("run_dos.csv", "run_window_0.csv", "run_window_1.csv", "run_thermo.csv",
                 "run_workers_stat.csv"):
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
