"""CPU tests of the host C++ shell (configuration, windows, checkpoint layout, summaries,
stitching, output formats) and of the wl_run command line that needs no GPU."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "replica_exchange_wang_landau_amd" / "bin"
CHECKS = ROOT / "tests" / "host" / "_bin" / "host_checks"


def run(*args, **kw):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, **kw)


def test_host_checks():
    out = run(CHECKS)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ALL PASS" in out.stdout


def test_wl_run_help_and_usage_errors(tmp_path):
    out = run(BIN / "wl_run", "--help")
    assert out.returncode == 0 and out.stdout.startswith("Usage: ")
    out = run(BIN / "wl_run", "--bogus", "1", cwd=tmp_path)
    assert out.returncode == 1 and out.stderr == "wl_run: Unknown option: --bogus\n"
    out = run(BIN / "wl_run", "--seed", "3", cwd=tmp_path)
    assert out.returncode == 1
    assert out.stderr.endswith("wl_run: Production runs require --emin, --emax, and --bin-width\n")


def test_wl_run_rejects_bad_geometry_files(tmp_path):
    csv = tmp_path / "bad.csv"
    csv.write_text("x,y,z,mx,my\n0,0,0,0,0\n")
    out = run(BIN / "wl_run", "--geometry", csv, "--emin", "-4", "--emax", "4", "--bin-width",
              "0.25", "--seed", "1", cwd=tmp_path)
    assert out.returncode == 1
    assert out.stderr == "wl_run: CSV geometry header must be x,y,z,mx,my,mz\n"
    csv.write_text("x,y,z,mx,my,mz\n0,0,0,0,0,1\n1,0,0,0,0,2\n")
    out = run(BIN / "wl_run", "--geometry", csv, "--emin", "-4", "--emax", "4", "--bin-width",
              "0.25", "--seed", "1", cwd=tmp_path)
    assert out.stderr == "wl_run: Ising axes must be normalized\n"


def host_sites(path, box, periodic):
    """The host loaders (host_checks geometry mode): (positions, axes) or the error text."""
    out = run(CHECKS, "geometry", path, *[repr(float(b)) for b in box], "1" if periodic else "0")
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    if lines[0].startswith("error "):
        return lines[0][len("error "):]
    n = int(lines[0].split()[1])
    values = np.array([[float.fromhex(v) for v in line.split()] for line in lines[1:1 + n]])
    values = values.reshape(n, 6)
    return values[:, :3], values[:, 3:]


DAT_CASES = {
    # comments, blank and indented lines, CRLF endings, tabs, extra columns (ignored by both),
    # axes of any length (normalised on load), scientific and negative numbers
    "plain.dat": "0 0 0 0 0 1\n1 0 0 0 0 2\n0 1 0 3 4 0\n",
    "comments.dat": "# sites\n\n   # indented comment\n0.5 0.25 0 1 1 1\r\n\t2 2 2 0 -1e-3 7\r\n",
    "extra_columns.dat": "0 0 0 0 0 1 99 junk\n1.5e0 -2.25 3 1 0 0\n",
    "skewed_axes.dat": "0 0 0 0.1 0.2 0.3\n1 1 1 -5 0 12\n2 0 1 1e-8 1e-8 1\n",
    # error paths: a short line, a non-number, a zero axis (normalise/validate), coincident
    # positions are allowed by the loader (the pair build rejects them)
    "short_line.dat": "0 0 0 0 0 1\n1 0 0 0 1\n",
    "not_a_number.dat": "0 0 0 0 0 1\nx 0 0 0 0 1\n",
    "zero_axis.dat": "0 0 0 0 0 0\n",
    "empty.dat": "# nothing\n",
}


@pytest.mark.parametrize("name", sorted(DAT_CASES))
@pytest.mark.parametrize("box,periodic", [((0.0, 0.0, 0.0), False), ((4.0, 4.0, 0.0), True)])
def test_dat_loader_matches_reference(ref, tmp_path, name, box, periodic):
    """Geometry::load_xyz_axes (physics.cpp:77-97): the host loader reads the whitespace .dat
    format into the same sites, bit for bit (axes normalised), and fails with the same text."""
    path = tmp_path / name
    path.write_bytes(DAT_CASES[name].encode())
    mine, theirs = host_sites(path, box, periodic), ref.load_geometry(path, box, periodic)
    if isinstance(theirs, str):
        assert mine == theirs
        return
    assert not isinstance(mine, str), mine
    assert mine[0].tobytes() == theirs[0].tobytes() and mine[1].tobytes() == theirs[1].tobytes()
    assert np.allclose(np.linalg.norm(mine[1], axis=1), 1.0, rtol=0, atol=1e-15)


def test_csv_loader_matches_reference(ref, tmp_path):
    """Geometry::load_csv (physics.cpp:99-149) on a file with a BOM, comments and spaces."""
    path = tmp_path / "sites.csv"
    path.write_bytes("﻿X, Y ,z,MX,my,mz\n# c\n0,0,0,0,0,1\n 1.25 ,0,0,0.6,0,0.8\n".encode())
    mine, theirs = host_sites(path, (0, 0, 0), False), ref.load_geometry(path, (0, 0, 0), False,
                                                                         csv=True)
    assert not isinstance(theirs, str), theirs
    assert mine[0].tobytes() == theirs[0].tobytes() and mine[1].tobytes() == theirs[1].tobytes()
