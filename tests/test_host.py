"""CPU tests of the host C++ shell (configuration, windows, checkpoint layout, summaries,
stitching, output formats) and of the wl_run command line that needs no GPU."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "replica_exchange_wang_landau_amd" / "bin"


def run(*args, **kw):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, **kw)


def test_host_selftest():
    out = run(BIN / "host_selftest")
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
