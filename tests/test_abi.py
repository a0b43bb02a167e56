"""CPU checks of the C ABI boundary: the library loads and exports every declared entry point,
and the ctypes structs have the header's layout (no GPU calls)."""
from __future__ import annotations

import ctypes
import subprocess
import tempfile
from pathlib import Path

import pytest

from replica_exchange_wang_landau_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    declared = abi.header_declarations()
    assert declared == sorted(abi.EXPORTS)
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing


def test_struct_layouts_match_header():
    source = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "rewl.h"
    int main(void) {
        printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(wlx_params), sizeof(wlx_config),
               sizeof(wlx_walker_info), sizeof(wlx_histogram_stats), sizeof(wlx_run_status),
               offsetof(wlx_config, explicit_windows), offsetof(wlx_config, wl),
               offsetof(wlx_walker_info, rng_state));
        return 0;
    }
    """
    with tempfile.TemporaryDirectory() as tmp:
        src = Path(tmp) / "layout.c"
        exe = Path(tmp) / "layout"
        src.write_text(source)
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
        sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                                check=True).stdout.split()]
    assert sizes == [ctypes.sizeof(abi.Params), ctypes.sizeof(abi.Config),
                     ctypes.sizeof(abi.WalkerInfo), ctypes.sizeof(abi.HistogramStats),
                     ctypes.sizeof(abi.RunStatus), abi.Config.explicit_windows.offset,
                     abi.Config.wl.offset, abi.WalkerInfo.rng_state.offset]


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        abi.load_library(tmp_path / "librewl.so")


def test_library_is_built_for_gfx950():
    lib = abi.LIBRARY_PATH.read_bytes()
    assert b"gfx950" in lib
