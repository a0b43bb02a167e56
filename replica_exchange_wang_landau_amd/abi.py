"""ctypes view of the engine's C ABI (include/rewl.h).

The structs mirror the header member for member (all 8-byte fields, no padding). `EXPORTS`
lists every entry point the header declares; the library must export all of them.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
LIBRARY_PATH = PACKAGE_DIR / "librewl.so"

WLX_OK = 0
WLX_INVALID_ARGUMENT = 1
WLX_RUNTIME_ERROR = 2
WLX_LOGIC_ERROR = 3
WLX_OUT_OF_RANGE = 4
WLX_DEVICE_ERROR = 5

COUPLINGS_FP64 = 0
COUPLINGS_FP32 = 1
STAGE_WANG_LANDAU = 0
STAGE_INVERSE_TIME = 1
STAGE_FROZEN = 2

RNG_XOSHIRO = 0
RNG_PHILOX = 1
FIELDS_FP64 = 0
FIELDS_FP32_KAHAN = 1

KERNEL_AUTO = 0
KERNEL_LOOKAHEAD = 1
KERNEL_SEQUENTIAL = 2
KERNEL_LDS = 3
KERNEL_STAGED = 4

TRACE_ACCEPTED = 1
TRACE_FORCED = 2
TRACE_UNIFORM = 4
TRACE_IN_WINDOW = 8

c_u64 = ctypes.c_uint64
c_f64 = ctypes.c_double


class Params(ctypes.Structure):
    """WlParameters (wang_landau.hpp:40-56)."""

    _fields_ = [
        ("flatness", c_f64),
        ("minimum_visits", c_u64),
        ("final_factor", c_f64),
        ("check_interval_attempts", c_u64),
        ("force_accept_after_attempts", c_u64),
        ("inverse_time_enabled", c_u64),
        ("initialization_max_attempts", c_u64),
        ("initialization_target_fraction", c_f64),
        ("initialization_temperature_fraction", c_f64),
        ("initialization_stall_attempts_per_spin", c_u64),
        ("initialization_temperature_multiplier", c_f64),
        ("initialization_max_temperature_fraction", c_f64),
        ("nalivaiko_mod", c_u64),
        ("collect_window_statistics", c_u64),
        ("round_trip_margin_fraction", ctypes.c_double),
    ]


class Config(ctypes.Structure):
    """The run_rewl part of RunConfig (io.hpp:12-58), intervals resolved to attempts."""

    _fields_ = [
        ("emin", c_f64),
        ("emax", c_f64),
        ("bin_width", c_f64),
        ("windows", c_u64),
        ("walkers_per_window", c_u64),
        ("overlap", c_f64),
        ("seed", c_u64),
        ("exchange_interval_attempts", c_u64),
        ("max_attempts", c_u64),
        ("rng", c_u64),
        ("fields", c_u64),
        ("explicit_windows", ctypes.POINTER(c_u64)),
        ("wl", Params),
        ("resync_interval_attempts", c_u64),
    ]


class WalkerInfo(ctypes.Structure):
    _fields_ = [
        ("walker_id", c_u64),
        ("window", c_u64),
        ("window_begin", c_u64),
        ("window_end", c_u64),
        ("energy", c_f64),
        ("factor", c_f64),
        ("attempted", c_u64),
        ("accepted", c_u64),
        ("forced_accepted", c_u64),
        ("last_accepted_attempt", c_u64),
        ("stage", c_u64),
        ("active_bins", c_u64),
        ("refinement_active_bins", c_u64),
        ("rng_state", c_u64 * 4),
        ("last_flatness_check", c_u64),
        ("last_flatness", c_f64),
        ("clock_offset", c_u64),
    ]


class HistogramStats(ctypes.Structure):
    _fields_ = [
        ("active_bins", c_u64),
        ("covered_bins", c_u64),
        ("minimum", c_u64),
        ("mean", c_f64),
        ("min_over_mean", c_f64),
        ("coverage", c_f64),
    ]


class RunStatus(ctypes.Structure):
    _fields_ = [
        ("epoch", c_u64),
        ("done", c_u64),
        ("converged", c_u64),
        ("attempted", c_u64),
        ("accepted", c_u64),
        ("forced_accepted", c_u64),
        ("exchange_attempted", c_u64),
        ("exchange_accepted", c_u64),
        ("ambiguous_decisions", c_u64),
    ]


EXPORTS = [
    "wlx_last_error",
    "wlx_system_create",
    "wlx_system_destroy",
    "wlx_system_info",
    "wlx_system_read_dense",
    "wlx_system_read_csr",
    "wlx_system_fields",
    "wlx_system_fields_noalloc",
    "wlx_system_pilot",
    "wlx_systems_pilot",
    "wlx_device_info",
    "wlx_engine_create",
    "wlx_engine_destroy",
    "wlx_engine_initialize",
    "wlx_engine_sample",
    "wlx_engine_check",
    "wlx_engine_exchange",
    "wlx_engine_run",
    "wlx_engine_status",
    "wlx_engine_set_graph",
    "wlx_engine_set_run_state",
    "wlx_engine_walkers",
    "wlx_engine_layout",
    "wlx_engine_read_walker",
    "wlx_engine_read_infos",
    "wlx_engine_write_walker",
    "wlx_engine_set_walker_cadence",
    "wlx_engine_resync",
    "wlx_engine_drift",
    "wlx_engine_read_sampling",
    "wlx_engine_warm_start",
    "wlx_engine_set_kernel",
    "wlx_engine_kernel",
    "wlx_engine_autotune",
    "wlx_engine_set_lookahead_width",
    "wlx_batch_create",
    "wlx_batch_destroy",
    "wlx_batch_set_graph",
    "wlx_batch_run",
    "wlx_batch_kernel",
    "wlx_batch_set_lds_budget",
    "wlx_batch_layout",
    "wlx_batch_autotune",
    "wlx_mlp_params",
    "wlx_mlp_create",
    "wlx_mlp_destroy",
    "wlx_mlp_get_params",
    "wlx_mlp_set_params",
    "wlx_mlp_predict",
    "wlx_mlp_gradient",
    "wlx_mlp_train",
    "wlx_engine_histogram_statistics",
    "wlx_engine_ready",
    "wlx_engine_next_iteration",
    "wlx_engine_trace",
    "wlx_engine_read_trace",
    "wlx_engine_set_timing",
    "wlx_engine_kernel_time",
]


def header_declarations(header: Path | None = None) -> list[str]:
    """Names of the entry points declared in include/rewl.h."""
    import re

    header = header or PACKAGE_DIR.parent / "include" / "rewl.h"
    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(wlx_\w+)\s*\(", text, re.M)))


_LIBRARY: ctypes.CDLL | None = None


def load_library(path: os.PathLike | None = None) -> ctypes.CDLL:
    """Loads librewl.so. Fails loudly when it has not been built: there is no fallback."""
    global _LIBRARY
    if path is None and _LIBRARY is not None:
        return _LIBRARY
    target = Path(path) if path else LIBRARY_PATH
    if not target.exists():
        raise ImportError(
            f"{target} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = ctypes.CDLL(str(target))
    lib.wlx_last_error.restype = ctypes.c_size_t
    lib.wlx_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    if path is None:
        _LIBRARY = lib
    return lib
