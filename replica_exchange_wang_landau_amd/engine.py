"""Python face of the engine's C ABI: coupling systems and device-resident REWL walkers.

`System` mirrors the reference's Couplings backends (physics.hpp:41-84) and `Engine` the
walker/driver API the reference tests exercise (WangLandauWalker, wang_landau.hpp:91-138;
run_rewl, parallel.hpp:35-36). Both call through a `Backend`, by default the HIP library
(librewl.so). Errors raise the Python counterpart of the reference's exception class.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, fields as dc_fields
from typing import Any, Sequence

import numpy as np

from . import abi


class InvalidArgument(ValueError):
    """std::invalid_argument"""


class LogicError(RuntimeError):
    """std::logic_error"""


class DeviceError(RuntimeError):
    """HIP runtime failure (no reference counterpart)"""


_ERRORS = {
    abi.WLX_INVALID_ARGUMENT: InvalidArgument,
    abi.WLX_RUNTIME_ERROR: RuntimeError,
    abi.WLX_LOGIC_ERROR: LogicError,
    abi.WLX_OUT_OF_RANGE: IndexError,
    abi.WLX_DEVICE_ERROR: DeviceError,
}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_i8p = ctypes.POINTER(ctypes.c_int8)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def _ptr(array: np.ndarray | None, ctype):
    if array is None:
        return ctypes.cast(None, ctypes.POINTER(ctype))
    return array.ctypes.data_as(ctypes.POINTER(ctype))


class Backend:
    """A C library exposing the include/rewl.h calls under `prefix`."""

    prefix = "wlx"

    def __init__(self, lib: ctypes.CDLL):
        self.lib = lib
        getattr(lib, f"{self.prefix}_last_error").restype = ctypes.c_size_t

    def last_error(self) -> str:
        fn = getattr(self.lib, f"{self.prefix}_last_error")
        size = fn(None, 0)
        buffer = ctypes.create_string_buffer(size + 1)
        fn(buffer, size + 1)
        return buffer.value.decode(errors="replace")

    def check(self, status: int) -> None:
        if status != abi.WLX_OK:
            raise _ERRORS.get(status, RuntimeError)(self.last_error())

    def call(self, name: str, *args) -> None:
        self.check(getattr(self.lib, f"{self.prefix}_{name}")(*args))

    # constructors differ between backends
    def system_create(self, positions, axes, n, box, periodic, scale, cutoff, precision=0):
        out = ctypes.c_void_p()
        self.call("system_create", _ptr(positions, ctypes.c_double), _ptr(axes, ctypes.c_double),
                  ctypes.c_uint64(n), _ptr(box, ctypes.c_double), ctypes.c_int(periodic),
                  ctypes.c_double(scale), ctypes.c_double(cutoff), ctypes.c_int(precision),
                  ctypes.byref(out))
        return out

    def system_destroy(self, handle):
        self.call("system_destroy", handle)

    def engine_create(self, system_handle, config: abi.Config):
        out = ctypes.c_void_p()
        self.call("engine_create", system_handle, ctypes.byref(config), ctypes.byref(out))
        return out

    def engine_destroy(self, handle):
        self.call("engine_destroy", handle)


class GpuBackend(Backend):
    prefix = "wlx"

    def __init__(self, lib: ctypes.CDLL | None = None):
        super().__init__(lib or abi.load_library())


_DEFAULT_BACKEND: Backend | None = None


def default_backend() -> Backend:
    global _DEFAULT_BACKEND
    if _DEFAULT_BACKEND is None:
        _DEFAULT_BACKEND = GpuBackend()
    return _DEFAULT_BACKEND


# --------------------------------------------------------------------------------------------
# configuration: defaults of WlParameters (wang_landau.hpp:40-56) and RunConfig (io.hpp:12-58)
# --------------------------------------------------------------------------------------------
@dataclass
class WlParameters:
    flatness: float = 0.8
    minimum_visits: int = 100
    final_factor: float = 1e-8
    check_interval_attempts: int = 10_000
    force_accept_after_attempts: int = 0
    inverse_time_enabled: bool = True
    initialization_max_attempts: int = 1_000_000
    initialization_target_fraction: float = 0.5
    initialization_temperature_fraction: float = 0.05
    initialization_stall_attempts_per_spin: int = 1_000
    initialization_temperature_multiplier: float = 2.0
    initialization_max_temperature_fraction: float = 0.5
    nalivaiko_mod: bool = False
    collect_window_statistics: bool = False
    round_trip_margin_fraction: float = 0.1

    def to_c(self) -> abi.Params:
        p = abi.Params()
        for f in dc_fields(self):
            setattr(p, f.name, getattr(self, f.name))
        return p


@dataclass
class RunConfig:
    emin: float = -20.0
    emax: float = 20.0
    bin_width: float = 0.1
    windows: int = 1
    walkers_per_window: int = 1
    overlap: float = 0.75
    seed: int = 1
    exchange_interval_attempts: int = 1000
    max_attempts: int = 10_000_000
    rng: int = abi.RNG_XOSHIRO
    fields: int = abi.FIELDS_FP64
    explicit_windows: Sequence[tuple[int, int]] | None = None
    resync_interval_attempts: int = 0
    wl: WlParameters = field(default_factory=WlParameters)

    def to_c(self) -> tuple[abi.Config, Any]:
        c = abi.Config()
        for name in ("emin", "emax", "bin_width", "windows", "walkers_per_window", "overlap",
                     "seed", "exchange_interval_attempts", "max_attempts", "rng", "fields",
                     "resync_interval_attempts"):
            setattr(c, name, getattr(self, name))
        keep = None
        if self.explicit_windows is not None:
            keep = np.ascontiguousarray(np.asarray(self.explicit_windows, dtype=np.uint64).ravel())
            c.explicit_windows = keep.ctypes.data_as(_u64p)
        c.wl = self.wl.to_c()
        return c, keep

    @property
    def bins(self) -> int:
        return int(round((self.emax - self.emin) / self.bin_width))


# --------------------------------------------------------------------------------------------
# geometry helpers (host): Geometry::simple_cubic (physics.cpp:58-75)
# --------------------------------------------------------------------------------------------
def simple_cubic(nx: int, ny: int, nz: int, spacing: float = 1.0,
                 axis: Sequence[float] = (0.0, 0.0, 1.0), periodic: bool = True):
    """Sites x + nx*(y + ny*z), box n*spacing, one normalised easy axis for every site."""
    if nx == 0 or ny == 0 or nz == 0 or not spacing > 0.0:
        raise InvalidArgument("Invalid simple-cubic dimensions or spacing")
    ax, ay, az = (float(v) for v in axis)
    length = (ax * ax + ay * ay + az * az) ** 0.5
    if not (length > 0.0) or length != length or length == float("inf"):
        raise InvalidArgument("Ising axes must be finite non-zero vectors")
    unit = (ax / length, ay / length, az / length)
    n = nx * ny * nz
    positions = np.empty((n, 3))
    k = 0
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                positions[k] = (float(x) * spacing, float(y) * spacing, float(z) * spacing)
                k += 1
    axes = np.tile(np.asarray(unit, dtype=np.float64), (n, 1))
    box = np.array([float(nx) * spacing, float(ny) * spacing, float(nz) * spacing])
    return positions, axes, box, bool(periodic)


# --------------------------------------------------------------------------------------------
class System:
    """Coupling system J_ij on the device (kernel a); dense if cutoff <= 0, else CSR.
    precision: abi.COUPLINGS_FP64 (the reference's), or abi.COUPLINGS_FP32 (every coupling
    rounded to binary32 after the FP64 build)."""

    def __init__(self, positions, axes, box, periodic: bool = True, coupling_scale: float = 1.0,
                 cutoff: float = 0.0, backend: Backend | None = None, precision: int = 0):
        self.backend = backend or default_backend()
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.axes = np.ascontiguousarray(axes, dtype=np.float64).reshape(-1, 3)
        self.box = np.ascontiguousarray(box, dtype=np.float64).reshape(3)
        self.periodic = bool(periodic)
        self.coupling_scale = float(coupling_scale)
        self.cutoff = float(cutoff)
        self.precision = int(precision)
        self.handle = self.backend.system_create(self.positions, self.axes, len(self.positions),
                                                 self.box, int(self.periodic),
                                                 self.coupling_scale, self.cutoff, self.precision)
        n, backend_id, nnz = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self.backend.call("system_info", self.handle, ctypes.byref(n), ctypes.byref(backend_id),
                          ctypes.byref(nnz))
        self.n = n.value
        self.backend_name = "csr" if backend_id.value else "dense"
        self.nnz = nnz.value

    @classmethod
    def simple_cubic(cls, nx, ny, nz, spacing=1.0, axis=(0.0, 0.0, 1.0), periodic=True,
                     coupling_scale=1.0, cutoff=0.0, backend=None, precision=0):
        pos, axes, box, per = simple_cubic(nx, ny, nz, spacing, axis, periodic)
        return cls(pos, axes, box, per, coupling_scale, cutoff, backend, precision)

    def dense(self) -> np.ndarray:
        out = np.empty((self.n, self.n))
        self.backend.call("system_read_dense", self.handle, _ptr(out, ctypes.c_double))
        return out

    def csr(self):
        offsets = np.empty(self.n + 1, dtype=np.uint64)
        columns = np.empty(max(1, self.nnz), dtype=np.uint64)
        values = np.empty(max(1, self.nnz))
        self.backend.call("system_read_csr", self.handle, _ptr(offsets, ctypes.c_uint64),
                          _ptr(columns, ctypes.c_uint64), _ptr(values, ctypes.c_double))
        return offsets, columns[: self.nnz], values[: self.nnz]

    def pilot(self, seed: int, attempts: int) -> tuple[float, float]:
        """The wl_run --pilot sweep (wl_run.cpp:84-104): sampled energy range."""
        lo, hi = ctypes.c_double(), ctypes.c_double()
        self.backend.call("system_pilot", self.handle, ctypes.c_uint64(seed),
                          ctypes.c_uint64(attempts), ctypes.byref(lo), ctypes.byref(hi))
        return lo.value, hi.value

    @staticmethod
    def pilot_many(systems, seed: int, attempts) -> tuple[np.ndarray, np.ndarray]:
        """The pilot sweep of several systems in one launch (wlx_systems_pilot); attempts per
        system. Returns (minima, maxima)."""
        count = len(systems)
        handles = (ctypes.c_void_p * count)(*[s.handle.value for s in systems])
        att = np.ascontiguousarray(attempts, dtype=np.uint64)
        lo, hi = np.empty(count), np.empty(count)
        systems[0].backend.call("systems_pilot", handles, ctypes.c_uint64(count),
                                ctypes.c_uint64(seed), _ptr(att, ctypes.c_uint64),
                                _ptr(lo, ctypes.c_double), _ptr(hi, ctypes.c_double))
        return lo, hi

    def fields(self, spins) -> tuple[np.ndarray, np.ndarray]:
        """local_fields + total_energy (physics.cpp:238-259) per configuration."""
        s = np.ascontiguousarray(spins, dtype=np.int8)
        single = s.ndim == 1
        s = s.reshape(-1, self.n)
        fields = np.empty((len(s), self.n))
        energies = np.empty(len(s))
        self.backend.call("system_fields", self.handle, _ptr(s, ctypes.c_int8),
                          ctypes.c_uint64(len(s)), _ptr(fields, ctypes.c_double),
                          _ptr(energies, ctypes.c_double))
        return (fields[0], energies[0]) if single else (fields, energies)

    def close(self):
        if getattr(self, "handle", None):
            self.backend.system_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class WalkerSnapshot:
    """WalkerSnapshot (wang_landau.hpp:74-89) plus driver bookkeeping; arrays are full-grid."""

    info: abi.WalkerInfo
    spins: np.ndarray
    fields: np.ndarray
    log_g: np.ndarray
    histogram: np.ndarray
    active: np.ndarray
    refinement_active: np.ndarray

    def __getattr__(self, name):
        return getattr(self.__dict__["info"], name)


class Engine:
    """Device-resident REWL walkers (kernel b) for one coupling system."""

    def __init__(self, system: System, config: RunConfig):
        self.system = system
        self.backend = system.backend
        self.config = config
        c, self._keep = config.to_c()
        self.handle = self.backend.engine_create(system.handle, c)
        walkers, spins, bins = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self.backend.call("engine_walkers", self.handle, ctypes.byref(walkers),
                          ctypes.byref(spins), ctypes.byref(bins))
        self.walkers, self.n, self.bins = walkers.value, spins.value, bins.value

    def layout(self) -> tuple[int, bool]:
        """(LDS bytes of one LDS-walker workgroup, window bins kept in HBM)."""
        lds, hbm = ctypes.c_uint64(), ctypes.c_uint64()
        self.backend.call("engine_layout", self.handle, ctypes.byref(lds), ctypes.byref(hbm))
        return lds.value, bool(hbm.value)

    def initialize(self, initial_spins=None, supplied=None):
        if initial_spins is None:
            self.backend.call("engine_initialize", self.handle, _ptr(None, ctypes.c_int8),
                              _ptr(None, ctypes.c_uint8))
            return
        spins = np.ascontiguousarray(initial_spins, dtype=np.int8).reshape(self.walkers, self.n)
        mask = (np.ones(self.walkers, dtype=np.uint8) if supplied is None
                else np.ascontiguousarray(supplied, dtype=np.uint8))
        self.backend.call("engine_initialize", self.handle, _ptr(spins, ctypes.c_int8),
                          _ptr(mask, ctypes.c_uint8))

    def sample(self, attempts: int):
        self.backend.call("engine_sample", self.handle, ctypes.c_uint64(attempts))

    def check(self):
        self.backend.call("engine_check", self.handle)

    def exchange(self, epoch: int) -> tuple[int, int]:
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self.backend.call("engine_exchange", self.handle, ctypes.c_uint64(epoch), ctypes.byref(a),
                          ctypes.byref(b))
        return a.value, b.value

    def run(self, max_epochs: int = 2**63, status: bool = True) -> abi.RunStatus | None:
        """Runs up to `max_epochs` exchange epochs; returns after the device has finished them.
        With status=False the run's status is not read back (None is returned)."""
        st = abi.RunStatus() if status else None
        self.backend.call("engine_run", self.handle, ctypes.c_uint64(max_epochs),
                          ctypes.byref(st) if status else None)
        return st

    def status(self) -> abi.RunStatus:
        st = abi.RunStatus()
        self.backend.call("engine_status", self.handle, ctypes.byref(st))
        return st

    def resync(self):
        """FP64 recompute of every configuration's fields and energy (FP32 fields mode)."""
        self.backend.call("engine_resync", self.handle)

    def warm_start(self, log_g, factor: float):
        """Presets every walker's window log g and ln f = factor (wlx_engine_warm_start)."""
        lg = np.ascontiguousarray(log_g, dtype=np.float64)
        if lg.shape != (self.bins,):
            raise InvalidArgument("The warm-start log g must span the grid")
        self.backend.call("engine_warm_start", self.handle, _ptr(lg, ctypes.c_double),
                          ctypes.c_double(factor))

    def set_kernel(self, kernel: int):
        """Forces a walker kernel family (abi.KERNEL_*); all give the same trajectory."""
        self.backend.call("engine_set_kernel", self.handle, ctypes.c_uint64(kernel))

    def set_lookahead_width(self, width: int):
        """Attempts per lookahead batch of the Philox common-configuration kernel."""
        self.backend.call("engine_set_lookahead_width", self.handle, ctypes.c_uint64(width))

    def kernel(self) -> tuple[int, str]:
        """(family, kernel name) of the next sampling launch."""
        family = ctypes.c_uint64()
        name = ctypes.create_string_buffer(128)
        self.backend.call("engine_kernel", self.handle, ctypes.byref(family), name,
                          ctypes.c_size_t(128))
        return family.value, name.value.decode()

    def autotune(self, epochs: int) -> tuple[int, list[float]]:
        """Runs `epochs` epochs of the live run with each eligible kernel family and keeps the
        fastest: (chosen family, seconds per family, -1 where not tried)."""
        chosen = ctypes.c_uint64()
        seconds = (ctypes.c_double * 5)()
        self.backend.call("engine_autotune", self.handle, ctypes.c_uint64(epochs),
                          ctypes.byref(chosen), seconds)
        return chosen.value, list(seconds)

    def drift(self) -> tuple[float, float, int]:
        """(max field error, max energy error in bins, resyncs) over all resyncs so far."""
        f, e, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        self.backend.call("engine_drift", self.handle, ctypes.byref(f), ctypes.byref(e),
                          ctypes.byref(n))
        return f.value, e.value, n.value

    def set_walker_cadence(self, k: int, last_flatness_check: int, last_flatness: float):
        """The driver's flatness-check bookkeeping of walker k (parallel.cpp:364-365)."""
        self.backend.call("engine_set_walker_cadence", self.handle, ctypes.c_uint64(k),
                          ctypes.c_uint64(last_flatness_check), ctypes.c_double(last_flatness))

    def set_run_state(self, epoch: int, exchange_attempted: int, exchange_accepted: int):
        """Run counters of a resumed run (exchange epoch and tallies)."""
        self.backend.call("engine_set_run_state", self.handle, ctypes.c_uint64(epoch),
                          ctypes.c_uint64(exchange_attempted), ctypes.c_uint64(exchange_accepted))

    # ---- multi-walker checkpoint: every walker's snapshot plus the run counters -----------
    _SNAP_FIELDS = ("walker_id", "window", "window_begin", "window_end", "energy", "factor",
                    "attempted", "accepted", "forced_accepted", "last_accepted_attempt",
                    "stage", "active_bins", "refinement_active_bins", "last_flatness_check",
                    "last_flatness", "clock_offset")

    def save_checkpoint(self, path) -> None:
        """Writes every walker's WalkerSnapshot (wang_landau.hpp:74-89) and the run counters
        to an .npz file. The reference checkpoints only 1x1 runs (parallel.cpp:354,522)."""
        snaps = [self.read_walker(k) for k in range(self.walkers)]
        st = self.status()
        data = {name: np.array([getattr(s.info, name) for s in snaps])
                for name in self._SNAP_FIELDS}
        data["rng_state"] = np.array([list(s.info.rng_state) for s in snaps], dtype=np.uint64)
        for name in ("spins", "fields", "log_g", "histogram", "active"):
            data[name] = np.stack([getattr(s, name) for s in snaps])
        data["run"] = np.array([st.epoch, st.exchange_attempted, st.exchange_accepted],
                               dtype=np.uint64)
        data["shape"] = np.array([self.walkers, self.n, self.bins], dtype=np.uint64)
        np.savez_compressed(path, **data)

    def load_checkpoint(self, path, recompute_fields: bool = False) -> None:
        """Restores a save_checkpoint file into this (initialized) engine, walker by walker
        through WangLandauWalker::restore's checks (wang_landau.cpp:420-466).

        Those checks require the saved energy and fields to match a fresh FP64 recompute
        within 1e-9 relative. After billions of incremental updates the running values drift
        further than that, in the reference as well. recompute_fields=True restarts every
        configuration from the recompute instead of the saved values; the resumed run then
        differs from an uninterrupted one in the last bits."""
        with np.load(path, allow_pickle=False) as z:
            data = {k: z[k] for k in z.files}
        if "clock_offset" not in data:  # files written before warm starts were checkpointed
            data["clock_offset"] = np.zeros(self.walkers, dtype=np.uint64)
        if tuple(int(x) for x in data["shape"]) != (self.walkers, self.n, self.bins):
            raise InvalidArgument("Checkpoint does not match the engine's walkers, spins or bins")
        if recompute_fields:
            fields, energies = self.system.fields(data["spins"][:, : self.n])
            data["fields"] = fields
            data["energy"] = energies
        for k in range(self.walkers):
            snap = self.read_walker(k)
            for name in self._SNAP_FIELDS:
                setattr(snap.info, name, type(getattr(snap.info, name))(data[name][k]))
            for j in range(4):
                snap.info.rng_state[j] = int(data["rng_state"][k][j])
            snap.spins = data["spins"][k]
            snap.fields = data["fields"][k]
            snap.log_g = data["log_g"][k]
            snap.histogram = data["histogram"][k]
            snap.active = data["active"][k]
            self.write_walker(k, snap)
            self.set_walker_cadence(k, int(data["last_flatness_check"][k]),
                                    float(data["last_flatness"][k]))
        epoch, ex_att, ex_acc = (int(x) for x in data["run"])
        self.set_run_state(epoch, ex_att, ex_acc)

    def read_walker(self, k: int) -> WalkerSnapshot:
        info = abi.WalkerInfo()
        spins = np.empty(self.n, dtype=np.int8)
        fields = np.empty(self.n)
        log_g = np.empty(self.bins)
        hist = np.empty(self.bins, dtype=np.uint64)
        active = np.empty(self.bins, dtype=np.uint8)
        refine = np.empty(self.bins, dtype=np.uint8)
        self.backend.call("engine_read_walker", self.handle, ctypes.c_uint64(k),
                          ctypes.byref(info), _ptr(spins, ctypes.c_int8),
                          _ptr(fields, ctypes.c_double), _ptr(log_g, ctypes.c_double),
                          _ptr(hist, ctypes.c_uint64), _ptr(active, ctypes.c_uint8),
                          _ptr(refine, ctypes.c_uint8))
        return WalkerSnapshot(info, spins, fields, log_g, hist, active, refine)

    def write_walker(self, k: int, snap: WalkerSnapshot):
        """WangLandauWalker::restore (wang_landau.cpp:420-466)."""
        arrays = [np.ascontiguousarray(snap.spins, dtype=np.int8),
                  np.ascontiguousarray(snap.fields, dtype=np.float64),
                  np.ascontiguousarray(snap.log_g, dtype=np.float64),
                  np.ascontiguousarray(snap.histogram, dtype=np.uint64),
                  np.ascontiguousarray(snap.active, dtype=np.uint8)]
        self.backend.call("engine_write_walker", self.handle, ctypes.c_uint64(k),
                          ctypes.byref(snap.info), _ptr(arrays[0], ctypes.c_int8),
                          _ptr(arrays[1], ctypes.c_double), _ptr(arrays[2], ctypes.c_double),
                          _ptr(arrays[3], ctypes.c_uint64), _ptr(arrays[4], ctypes.c_uint8))

    def read_sampling(self, k: int) -> dict:
        """Pilot statistics of walker k (wlx_engine_read_sampling): squared displacements and
        samples per bin (full grid), round trips, and the nine representatives as
        (energy, spins) pairs, None where none was recorded."""
        sq = np.empty(self.bins)
        samples = np.empty(self.bins, dtype=np.uint64)
        trips = ctypes.c_uint64()
        energies = np.empty(9)
        present = np.empty(9, dtype=np.uint8)
        spins = np.empty((9, self.n), dtype=np.int8)
        self.backend.call("engine_read_sampling", self.handle, ctypes.c_uint64(k),
                          _ptr(sq, ctypes.c_double), _ptr(samples, ctypes.c_uint64),
                          ctypes.byref(trips), _ptr(energies, ctypes.c_double),
                          _ptr(present, ctypes.c_uint8), _ptr(spins, ctypes.c_int8))
        reps = [(float(energies[r]), spins[r].copy()) if present[r] else None for r in range(9)]
        return {"squared_displacement": sq, "displacement_samples": samples,
                "round_trips": trips.value, "representatives": reps}

    def statistics(self, k: int) -> abi.HistogramStats:
        out = abi.HistogramStats()
        self.backend.call("engine_histogram_statistics", self.handle, ctypes.c_uint64(k),
                          ctypes.byref(out))
        return out

    def ready(self, k: int) -> bool:
        r = ctypes.c_uint64()
        self.backend.call("engine_ready", self.handle, ctypes.c_uint64(k), ctypes.byref(r))
        return bool(r.value)

    def next_iteration(self, k: int):
        self.backend.call("engine_next_iteration", self.handle, ctypes.c_uint64(k))

    def trace(self, capacity: int):
        self.backend.call("engine_trace", self.handle, ctypes.c_uint64(capacity))
        self._trace_capacity = capacity

    def read_trace(self, k: int):
        count = ctypes.c_uint64()
        cap = getattr(self, "_trace_capacity", 0)
        sites = np.empty(max(1, cap), dtype=np.uint32)
        bins = np.empty(max(1, cap), dtype=np.uint32)
        flags = np.empty(max(1, cap), dtype=np.uint8)
        self.backend.call("engine_read_trace", self.handle, ctypes.c_uint64(k),
                          ctypes.byref(count), _ptr(sites, ctypes.c_uint32),
                          _ptr(bins, ctypes.c_uint32), _ptr(flags, ctypes.c_uint8))
        c = count.value
        return sites[:c], bins[:c], flags[:c]

    def set_timing(self, enable: bool = True, every: int = 1):
        """HIP events around every `every`-th sampling launch of run() (wlx_engine_set_timing);
        kernel_time() then averages the timed launches."""
        if every < 1:
            raise InvalidArgument("Timing period must be >= 1")
        self.backend.call("engine_set_timing", self.handle,
                          ctypes.c_int(int(every) if enable else 0))

    def set_graph(self, epochs_per_graph: int):
        """run() replays `epochs_per_graph` captured epochs per graph launch
        (wlx_engine_set_graph; 0: plain launches)."""
        self.backend.call("engine_set_graph", self.handle, ctypes.c_uint64(epochs_per_graph))

    def kernel_time(self) -> tuple[float, int]:
        ms, launches = ctypes.c_double(), ctypes.c_uint64()
        self.backend.call("engine_kernel_time", self.handle, ctypes.byref(ms),
                          ctypes.byref(launches))
        return ms.value, launches.value

    def close(self):
        if getattr(self, "handle", None):
            self.backend.engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several engines whose epochs run together (wlx_batch): one sampling launch over all
    their walkers and one exchange launch per epoch, optionally replayed from a HIP graph.
    The engines stay usable for reads; each one's results equal running it alone."""

    def __init__(self, engines):
        if not engines:
            raise InvalidArgument("A batch needs at least one engine")
        self.engines = list(engines)
        self.backend = self.engines[0].backend
        handles = (ctypes.c_void_p * len(self.engines))(*[e.handle.value for e in self.engines])
        out = ctypes.c_void_p()
        self.backend.call("batch_create", handles, ctypes.c_uint64(len(self.engines)),
                          ctypes.byref(out))
        self.handle = out

    def set_graph(self, epochs_per_graph: int):
        """Replays `epochs_per_graph` captured epochs per launch (0: plain launches)."""
        self.backend.call("batch_set_graph", self.handle, ctypes.c_uint64(epochs_per_graph))

    def run(self, max_epochs: int = 2**63) -> tuple[int, int]:
        """(epochs enqueued, engines done)."""
        epochs, done = ctypes.c_uint64(), ctypes.c_uint64()
        self.backend.call("batch_run", self.handle, ctypes.c_uint64(max_epochs),
                          ctypes.byref(epochs), ctypes.byref(done))
        return epochs.value, done.value

    def set_lds_budget(self, nbytes: int):
        """Engines needing more LDS than `nbytes` keep their window bins in HBM."""
        self.backend.call("batch_set_lds_budget", self.handle, ctypes.c_uint64(nbytes))

    def layout(self) -> dict:
        """Launch groups, engines with bins in HBM, largest per-walker LDS."""
        g, h, l = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self.backend.call("batch_layout", self.handle, ctypes.byref(g), ctypes.byref(h),
                          ctypes.byref(l))
        return {"groups": g.value, "bins_in_hbm": h.value, "max_lds_bytes": l.value}

    def autotune(self, epochs: int) -> tuple[int, list[float]]:
        """(chosen LDS budget, seconds per budget of 160, 64 and 32 KiB)."""
        chosen = ctypes.c_uint64()
        seconds = (ctypes.c_double * 3)()
        self.backend.call("batch_autotune", self.handle, ctypes.c_uint64(epochs),
                          ctypes.byref(chosen), seconds)
        return chosen.value, list(seconds)

    def kernel(self) -> str:
        """Name of the sampling kernel the batch launches."""
        name = ctypes.create_string_buffer(64)
        self.backend.call("batch_kernel", self.handle, name, ctypes.c_size_t(len(name)))
        return name.value.decode()

    def close(self):
        if getattr(self, "handle", None):
            self.backend.call("batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
