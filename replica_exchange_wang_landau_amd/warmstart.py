"""Learned warm start: a BF16 MLP (wlx_mlp_*) that maps (E/N, N, density, dilution, slab
thickness, axis mode) to log g / N, trained on converged densities of states this engine
produces for a family of dipolar slabs, and used to preset every walker's log g and its
initial ln f. Plain flat-start WL stays the fallback; a warm start changes only where the
walkers begin, and the run then refines log g with the same WL / 1-over-t schedule.

No reference counterpart: the reference always starts WL from log g = 0 and ln f = 1
(wang_landau.cpp:91-118).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import abi
from .engine import (Batch, Engine, GpuBackend, InvalidArgument, RunConfig, System, WlParameters,
                     _ptr, default_backend)

FEATURES = 6


class Mlp:
    """The warm-start network on the device (include/rewl.h, wlx_mlp_*)."""

    def __init__(self, seed: int = 1234, backend: GpuBackend | None = None):
        self.backend = backend or default_backend()
        out = ctypes.c_void_p()
        self.backend.call("mlp_create", ctypes.c_uint64(seed), ctypes.byref(out))
        self.handle = out
        count = ctypes.c_uint64()
        self.backend.call("mlp_params", ctypes.byref(count))
        self.param_count = count.value

    def params(self) -> np.ndarray:
        p = np.empty(self.param_count, dtype=np.float32)
        self.backend.call("mlp_get_params", self.handle, _ptr(p, ctypes.c_float))
        return p

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (self.param_count,):
            raise InvalidArgument("Wrong parameter count")
        self.backend.call("mlp_set_params", self.handle, _ptr(p, ctypes.c_float))

    def predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        y = np.empty(len(x), dtype=np.float32)
        self.backend.call("mlp_predict", self.handle, _ptr(x, ctypes.c_float),
                          ctypes.c_uint64(len(x)), _ptr(y, ctypes.c_float))
        return y

    def gradient(self, x, y) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        g = np.empty(self.param_count, dtype=np.float32)
        self.backend.call("mlp_gradient", self.handle, _ptr(x, ctypes.c_float),
                          _ptr(y, ctypes.c_float), ctypes.c_uint64(len(x)), ctypes.byref(loss),
                          _ptr(g, ctypes.c_float))
        return loss.value, g

    def train(self, x, y, steps: int, batch: int = 1024, learning_rate: float = 1e-3,
              seed: int = 1) -> np.ndarray:
        """Adam steps on shuffled minibatches; returns the loss before every step."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        y = np.ascontiguousarray(y, dtype=np.float32)
        losses = np.empty(steps, dtype=np.float32)
        self.backend.call("mlp_train", self.handle, _ptr(x, ctypes.c_float),
                          _ptr(y, ctypes.c_float), ctypes.c_uint64(len(x)),
                          ctypes.c_uint64(steps), ctypes.c_uint64(min(batch, len(x))),
                          ctypes.c_float(learning_rate), ctypes.c_uint64(seed),
                          _ptr(losses, ctypes.c_float))
        return losses

    def close(self):
        if getattr(self, "handle", None):
            self.backend.call("mlp_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the slab family ------------------------------------------------------------------------
@dataclass(frozen=True)
class Slab:
    """Lx x Ly x Lz simple-cubic sites (spacing 1), periodic in x and y (box Lx, Ly) and open
    in z (box length 0), each site removed with probability `dilution`; easy axes all +z
    (axis_mode 0) or random in the plane (axis_mode 1); dipolar D = 1, CSR cutoff 3."""
    lx: int
    lz: int
    dilution: float
    axis_mode: int
    seed: int

    def geometry(self):
        rng = np.random.default_rng(self.seed)
        grid = np.stack(np.meshgrid(np.arange(self.lx, dtype=float),
                                    np.arange(self.lx, dtype=float),
                                    np.arange(self.lz, dtype=float), indexing="ij"),
                        -1).reshape(-1, 3)
        keep = rng.random(len(grid)) >= self.dilution
        if keep.sum() < 8:
            keep[:8] = True
        pos = grid[keep]
        if self.axis_mode == 0:
            axes = np.tile([0.0, 0.0, 1.0], (len(pos), 1))
        else:
            phi = rng.uniform(0.0, 2.0 * math.pi, size=len(pos))
            axes = np.stack([np.cos(phi), np.sin(phi), np.zeros_like(phi)], -1)
        return pos, axes, np.array([float(self.lx), float(self.lx), 0.0]), True

    def system(self, backend=None) -> System:
        return System(*self.geometry(), backend=backend, cutoff=3.0)

    def features(self, n: int) -> tuple[float, float, float, float, float]:
        """(N, density, dilution, thickness, axis mode) of the network input."""
        density = n / float(self.lx * self.lx * self.lz)
        return float(n), density, self.dilution, float(self.lz), float(self.axis_mode)


def network_input(e_per_spin, n, density, dilution, thickness, axis_mode) -> np.ndarray:
    """The scaled feature rows of the network for energies per spin `e_per_spin`."""
    e = np.asarray(e_per_spin, dtype=np.float64)
    cols = [e / 2.0, np.full_like(e, n / 500.0 - 1.0), np.full_like(e, density),
            np.full_like(e, dilution * 2.5), np.full_like(e, thickness / 4.0 - 1.0),
            np.full_like(e, axis_mode)]
    return np.stack(cols, -1).astype(np.float32)


def energy_spread(system: System) -> float:
    """s = sigma^2 / N with sigma^2 = sum over pairs i < j of J_ij^2: the variance of E over
    uniformly random spins (the infinite-temperature ensemble, where the mean energy is 0 because
    J_ii = 0). Near the band centre log g(E) / N ~ ln 2 - (E/N)^2 / (2 s)."""
    if system.backend_name == "csr":
        _, _, values = system.csr()
        total = float(np.sum(values.astype(np.float64) ** 2))
    else:
        total = float(np.sum(system.dense().astype(np.float64) ** 2))
    return total / 2.0 / system.n  # each pair appears twice in the symmetric matrix


def gaussian_log_g(e_per_spin, spread: float) -> np.ndarray:
    """The Gaussian part of log g / N with its peak at 0: -(E/N)^2 / (2 s)."""
    e = np.asarray(e_per_spin, dtype=np.float64)
    return -(e * e) / (2.0 * spread)


def slab_family(count: int, seed: int = 1234) -> list[Slab]:
    """Lx = Ly in {6..12}, Lz in {2..6}, dilution uniform in [0, 0.4], axis mode 0 or 1."""
    rng = np.random.default_rng(seed)
    return [Slab(int(rng.integers(6, 13)), int(rng.integers(2, 7)), float(rng.uniform(0.0, 0.4)),
                 int(rng.integers(0, 2)), int(seed * 1000 + k)) for k in range(count)]


def slab_config(system: System, lo: float, hi: float, seed: int, final_factor: float,
                walkers_per_window: int = 8) -> RunConfig:
    """Grid from a pilot range (padded 5 %, bin 0.5), windows in proportion to N (4-16),
    exchange every MCS, Philox stream."""
    n = system.n
    pad = 0.05 * (hi - lo) + 0.5
    width = 0.5
    emin = width * math.floor((lo - pad) / width)
    emax = width * math.ceil((hi + pad) / width)
    windows = min(16, max(4, round(4 + 12 * (n - 43) / 821)))
    return RunConfig(emin=emin, emax=emax, bin_width=width, windows=windows,
                     walkers_per_window=walkers_per_window, overlap=0.75, seed=seed,
                     exchange_interval_attempts=n, max_attempts=0, rng=abi.RNG_PHILOX,
                     wl=WlParameters(final_factor=final_factor, check_interval_attempts=20 * n))


def converged_curves(slabs, final_factor: float = 1e-3, pilot_mcs: int = 2000,
                     graph: int = 16, report=None):
    """Runs every slab's REWL to ln f = final_factor in one batch and returns, per slab,
    (bin energies, stitched log g with the peak at 0, valid mask, system)."""
    from . import analysis

    systems = [s.system() for s in slabs]
    lo, hi = System.pilot_many(systems, 1234, [pilot_mcs * s.n for s in systems])
    engines = []
    for k, (system, a, b) in enumerate(zip(systems, lo, hi)):
        e = Engine(system, slab_config(system, a, b, 1234 + k, final_factor))
        e.initialize()
        e.set_kernel(abi.KERNEL_LDS)
        engines.append(e)
    batch = Batch(engines)
    batch.set_graph(graph)
    done = 0
    while done < len(engines):
        _, done = batch.run(2048)
        if report:
            report(done, len(engines))
    out = []
    for slab, system, e in zip(slabs, systems, engines):
        log_g, valid = analysis.engine_dos(e)
        centers = e.config.emin + (np.arange(e.bins) + 0.5) * e.config.bin_width
        out.append((centers, log_g, valid, system))
    batch.close()
    return out


def training_set(slabs, curves, gaussian: bool = False):
    """Network rows and targets (log g - max log g) / N over every valid bin of every curve;
    with `gaussian`, the target is what remains after each realization's own Gaussian part
    (gaussian_log_g with its energy_spread), so the network learns the non-Gaussian shape."""
    xs, ys = [], []
    for slab, (centers, log_g, valid, system) in zip(slabs, curves):
        ok = valid != 0
        n, density, dilution, thickness, axis_mode = slab.features(system.n)
        e = centers[ok] / n
        xs.append(network_input(e, n, density, dilution, thickness, axis_mode))
        y = log_g[ok] / n
        if gaussian:
            y = y - gaussian_log_g(e, energy_spread(system))
        ys.append(y.astype(np.float32))
    return np.concatenate(xs), np.concatenate(ys)


def predicted_log_g(mlp: Mlp, engine: Engine, features, spread: float | None = None) -> np.ndarray:
    """N * network(E/N) at every bin centre of the engine's grid, plus N times the Gaussian
    part with energy spread `spread` for a network trained on the residual."""
    n, density, dilution, thickness, axis_mode = features
    centers = engine.config.emin + (np.arange(engine.bins) + 0.5) * engine.config.bin_width
    x = network_input(centers / n, n, density, dilution, thickness, axis_mode)
    out = mlp.predict(x).astype(np.float64)
    if spread is not None:
        out = out + gaussian_log_g(centers / n, spread)
    return n * out


def warm_start(engine: Engine, log_g, factor: float) -> None:
    """Presets every walker of an initialized engine from `log_g` (full grid, any offset) with
    ln f = factor (Engine.warm_start / wlx_engine_warm_start)."""
    engine.warm_start(log_g, factor)
