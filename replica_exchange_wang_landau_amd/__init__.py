"""MI355X-native replica-exchange Wang-Landau engine for Ising dipoles.

The walker-step hot path of Faikert/replica_exchange_wang_landau (coupling build, WL walker,
even/odd replica exchange) runs as hand-written HIP kernels for gfx950 behind the C ABI of
include/rewl.h; this package is its Python face (ctypes) and the host-side helpers.
"""
from .abi import load_library  # noqa: F401
from .engine import (  # noqa: F401
    Backend,
    Batch,
    DeviceError,
    Engine,
    GpuBackend,
    InvalidArgument,
    LogicError,
    RunConfig,
    System,
    WalkerSnapshot,
    WlParameters,
    simple_cubic,
)
