"""jaxdd_amd: jaxdd's batched steady-state drift-diffusion J-V solve, MI355X-native.

The reference's API for this path (src/jaxdd/__init__.py:14-66): SIMsalabim input files in
(`read_setup`, `setup_from_texts`), a device built from them (`build_device`), its J-V curve
(`solve_jv`) or a batch's (`prepare`, `solve_batch`), the batched block sweep operator
(`block_sweep`) and a batch's impedance at a steady state (`solve_impedance`). The solve runs in hand-written HIP kernels for gfx950 (``libddhip.so``),
called through a C-ABI; torch tensors hold the buffers.
"""

from .device import Device, Structure, build_device, device_arrays, device_parameters, make_grid
from .simss_input import (
    Layer,
    Setup,
    check_supported,
    read_setup,
    read_tvg,
    setup_from_texts,
    simss_voltages,
)

# The GPU entry points import torch; they are loaded on first use, so that the host-only
# modules (the reader, the device builder, the synthetic workloads) import without it.
_LAZY = {
    "external_voltages": "batch", "prepare": "batch", "solve_batch": "batch",
    "solve_states": "batch",
    "JVResult": "solver", "bernoulli": "solver", "bernoulli_simss": "solver", "solve_jv": "solver",
    "block_sweep": "sweep",
    "ImpedanceResult": "impedance", "solve_impedance": "impedance",
    "TransientResult": "transient", "solve_transient": "transient",
}


def __getattr__(name):
    if name in _LAZY:
        import importlib

        return getattr(importlib.import_module(f".{_LAZY[name]}", __name__), name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")

__version__ = "0.1.0"

__all__ = [
    "Device",
    "ImpedanceResult",
    "JVResult",
    "TransientResult",
    "Layer",
    "Setup",
    "Structure",
    "__version__",
    "bernoulli",
    "bernoulli_simss",
    "block_sweep",
    "build_device",
    "check_supported",
    "device_arrays",
    "device_parameters",
    "external_voltages",
    "make_grid",
    "prepare",
    "read_setup",
    "read_tvg",
    "setup_from_texts",
    "simss_voltages",
    "solve_batch",
    "solve_impedance",
    "solve_jv",
    "solve_states",
    "solve_transient",
]
