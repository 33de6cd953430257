"""jaxdd_amd: jaxdd's batched steady-state drift-diffusion J-V solve, MI355X-native.

The reference's API for this path (src/jaxdd/__init__.py:14-66): SIMsalabim input files in
(`read_setup`, `setup_from_texts`), a device built from them (`build_device`), its J-V curve
(`solve_jv`) or a batch's (`prepare`, `solve_batch`), and the batched block sweep operator
(`block_sweep`). The solve runs in hand-written HIP kernels for gfx950 (``libddhip.so``),
called through a C-ABI; torch tensors hold the buffers.
"""

from .batch import external_voltages, prepare, solve_batch
from .device import Device, Structure, build_device, device_arrays, device_parameters, make_grid
from .simss_input import (
    Layer,
    Setup,
    check_supported,
    read_setup,
    setup_from_texts,
    simss_voltages,
)
from .solver import JVResult, bernoulli, bernoulli_simss, solve_jv
from .sweep import block_sweep

__version__ = "0.1.0"

__all__ = [
    "Device",
    "JVResult",
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
    "setup_from_texts",
    "simss_voltages",
    "solve_batch",
    "solve_jv",
]
