"""One device's J-V curve (`solve_jv`), its result type, and the Bernoulli functions.

Reference interface mirrored: src/jaxdd/solver.py:95-121 (`bernoulli`, `bernoulli_simss`),
:171-177 (`JVResult`), :2041-2060 (`solve_jv`). `solve_jv` runs the same HIP kernels as
`solve_batch`, for a batch of one, in the reference's one-device mode: the voltages in the
given order, each solved from the last converged state (no extrapolation), Newton stopped by
the step tolerance or the round-off floor only, and a failed pre-bias raising RuntimeError.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .batch import _run, prepare
from .device import Device

__all__ = ["JVResult", "bernoulli", "bernoulli_simss", "solve_jv"]

_SERIES_BELOW = 1e-2


def bernoulli(x):
    """``B(x) = x/(e^x - 1)``, by its series below |x| = 1e-2 (src/jaxdd/solver.py:95-105)."""
    x = np.asarray(x, dtype=float)
    small = np.abs(x) < _SERIES_BELOW
    safe = np.where(small, 1.0, x)
    series = 1.0 - x / 2.0 + x * x / 12.0 - x ** 4 / 720.0
    return np.where(small, series, safe / np.expm1(safe))


def bernoulli_simss(x):
    """SIMsalabim 5.36's piecewise `B` (src/jaxdd/solver.py:108-121): Taylor orders 1, 2, 4
    below |x| 0.05, 0.45, 1.5, the full expression above."""
    x = np.asarray(x, dtype=float)
    ax = np.abs(x)
    safe = np.where(ax < 1.5, 2.0, x)
    full = safe / np.expm1(safe)
    o1 = 1.0 - 0.5 * x
    o2 = 1.0 - x * (0.5 - x / 12.0)
    o4 = 1.0 + x * (-0.5 + x * (1.0 / 12.0 - x * x / 720.0))
    return np.where(ax < 0.05, o1, np.where(ax < 0.45, o2, np.where(ax < 1.5, o4, full)))


BERNOULLI = {"exact": bernoulli, "simss": bernoulli_simss}


@dataclass(frozen=True)
class JVResult:
    V: np.ndarray  # the internal voltage
    J: np.ndarray  # A/m^2, SIMsalabim's sign (photocurrent negative), with the shunt's current
    converged: np.ndarray
    iterations: np.ndarray
    V_ext: np.ndarray | None = None  # V + J * R_series


def solve_jv(dev: Device, voltages, g_frac: float | None = None, tol: float = 1e-10,
             max_iter: int = 200, bernoulli_name: str = "exact",
             precision: str = "fp64") -> JVResult:
    """The J-V curve at ``voltages`` (internal), each solved from the last converged one
    (reference: src/jaxdd/solver.py:2041-2060). With ``preCond`` the device is first solved at
    ``Vpre`` with its ions free, and with ``fixIons`` they are then held for the scan; a pre-bias
    that does not converge raises RuntimeError."""
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    V = np.asarray(voltages, dtype=float).ravel()
    batch = prepare([dev])
    if g_frac is not None:
        batch["g_frac"] = batch["g_frac"] * 0 + float(g_frac)
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), 1,
                     0 if bernoulli_name == "exact" else 1, _lib.PRECISIONS[precision], 1)
    J, ok, its, status = _run(batch, V, opts, _lib.PRECISIONS[precision])
    if int(status[0].item()) != 0:
        raise RuntimeError(f"no convergence at the pre-bias {dev.V_pre} V")
    J = J[0].cpu().numpy()
    if dev.R_shunt > 0:
        J = J + V / dev.R_shunt
    return JVResult(V, J, ok[0].cpu().numpy().astype(bool), its[0].cpu().numpy().astype(np.int64),
                    V + J * dev.R_series)
