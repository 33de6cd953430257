"""Impedance in the frequency domain from the steady state's Jacobian (reference:
src/jaxdd/impedance.py:47-61 `ImpedanceResult`, :371-409 `solve_impedance`).

At the steady state at the internal voltage `voltage` (the solve kernel, keeping its final
state) a small voltage `dV e^(iwt)` drives `(J + iw M) dy = -(dF/dV) dV`: `J` the steady
residual's Jacobian, `M` implicit Euler's time terms, the bulk traps' occupations lagging at
their own rates. `ddh_impedance` builds a per-device table of the frequency-independent terms
(one wavefront per device), then solves every (device, frequency) system on a lane of its own,
contracting the linearised transient current on the way; `Y = dJ/dV` in SIMsalabim's sign and
the external `Z = 1/(Y + G_shunt) + R_series`. When any device of the batch has mobile ions,
the whole batch goes through `ddh_impedance_ions`, as the reference's: the ions' logarithms are
unknowns of their own (5 x 5 blocks), each species' number kept by a slot, and the ions'
currents join the linearised current.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .batch import SC_FIELDS, _sc_now, refuse_charged_traps

#: The ions' fluxes the reference's transients take (src/jaxdd/solver.py:1312).
ION_FLUXES = ("electrochemical", "modified-drift")


@dataclass(frozen=True)
class ImpedanceResult:
    """Per device and frequency (batch x frequencies): ``Z`` the external impedance (Ohm m^2,
    with the series resistance and the shunt), ``Y`` the internal admittance dJ/dV_int (S/m^2,
    SIMsalabim's sign); per device the steady state's ``J`` (A/m^2, with the shunt) and
    ``converged``. ``capacitance`` is Im(Y)/w (F/m^2)."""

    frequencies: np.ndarray
    Z: np.ndarray
    Y: np.ndarray
    J: np.ndarray
    converged: np.ndarray

    @property
    def capacitance(self) -> np.ndarray:
        return self.Y.imag / (2 * np.pi * self.frequencies)


def laps_of(frequencies, dt=None) -> np.ndarray:
    """`iw`, or what implicit Euler's steps of ``dt`` make of it, `(1 - e^(-iw dt))/dt`
    (reference: src/jaxdd/impedance.py:397-398)."""
    w = 2 * np.pi * np.asarray(frequencies, dtype=float)
    return 1j * w if dt is None else -np.expm1(-1j * w * dt) / dt


def external_impedance(Y, G_shunt, R_series) -> np.ndarray:
    """`Z = 1/(Y + G_shunt) + R_series` per device (rows) and frequency (reference:
    src/jaxdd/impedance.py:405)."""
    Y = np.asarray(Y)
    return 1.0 / (Y + np.asarray(G_shunt)[:, None]) + np.asarray(R_series)[:, None]


def solve_impedance(batch, frequencies, voltage=0.0, tol=1e-10, max_iter=60,
                    bernoulli_name="exact", ion_flux="electrochemical", dt=None, lanes=None,
                    events=None, return_states=False) -> ImpedanceResult:
    """The impedance of a batch (`jaxdd_amd.batch.prepare`) at the internal voltage ``voltage``
    (V) at each of ``frequencies`` (Hz), from the steady state there (solved with ``tol`` and
    ``max_iter``, as `solve_batch` at that one voltage) by one linear solve per frequency. The
    bulk traps' occupations follow at their own rates; mobile ions move with the voltage, their
    steady state the free ions' (Boltzmann's in each region). ``dt`` (s): instead of the
    continuous response, what implicit Euler's steps of ``dt`` see (the time terms at
    `(1 - e^(-iw dt))/dt`, the traps' occupations over each step). ``lanes``: the steady solve's
    (`solve_batch`'s; default four, the grid in quarters, for NP >= 16 and no mobile ions, else
    two). ``events``: four torch.cuda.Event recorded on the launch stream around the steady
    solve and around the impedance kernels. ``return_states``: also the steady states
    (n, NP, 3) the admittance was taken at.

    Raises `ValueError` for an unknown ``ion_flux`` or a ``dt`` that is not positive, and
    `NotImplementedError` for ions frozen at a pre-bias (as the reference), the modified-drift
    flux (ions bounded by their sites are outside this build's device model, and without sites
    both fluxes are Scharfetter-Gummel's), and devices whose ions reach a contact or fill
    several regions of one species."""
    if ion_flux not in ION_FLUXES:
        raise ValueError(f"ion_flux {ion_flux!r}: one of {ION_FLUXES}")
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    if dt is not None and not (np.isfinite(dt) and dt > 0):
        raise ValueError(f"dt {dt!r}: a positive step (s), or None for the continuous response")
    refuse_charged_traps(batch, "the impedance")
    sc = batch["_sc"]
    ions = bool(((sc[SC_FIELDS.index("a_total")] > 0) | (sc[SC_FIELDS.index("c_total")] > 0)).any())
    if ions and bool(((batch["pre_cond"] > 0) & (batch["fix_ions"] > 0)).any()):
        raise NotImplementedError("impedance with ions frozen at a pre-bias")
    if ions and ion_flux != "electrochemical":
        raise NotImplementedError("the modified drift differs from the electrochemical flux "
                                  "only for ions bounded by their sites, outside this path")
    if lanes is None:
        lanes = 4 if batch["NP"] >= 16 and not ions else 2
    if lanes not in (1, 2, 4):
        raise ValueError(f"lanes {lanes!r}: 1, 2 or 4")
    frequencies = np.atleast_1d(np.asarray(frequencies, dtype=float))
    if lanes == 4:
        from .batch import check_four_lanes

        check_four_lanes(batch)
    lib = _lib.load()
    dev = batch["_pt"].device
    if dev.type != "cuda":
        raise RuntimeError("solve_impedance runs on the GPU: prepare(devices, device='cuda')")
    n, NP, nf = batch["n"], batch["NP"], len(frequencies)
    bern = 0 if bernoulli_name == "exact" else 1
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), 0, bern, _lib.DDH_FP64, 0,
                     int(lanes))
    scn = _sc_now(batch)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), scn.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    Vd = torch.full((1,), float(voltage), dtype=torch.float64, device=dev)
    J = torch.empty((n, 1), dtype=torch.float64, device=dev)
    ok = torch.zeros((n, 1), dtype=torch.uint8, device=dev)
    its = torch.zeros((n, 1), dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    states = torch.empty((1, NP, 3, n), dtype=torch.float64, device=dev)
    nbytes = lib.ddh_workspace_bytes(n, NP, 1, _lib.DDH_FP64)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    laps = laps_of(frequencies, dt)
    lap_d = torch.as_tensor(np.ascontiguousarray(np.stack([laps.real, laps.imag], -1)),
                            device=dev)
    Vint = torch.full((n,), float(voltage), dtype=torch.float64, device=dev)
    Y = torch.empty((n, nf, 2), dtype=torch.float64, device=dev)
    istatus = torch.zeros(n, dtype=torch.int32, device=dev)
    ibytes = (lib.ddh_impedance_ions_workspace_bytes if ions
              else lib.ddh_impedance_workspace_bytes)(n, NP)
    iws = torch.empty(ibytes, dtype=torch.uint8, device=dev)
    if events is not None:
        events[0].record()
    _lib.check(lib.ddh_solve_batch_states(ctypes.byref(b), Vd.data_ptr(), 1, ctypes.byref(opts),
                                          J.data_ptr(), ok.data_ptr(), its.data_ptr(),
                                          status.data_ptr(), states.data_ptr(), None,
                                          ws.data_ptr(), nbytes, stream), "ddh_solve_batch_states")
    if events is not None:
        events[1].record()
        events[2].record()
    dt_ = 0.0 if dt is None else float(dt)
    if ions:
        _lib.check(lib.ddh_impedance_ions(ctypes.byref(b), batch["_imu"].data_ptr(),
                                          states.data_ptr(), Vint.data_ptr(),
                                          batch["_iface"].data_ptr(), lap_d.data_ptr(), nf, dt_,
                                          bern, Y.data_ptr(), istatus.data_ptr(), iws.data_ptr(),
                                          ibytes, stream), "ddh_impedance_ions")
    else:
        _lib.check(lib.ddh_impedance(ctypes.byref(b), states.data_ptr(), Vint.data_ptr(),
                                     batch["_iface"].data_ptr(), lap_d.data_ptr(), nf, dt_, bern,
                                     Y.data_ptr(), istatus.data_ptr(), iws.data_ptr(), ibytes,
                                     stream), "ddh_impedance")
    if events is not None:
        events[3].record()
    torch.cuda.current_stream(dev).synchronize()
    if int(istatus.max()) != 0:
        raise NotImplementedError("a device whose ions reach a contact or fill several regions "
                                  "of one species is outside this path")
    Yc = Y.cpu().numpy()
    Yc = Yc[..., 0] + 1j * Yc[..., 1]
    G_shunt = batch["G_shunt"].cpu().numpy()
    R_series = batch["R_series"].cpu().numpy()
    Jh = J[:, 0].cpu().numpy() + G_shunt * float(voltage)
    res = ImpedanceResult(frequencies=frequencies, Z=external_impedance(Yc, G_shunt, R_series),
                          Y=Yc, J=Jh, converged=ok[:, 0].bool().cpu().numpy())
    if return_states:
        return res, states[0].permute(2, 0, 1).contiguous()
    return res


__all__ = ["ION_FLUXES", "ImpedanceResult", "external_impedance", "laps_of", "solve_impedance"]
