"""Transients of a batch: ZimT 5.36's time steps solved by the HIP transient kernel (reference:
src/jaxdd/transient.py:537-637 `solve_transient`, `TransientResult` :94-148).

Each step is one Newton solve of the whole system with implicit Euler in time from the state of
the step before; a step at t = 0 is a steady state. The mobile ions are unknowns of their own
(5 x 5 blocks) with one normalisation slot per species; the bulk traps' occupations are carried
from step to step, each solved exactly over its step. The current is the integral form with the
densities' change, the ions' currents and the displacement current. ZimT's `failureMode`
decides which steps are kept and each step's dt follows from the last step kept.

Outside this path (refused with `NotImplementedError`): an open circuit or a series resistance
(the internal voltage as an unknown), ZimT's automatic stop (`Track` other than 0), a pre-bias,
ions reaching a contact, a start from a given state.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .batch import SC_FIELDS, _run, _sc_now, refuse_charged_traps
from .simss_input import TVG

__all__ = ["TransientResult", "solve_transient"]

#: The HIP transient kernel stays off until the memory access fault of its first GPU run has a
#: cause shown and fixed; the same per-lane code runs in the CPU replay of the tests.
GPU_TRANSIENT_ENABLED = False


@dataclass(frozen=True)
class TransientResult:
    """Per device and step (batch x steps): ``J`` the internal current density (A/m^2, ZimT's
    `Jint`), ``J_ext`` with the shunt (`Jext`), the internal and external voltages, converged,
    the Newton steps, ``accepted`` the steps ZimT keeps (the rows of its tJFile), ``t`` each
    step's time (s). NumPy arrays, as the reference returns them."""

    J: np.ndarray
    J_ext: np.ndarray
    V_int: np.ndarray
    V_ext: np.ndarray
    converged: np.ndarray
    iterations: np.ndarray
    accepted: np.ndarray
    t: np.ndarray


def _steps(tvg, n):
    tvgs = list(tvg) if isinstance(tvg, (list, tuple)) else [tvg] * n
    if len(tvgs) != n or len({len(x.t) for x in tvgs}) != 1:
        raise ValueError("one tVG for the batch, or one per device, all of one length")
    for x in tvgs:
        if np.any(x.oc):
            raise NotImplementedError("an open circuit (the internal voltage as an unknown) is "
                                      "outside this path")
        if np.any(np.asarray(x.track) != 0):
            raise NotImplementedError("ZimT's automatic stop (Track 1-3) is outside this path")
    if len({x.failure_mode for x in tvgs}) != 1:
        raise NotImplementedError("one failureMode for the whole batch")
    t, V, G = (np.ascontiguousarray(np.stack([np.asarray(getattr(x, k), float) for x in tvgs]))
               for k in ("t", "Vext", "G_frac"))
    if np.any(t[:, 0] != 0):
        raise ValueError("a transient starts with a steady state (t = 0)")
    return t, V, G, tvgs[0].failure_mode


def solve_transient(batch, tvg: TVG | list[TVG], tol=1e-10, max_iter=200,
                    bernoulli_name="exact", zimt_current=False, events=None) -> TransientResult:
    """A batch (`batch.prepare`) through ZimT's time steps ``tvg`` (one for every device or one
    per device, all of one length; `simss_input.read_tvg`). Each step's Newton to ``tol`` or
    ``max_iter``. ``zimt_current``: the integral current's sums as ZimT 5.36 has them (each
    point's change over the interval after it); by default over each point's control volume.
    ``events``: (start, end) torch.cuda.Event pair timing the transient kernel. Returns a
    `TransientResult`. The steady start at the first step's voltage is solved first
    (`solve_batch`'s start, the same kernels)."""
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    refuse_charged_traps(batch, "the transient")
    if not GPU_TRANSIENT_ENABLED:
        raise NotImplementedError(
            "the GPU transient kernel is switched off: its first run on the MI355X ended in a "
            "memory access fault whose cause is not established (its code is verified on the "
            "CPU against the oracle, tests/test_transient.py)")
    lib = _lib.load()
    n, NP = batch["n"], batch["NP"]
    dev = batch["_pt"].device
    if dev.type != "cuda":
        raise RuntimeError("solve_transient runs on the GPU: prepare(devices, device='cuda')")
    t, V, G, failure_mode = _steps(tvg, n)
    nt = t.shape[1]
    if bool((batch["R_series"] > 0).any()):
        raise NotImplementedError("a series resistance (the internal voltage as an unknown) is "
                                  "outside this path")
    if bool((batch["pre_cond"] > 0).any()):
        raise NotImplementedError("a pre-bias is not a transient's start")
    V0 = np.unique(V[:, 0])
    if len(V0) != 1:
        raise NotImplementedError("every device's first step at one voltage")
    bern = 0 if bernoulli_name == "exact" else 1
    # the steady start at the first voltage and generation fraction (transient.py:307-313)
    start = dict(batch, g_frac=torch.as_tensor(G[:, 0], device=dev))
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), 0, bern, _lib.DDH_FP64, 0, 2)
    y0 = torch.empty((1, NP, 3, n), dtype=torch.float64, device=dev)
    _run(start, V0, opts, _lib.DDH_FP64, states=y0)
    sc = _sc_now(start)
    ions = int(bool(((sc[SC_FIELDS.index("a_total")] > 0)
                     | (sc[SC_FIELDS.index("c_total")] > 0)).any()))
    td, Vd, Gd = (torch.as_tensor(a, device=dev) for a in (t, V, G))
    J = torch.empty((n, nt), dtype=torch.float64, device=dev)
    Jx = torch.empty_like(J)
    ok = torch.zeros((n, nt), dtype=torch.uint8, device=dev)
    kept = torch.zeros_like(ok)
    its = torch.zeros((n, nt), dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    nbytes = lib.ddh_transient_workspace_bytes(n, NP)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    to = _lib.TOpts(float(tol), int(max_iter), bern, int(bool(zimt_current)), int(failure_mode))
    stream = torch.cuda.current_stream(dev).cuda_stream
    if events is not None:
        events[0].record()
    rc = lib.ddh_transient(ctypes.byref(b), batch["_imu"].data_ptr(), batch["_iface"].data_ptr(),
                           td.data_ptr(), Vd.data_ptr(), Gd.data_ptr(), nt, y0.data_ptr(),
                           ctypes.byref(to), ions, J.data_ptr(), Jx.data_ptr(), ok.data_ptr(),
                           its.data_ptr(), kept.data_ptr(), status.data_ptr(), ws.data_ptr(),
                           nbytes, stream)
    _lib.check(rc, "ddh_transient")
    if events is not None:
        events[1].record()
    torch.cuda.current_stream(dev).synchronize()
    if int(status.max()) != 0:
        raise NotImplementedError("a device outside the transient's path (ions reaching a "
                                  "contact)")
    Jx = Jx.cpu().numpy()
    R = batch["R_series"].cpu().numpy()[:, None]
    return TransientResult(J=J.cpu().numpy(), J_ext=Jx, V_int=V.copy(), V_ext=V + Jx * R,
                           converged=ok.cpu().numpy().astype(bool), iterations=its.cpu().numpy(),
                           accepted=kept.cpu().numpy().astype(bool), t=t.copy())
