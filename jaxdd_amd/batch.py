"""Many devices at once: `prepare` and `solve_batch`, the reference's batched API
(src/jaxdd/batch.py:151-223 `prepare`, :1136-1162 `solve_batch`, :1104-1108
`external_voltages`), solved by the HIP kernels of ``libddhip.so``.

`prepare` packs a list of devices into device-resident (HBM) torch tensors in the layout the
kernels read (jaxdd_amd/csrc/ddhip_layout.h): per grid point the spacing, position,
generation and fixed ions; per segment (a run of points with one material and one interval
mobility) the material parameters; per device the scalars. The returned dict also holds
the reference's per-device keys (``J_scale``, ``g_frac``, ``R_series``, ``G_shunt``, ...),
which `solve_batch` reads at every call, so ``dict(batch, g_frac=...)`` works as there.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .constants import Q
from .device import Device

__all__ = ["external_voltages", "from_packed", "merge", "pack", "prepare", "solve_batch",
           "solve_states"]

# Field orders of jaxdd_amd/csrc/ddhip_layout.h.
PT_FIELDS = ("h", "x", "G", "fix")
SEG_FIELDS = ("eps", "offn", "offp", "Nc", "dop", "kd", "ni2", "Nt", "Cn", "Cp", "nt0", "pt0",
              "am", "cm", "mun", "mup", "Ec", "ni", "sP", "sCs", "imun", "imup", "tt")
SC_FIELDS = ("L", "Vt", "g_frac", "V0", "VL", "W_L", "W_R", "ln_n_left", "ln_n_right",
             "ln_p_left", "ln_p_right", "left_elec", "a_total", "c_total", "V_pre", "J_scale",
             "G_shunt", "n_floor", "pre_cond", "fix_ions")
DENSITY_FLOOR = 1e-12  # src/jaxdd/solver.py:54
MAX_SEGMENTS = 14  # DDH_MAX_SEGMENTS of include/ddhip.h


def _point_fields(dev: Device) -> np.ndarray:
    """(NP + 2, 23) material parameters per point, in `SEG_FIELDS` order: the device's own
    and, derived from them, the row scales (src/jaxdd/solver.py:245-248, the continuity
    scale without its 1/h), the inverse interval mobilities and the bulk traps' charge type
    (1 donor, -1 acceptor, 0 neutral or none)."""
    with np.errstate(divide="ignore"):
        return np.stack([
            dev.eps, dev.offset_n, dev.offset_p, dev.N_c, dev.N_A - dev.N_D, dev.k_direct,
            dev.ni ** 2, dev.srh_Nt, dev.srh_Cn, dev.srh_Cp, dev.srh_nt0, dev.srh_pt0,
            np.asarray(dev.anion_mask, dtype=float), np.asarray(dev.cation_mask, dtype=float),
            dev.mu_n, dev.mu_p, dev.E_c, dev.ni,
            1.0 / (Q * dev.N_c), 1.0 / (np.maximum(dev.mu_n, dev.mu_p) * dev.Vt * dev.N_c),
            1.0 / dev.mu_n, 1.0 / dev.mu_p,
            np.zeros(len(dev.x)) if dev.srh_type is None else dev.srh_type,
        ], axis=-1).astype(np.float64)


def j_scale(dev: Device) -> float:
    """The current scale (src/jaxdd/batch.py:96, 115): the generation's current, >= 1 A/m^2."""
    generated = Q * dev.L * np.sum(dev.h[:-1] * dev.G[:-1]) / np.sum(dev.h[:-1])
    return max(float(generated), 1.0)


def _scalars(dev: Device) -> list[float]:
    return [
        dev.L, dev.Vt, dev.G_frac, dev.V0, dev.VL, dev.W_L, dev.W_R,
        np.log(dev.n_left), np.log(dev.n_right), np.log(dev.p_left), np.log(dev.p_right),
        float(dev.left_elec), dev.anion_total, dev.cation_total, dev.V_pre, j_scale(dev),
        1.0 / dev.R_shunt if dev.R_shunt > 0 else 0.0,
        DENSITY_FLOOR * float(np.max(dev.N_c)), float(dev.pre_cond), float(dev.fix_ions),
    ]


def charged_traps(batch) -> bool:
    """Whether any device of a batch (`prepare`'s dict, or a list of devices) holds charged
    bulk traps."""
    if isinstance(batch, dict):
        seg = batch.get("_seg")
        return seg is not None and bool((seg[:, SEG_FIELDS.index("tt"), :] != 0).any())
    return any(d.srh_type is not None and np.any(d.srh_type != 0) for d in batch)


def refuse_charged_traps(batch, what: str) -> None:
    """`NotImplementedError` where a batch holds charged bulk traps and ``what`` does not take
    them (their charge enters the steady solve's Poisson rows only)."""
    if charged_traps(batch):
        raise NotImplementedError(f"{what}: charged bulk traps (bulkTrapType 1 or -1) are "
                                  "solved in the steady state only")


def _check(devices) -> None:
    for d in devices:
        if d.statistics != "boltzmann" or any(d.free) or d.curr_diff:
            raise NotImplementedError("outside this path: Boltzmann carriers, no free ends, the "
                                      "integral current")


def pack(devices: list[Device]) -> dict[str, np.ndarray]:
    """The kernels' arrays of a list of devices (host NumPy, lane-innermost)."""
    sizes = {len(d.x) for d in devices}
    if len(sizes) != 1:
        raise ValueError(f"a batch needs one grid size, got NP + 2 in {sorted(sizes)}")
    _check(devices)
    n = len(devices)
    P = next(iter(sizes))
    fields = np.stack([_point_fields(d) for d in devices])  # (n, P, 23)
    # the mobile ions' mobilities (read by the transient only) also split segments
    ion_mu = np.stack([np.stack([np.zeros(P) if d.mu_anion is None else d.mu_anion,
                                 np.zeros(P) if d.mu_cation is None else d.mu_cation], -1)
                       for d in devices])  # (n, P, 2)
    change = np.any(np.concatenate([fields, ion_mu], axis=2)[:, 1:]
                    != np.concatenate([fields, ion_mu], axis=2)[:, :-1], axis=2)
    sid = np.concatenate([np.zeros((n, 1), dtype=np.int64), np.cumsum(change, axis=1)], axis=1)
    nseg = int(sid.max()) + 1
    if nseg > MAX_SEGMENTS:
        raise NotImplementedError(f"{nseg} segments of constant material: at most {MAX_SEGMENTS} "
                                  "(the kernels keep the segment tables in LDS)")
    seg = np.zeros((nseg, len(SEG_FIELDS), n))
    first = np.concatenate([np.ones((n, 1), dtype=bool), change], axis=1)  # (n, P)
    dev_idx, pt_idx = np.nonzero(first)
    seg[sid[dev_idx, pt_idx], :, dev_idx] = fields[dev_idx, pt_idx]
    imu = np.zeros((nseg, 2, n))
    imu[sid[dev_idx, pt_idx], :, dev_idx] = ion_mu[dev_idx, pt_idx]
    # segments beyond a device's own count repeat its last (never read)
    last = sid[:, -1]
    for s in range(nseg):
        pad = np.flatnonzero(last < s)
        if pad.size:
            seg[s, :, pad] = seg[last[pad], :, pad]
            imu[s, :, pad] = imu[last[pad], :, pad]
    pt = np.stack([
        np.stack([d.h, d.x, d.G, d.anion_fixed - d.cation_fixed], axis=0) for d in devices
    ], axis=-1)  # (4, P, n)
    pt = np.ascontiguousarray(pt.transpose(1, 0, 2))  # (P, 4, n)
    sc = np.ascontiguousarray(np.array([_scalars(d) for d in devices]).T)  # (20, n)
    # the intervals that cross from one layer to the next (the current's interfaces,
    # src/jaxdd/batch.py:375-376)
    iface = np.zeros((P - 1, n), dtype=np.uint8)
    for k, d in enumerate(devices):
        iface[list(d.i1[:-1]), k] = 1
    # each species' fixed density apart (the kernels read their difference, `pt`'s "fix"): the
    # states' anion and cation profiles (`solve_states`)
    fixed = np.stack([np.stack([d.anion_fixed, d.cation_fixed]) for d in devices], axis=-1)
    return {"pt": pt, "sid": np.ascontiguousarray(sid.T.astype(np.uint8)), "seg": seg,
            "sc": sc, "nseg": nseg, "R_series": np.array([d.R_series for d in devices]),
            "iface": iface, "fixed": np.ascontiguousarray(fixed), "imu": imu}


def merge(parts: list[dict]) -> dict:
    """Packed arrays of several device lists joined along the lane axis (the segment tables
    padded to the largest count, the extra segments never read)."""
    nseg = max(p["nseg"] for p in parts)

    def padded(key):
        out = []
        for p in parts:
            a = p[key]
            if a.shape[0] < nseg:
                a = np.concatenate([a, np.repeat(a[-1:], nseg - a.shape[0], axis=0)])
            out.append(a)
        return np.concatenate(out, axis=2)

    out = {"nseg": nseg, "seg": padded("seg"), "imu": padded("imu")}
    for k in ("pt", "sid", "sc", "iface", "fixed"):
        out[k] = np.ascontiguousarray(np.concatenate([p[k] for p in parts], axis=-1))
    out["R_series"] = np.concatenate([p["R_series"] for p in parts])
    return out


def prepare(devices: list[Device], device: str | torch.device = "cuda") -> dict:
    """A batch (reference: src/jaxdd/batch.py:151-223): the devices packed for the kernels, in
    HBM on ``device``, with the reference's per-device keys. Raises `ValueError` on devices of
    different grid sizes, as the reference does."""
    return from_packed(pack(devices), device)


def from_packed(packed: dict, device: str | torch.device = "cuda") -> dict:
    """The batch of packed arrays (`pack`, `merge`), moved to ``device``."""
    dev = torch.device(device)
    n = packed["sc"].shape[1]
    NP = packed["pt"].shape[0] - 2

    def t(a, dtype=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)

    sc = packed["sc"]
    col = {name: sc[k] for k, name in enumerate(SC_FIELDS)}
    return {
        "n": n, "NP": NP, "nseg": packed["nseg"],
        "_pt": t(packed["pt"]), "_sid": t(packed["sid"], torch.uint8), "_seg": t(packed["seg"]),
        "_sc": t(sc), "_iface": t(packed["iface"], torch.uint8), "_imu": t(packed["imu"]),
        # the reference's per-device keys (read by solve_batch at each call)
        "J_scale": t(col["J_scale"]), "g_frac": t(col["g_frac"]), "Vt": t(col["Vt"]),
        "L": t(col["L"]), "V_pre": t(col["V_pre"]), "pre_cond": t(col["pre_cond"]),
        "fix_ions": t(col["fix_ions"]), "left_elec": t(col["left_elec"]),
        "G_shunt": t(col["G_shunt"]), "R_series": t(packed["R_series"]),
        "x": t(packed["pt"][:, 1, :].T), "h": t(packed["pt"][:, 0, :].T),
        # host only: (2, NP + 2, n) fixed anions and cations, read by solve_states
        "_fixed": packed.get("fixed"),
    }


def check_four_lanes(batch: dict) -> None:
    """`ValueError` unless the four-lane kernel solves every device of the batch: no mobile
    ions, NP >= 16 (the kernel marks a device with ions status 2 and leaves it unsolved)."""
    sc = batch["_sc"]
    ions = (sc[SC_FIELDS.index("a_total")] > 0) | (sc[SC_FIELDS.index("c_total")] > 0)
    if bool(ions.any()):
        raise ValueError("lanes=4 solves devices without mobile ions only")
    if batch["NP"] < 16:
        raise ValueError(f"lanes=4 needs NP >= 16, got {batch['NP']}")


def _sc_now(batch: dict) -> torch.Tensor:
    """The packed scalars with the per-device keys the caller may have replaced."""
    sc = batch["_sc"].clone()
    for name in ("g_frac", "J_scale", "V_pre"):
        sc[SC_FIELDS.index(name)] = batch[name].to(sc.device, torch.float64)
    return sc


def _run(batch: dict, V: np.ndarray, opts: _lib.Opts, precision: int, events=None,
         states: torch.Tensor | None = None):
    """One call of ddh_solve_batch: (J, ok, its, status) as (n, nV) tensors on the batch's
    device, V as given (sorted by the caller in batch mode). ``events``: a pair of
    torch.cuda.Event recorded on the launch stream just before and after the solve kernel.
    ``states``: a (nV, NP, 3, n) float64 tensor on the batch's device to receive each
    voltage's final state (ddh_solve_batch_states)."""
    lib = _lib.load()
    dev = batch["_pt"].device
    if dev.type != "cuda":
        raise RuntimeError("solve_batch runs on the GPU: prepare(devices, device='cuda')")
    n, NP, nV = batch["n"], batch["NP"], len(V)
    sc = _sc_now(batch)
    Vd = torch.as_tensor(np.ascontiguousarray(V, dtype=np.float64), device=dev)
    J = torch.empty((n, nV), dtype=torch.float64, device=dev)
    ok = torch.zeros((n, nV), dtype=torch.uint8, device=dev)
    its = torch.zeros((n, nV), dtype=torch.int32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    nbytes = lib.ddh_workspace_bytes(n, NP, nV, precision)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    if events is not None:
        events[0].record()
    if states is None:
        rc = lib.ddh_solve_batch(ctypes.byref(b), Vd.data_ptr(), nV, ctypes.byref(opts),
                                 J.data_ptr(), ok.data_ptr(), its.data_ptr(), status.data_ptr(),
                                 ws.data_ptr(), nbytes, stream)
        _lib.check(rc, "ddh_solve_batch")
    else:
        if tuple(states.shape) != (nV, NP, 3, n) or states.dtype != torch.float64:
            raise ValueError(f"states: float64 of shape {(nV, NP, 3, n)}")
        rc = lib.ddh_solve_batch_states(ctypes.byref(b), Vd.data_ptr(), nV, ctypes.byref(opts),
                                        J.data_ptr(), ok.data_ptr(), its.data_ptr(),
                                        status.data_ptr(), states.data_ptr(), None,
                                        ws.data_ptr(), nbytes, stream)
        _lib.check(rc, "ddh_solve_batch_states")
    if events is not None:
        events[1].record()
    # the workspace and the inputs live until the kernels are done
    batch["_inflight"] = (ws, sc, Vd)
    return J, ok, its, status


def solve_batch(batch, voltages, tol=1e-10, max_iter=60, first_max_iter=6,
                bernoulli_name="exact", strict=False, devices=None, precision="fp64",
                events=None, lanes=2):
    """J-V curves of a batch at ``voltages`` (reference: src/jaxdd/batch.py:1136-1162):
    ``(J, converged, iterations)``, each batch x voltages, in the order of ``voltages``, as
    torch tensors on the batch's GPU (float64, bool, int64).

    Each device's scan runs outwards from its pre-bias, as the reference's. The reference's two
    passes (``first_max_iter``, then the stragglers with ``max_iter``) give every device the
    result of one pass with ``max_iter``; here each device runs on its own lane of a
    persistent kernel, so no device waits for another and one pass with ``max_iter`` is run.
    ``precision``: "fp64", or "mixed" (the Newton correction's block factorisation in FP32 on
    row-equilibrated blocks; residual, exponentials, current and convergence tests in FP64).
    ``devices`` is accepted for the reference's signature: one GPU solves the whole batch.
    ``events``: optional (start, end) torch.cuda.Event pair timing the solve kernel.
    ``lanes``: 2 (two GPU lanes per device, each Newton solve split at the grid's middle),
    1 (one lane sweeping the whole grid, the reference's order of elimination) or 4 (four
    lanes, the grid split in quarters: for batches too small to fill the GPU at two lanes;
    devices without mobile ions only, `ValueError` otherwise)."""
    del first_max_iter, devices
    if precision not in _lib.PRECISIONS:
        raise ValueError(f"precision {precision!r}: one of {tuple(_lib.PRECISIONS)}")
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    V = np.asarray(voltages, dtype=float).ravel()
    order = np.argsort(V, kind="stable")
    if lanes not in (1, 2, 4):
        raise ValueError(f"lanes {lanes!r}: 1, 2 or 4")
    if lanes == 4:
        check_four_lanes(batch)
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), int(bool(strict)),
                     0 if bernoulli_name == "exact" else 1, _lib.PRECISIONS[precision], 0,
                     int(lanes))
    J, ok, its, _ = _run(batch, V[order], opts, _lib.PRECISIONS[precision], events)
    inv = torch.as_tensor(np.argsort(order, kind="stable"), device=J.device)
    J, ok, its = J[:, inv], ok[:, inv].bool(), its[:, inv].long()
    Vt = torch.as_tensor(V, device=J.device)
    return J + batch["G_shunt"][:, None] * Vt[None, :], ok, its


def _point_field(batch: dict, name: str) -> torch.Tensor:
    """(n, NP + 2) one material field (`SEG_FIELDS`) on every point of every device, read
    through the points' segment ids."""
    seg = batch["_seg"][:, SEG_FIELDS.index(name), :]  # (nseg, n)
    return torch.gather(seg, 0, batch["_sid"].long()).T  # (n, NP + 2)


def solve_states(batch, voltages=0.0, tol=1e-10, max_iter=60, first_max_iter=6,
                 bernoulli_name="exact", strict=False, devices=None, precision="fp64", lanes=2):
    """The steady states of a batch at ``voltages`` (one or several, internal) and what
    photoluminescence measures of them (reference: src/jaxdd/batch.py:1165-1191
    `solve_states`, with `_observed` :1009-1030 and src/jaxdd/solver.py:2096-2102
    `_ions_on_grid`). A dict of tensors on the batch's GPU: ``V``, ``n``, ``p``, ``qfls``,
    ``anion``, ``cation`` (batch x voltages x NP + 2: the contacts' points their boundary
    values, the ions' their neighbours'; ``V`` in its gauge, densities in m^-3, ``qfls`` in V),
    ``implied_voc`` (V) and ``plqy`` of direct recombination, ``J``, ``converged``,
    ``iterations`` (batch x voltages). Solved as `solve_batch` (same arguments and results):
    the solve kernel keeps each voltage's final state (ddh_solve_batch_states). Mobile ions
    take Boltzmann's normalisation of their region's total at the solved potential (closed
    form, src/jaxdd/solver.py:1695-1712); ions frozen after a pre-bias take the profile of the
    free-ion solve at the pre-bias, as the reference's freeze does (batch.py:859-872)."""
    del first_max_iter, devices
    if precision not in _lib.PRECISIONS:
        raise ValueError(f"precision {precision!r}: one of {tuple(_lib.PRECISIONS)}")
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    if lanes not in (1, 2, 4):
        raise ValueError(f"lanes {lanes!r}: 1, 2 or 4")
    if lanes == 4:
        check_four_lanes(batch)
    V = np.atleast_1d(np.asarray(voltages, dtype=float)).ravel()
    order = np.argsort(V, kind="stable")
    n, NP, nV = batch["n"], batch["NP"], len(V)
    dev = batch["_pt"].device
    prec = _lib.PRECISIONS[precision]
    opts = _lib.Opts(float(tol), int(max_iter), int(max_iter), int(bool(strict)),
                     0 if bernoulli_name == "exact" else 1, prec, 0, int(lanes))
    st = torch.empty((nV, NP, 3, n), dtype=torch.float64, device=dev)
    J, ok, its, _ = _run(batch, V[order], opts, prec, states=st)
    inv = torch.as_tensor(np.argsort(order, kind="stable"), device=dev)
    J, ok, its = J[:, inv], ok[:, inv].bool(), its[:, inv].long()
    Vt_in = torch.as_tensor(V, device=dev)
    J = J + batch["G_shunt"][:, None] * Vt_in[None, :]
    y = st[inv.cpu()].permute(3, 0, 1, 2)  # (n, nV, NP, 3) in the caller's order
    full = _full_states(batch, y, Vt_in)
    an, ca = _ions_on_grid(batch, full[..., 0])
    frozen = _sc_now(batch)[SC_FIELDS.index("pre_cond")] * batch["fix_ions"] > 0
    if bool(frozen.any()):  # the pre-bias's free-ion profile, held at every voltage
        an_f, ca_f = _frozen_ions(batch, frozen, tol, max_iter, bernoulli_name, strict,
                                  precision, lanes)
        an = torch.where(frozen[:, None, None], an_f[:, None, :], an)
        ca = torch.where(frozen[:, None, None], ca_f[:, None, :], ca)
    lnn, lnp = full[..., 1], full[..., 2]
    nn, pp = torch.exp(lnn), torch.exp(lnp)
    ni2 = _point_field(batch, "ni2")[:, None, :]
    Vt = batch["Vt"][:, None]
    qfls = Vt[..., None] * torch.log(nn * pp / ni2)
    # control volumes of the interior points, direct recombination and generation
    h, L = batch["h"], batch["L"][:, None]
    vol = 0.5 * L * (h[:, :-2] + h[:, 1:-1])  # (n, NP)
    kd = _point_field(batch, "kd")[:, 1:-1]
    G = batch["_pt"][:, PT_FIELDS.index("G"), :].T[:, 1:-1]
    rad = torch.sum((vol * kd)[:, None, :] * nn[..., 1:-1] * pp[..., 1:-1], dim=-1)
    rad0 = torch.sum(vol * kd * ni2[:, 0, 1:-1], dim=-1)[:, None]
    gen = (batch["g_frac"] * torch.sum(vol * G, dim=-1))[:, None]
    nan = torch.full_like(rad, float("nan"))
    implied = torch.where(rad0 > 0, Vt * torch.log(rad / torch.where(rad0 > 0, rad0, 1.0)), nan)
    plqy = torch.where(gen > 0, (rad - rad0) / torch.where(gen > 0, gen, 1.0), nan)
    return {"V": full[..., 0], "n": nn, "p": pp, "qfls": qfls, "anion": an, "cation": ca,
            "implied_voc": implied, "plqy": plqy, "J": J, "converged": ok, "iterations": its}


def _full_states(batch: dict, y: torch.Tensor, Vint: torch.Tensor) -> torch.Tensor:
    """(n, nV, NP + 2, 3) states on every point: the interior's ``y`` (n, nV, NP, 3) between
    the contacts' boundary values at internal voltages ``Vint`` (nV,) (src/jaxdd/batch.py:255-267
    `_bounds`, `_full`)."""
    sc = batch["_sc"]

    def col(name):
        return sc[SC_FIELDS.index(name)][:, None]

    elec = col("left_elec")
    n, nV = y.shape[0], y.shape[1]
    left = torch.stack([col("V0") + 0.5 * elec * Vint[None, :],
                        col("ln_n_left").expand(n, nV), col("ln_p_left").expand(n, nV)], -1)
    right = torch.stack([col("VL") - 0.5 * elec * Vint[None, :],
                         col("ln_n_right").expand(n, nV), col("ln_p_right").expand(n, nV)], -1)
    return torch.cat([left[:, :, None], y, right[:, :, None]], dim=2)


def _ions_on_grid(batch: dict, Vfull: torch.Tensor):
    """Anions and cations (n, nV, NP + 2) at potentials ``Vfull`` (n, nV, NP + 2): fixed,
    plus Boltzmann's profile where mobile with each region's closed-form normalisation of
    its total, trapezoidal over the region (src/jaxdd/solver.py:951-961, :1241-1252,
    :1695-1712); the contacts' points their neighbours' (:2096-2102)."""
    sc = _sc_now(batch)
    Vt = sc[SC_FIELDS.index("Vt")][:, None, None]
    hL = (batch["h"] * batch["L"][:, None])[:, None, :-1]  # (n, 1, NP + 1)
    fixed = batch.get("_fixed")
    out = []
    for sign, mask_name, total_name, k in ((1.0, "am", "a_total", 0), (-1.0, "cm", "c_total", 1)):
        m = _point_field(batch, mask_name)[:, None, :]  # (n, 1, NP + 2)
        total = sc[SC_FIELDS.index(total_name)][:, None]
        prof = m * torch.exp(sign * Vfull / Vt)
        pair = (m[..., 1:] * m[..., :-1])
        number = torch.sum(0.5 * (prof[..., 1:] + prof[..., :-1]) * pair * hL, dim=-1)
        has = total > 0
        scale = torch.where(has, total / torch.where(has & (number > 0), number, 1.0), 0.0)
        dens = prof * scale[..., None]
        if fixed is not None:
            dens = dens + torch.as_tensor(fixed[k].T, device=dens.device)[:, None, :]
        dens = torch.cat([dens[..., 1:2], dens[..., 1:-1], dens[..., -2:-1]], dim=-1)
        out.append(dens)
    return out[0], out[1]


def _frozen_ions(batch, frozen, tol, max_iter, bernoulli_name, strict, precision, lanes):
    """(n, NP + 2) anions and cations of the free-ion steady state at each device's pre-bias:
    the profiles the reference's freeze holds (src/jaxdd/batch.py:859-872, 955-960)."""
    sc = batch["_sc"].clone()
    sc[SC_FIELDS.index("pre_cond")] = 0.0
    sc[SC_FIELDS.index("fix_ions")] = 0.0
    free = dict(batch, _sc=sc, pre_cond=torch.zeros_like(batch["pre_cond"]),
                fix_ions=torch.zeros_like(batch["fix_ions"]))
    # one solve per distinct pre-bias (the reference's V_pre is a per-device scalar)
    V_pre = batch["V_pre"]
    an = torch.zeros((batch["n"], batch["NP"] + 2), dtype=torch.float64, device=V_pre.device)
    ca = torch.zeros_like(an)
    for v in torch.unique(V_pre[frozen]).tolist():
        out = solve_states(free, [v], tol, max_iter, None, bernoulli_name, strict, None,
                           precision, lanes)
        sel = frozen & (V_pre == v)
        an[sel], ca[sel] = out["anion"][sel, 0], out["cation"][sel, 0]
    return an, ca


def assemble(batch, y, Vint, logs=None, bernoulli_name="exact"):
    """The residual, Jacobian blocks, ion columns and current of every device at a state
    (reference: src/jaxdd/solver.py:1021-1187 `_assemble` per device, with
    src/jaxdd/batch.py:518-549 `_system`'s ion columns and src/jaxdd/batch.py:302-403's
    current). ``y``: (n, NP, 3) interior (V, ln n, ln p); ``Vint``: (n,) internal voltages;
    ``logs``: (n, 2) ion normalisations or None. Returns a dict of tensors on the batch's
    GPU: F (n, NP, 3) row-scaled residual, A, B, C (n, NP, 3, 3), U (n, NP, 3, 2),
    J (n,) current with the shunt term (``G_shunt * Vint``) added, as solve_batch's."""
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    lib = _lib.load()
    dev = batch["_pt"].device
    if dev.type != "cuda":
        raise RuntimeError("assemble runs on the GPU: prepare(devices, device='cuda')")
    n, NP = batch["n"], batch["NP"]
    yt = torch.as_tensor(y, dtype=torch.float64, device=dev)
    if tuple(yt.shape) != (n, NP, 3):
        raise ValueError(f"y: shape (n, NP, 3) = {(n, NP, 3)}, got {tuple(yt.shape)}")
    yl = yt.permute(1, 2, 0).contiguous()  # (NP, 3, n): lane-innermost
    Vd = torch.as_tensor(Vint, dtype=torch.float64, device=dev).reshape(n).contiguous()
    lg = None
    if logs is not None:
        lg = torch.as_tensor(logs, dtype=torch.float64, device=dev).reshape(n, 2).T.contiguous()
    sc = _sc_now(batch)
    F = torch.empty((NP, 3, n), dtype=torch.float64, device=dev)
    A, B, C = (torch.empty((NP, 9, n), dtype=torch.float64, device=dev) for _ in range(3))
    U = torch.empty((NP, 6, n), dtype=torch.float64, device=dev)
    J = torch.empty(n, dtype=torch.float64, device=dev)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = lib.ddh_assemble(ctypes.byref(b), yl.data_ptr(), Vd.data_ptr(),
                          lg.data_ptr() if lg is not None else None,
                          0 if bernoulli_name == "exact" else 1, F.data_ptr(), A.data_ptr(),
                          B.data_ptr(), C.data_ptr(), U.data_ptr(), J.data_ptr(), stream)
    _lib.check(rc, "ddh_assemble")
    torch.cuda.current_stream(dev).synchronize()
    return {
        "F": F.permute(2, 0, 1), "A": A.reshape(NP, 3, 3, n).permute(3, 0, 1, 2),
        "B": B.reshape(NP, 3, 3, n).permute(3, 0, 1, 2),
        "C": C.reshape(NP, 3, 3, n).permute(3, 0, 1, 2),
        "U": U.reshape(NP, 3, 2, n).permute(3, 0, 1, 2),
        "J": J + batch["G_shunt"] * Vd,
    }


def assemble_ions(batch, y, Vint, logs, bernoulli_name="exact"):
    """The steady 5x5 system with the mobile ions as unknowns of every device (reference:
    src/jaxdd/solver.py:1364-1473, `_assemble_ions` with dti = 0). ``y``: (n, NP, 5) interior
    (V, ln n, ln p, ln a, ln c); ``Vint``: (n,); ``logs``: (n, 2) the anion and cation slots'
    normalisations. Returns tensors on the batch's GPU: F (n, NP, 5), A, B, C (n, NP, 5, 5),
    U (n, NP, 5, 2)."""
    if bernoulli_name not in ("exact", "simss"):
        raise ValueError(f"bernoulli_name {bernoulli_name!r}: 'exact' or 'simss'")
    lib = _lib.load()
    dev = batch["_pt"].device
    if dev.type != "cuda":
        raise RuntimeError("assemble_ions runs on the GPU: prepare(devices, device='cuda')")
    n, NP = batch["n"], batch["NP"]
    yt = torch.as_tensor(y, dtype=torch.float64, device=dev)
    if tuple(yt.shape) != (n, NP, 5):
        raise ValueError(f"y: shape (n, NP, 5) = {(n, NP, 5)}, got {tuple(yt.shape)}")
    yl = yt.permute(1, 2, 0).contiguous()  # (NP, 5, n): lane-innermost
    Vd = torch.as_tensor(Vint, dtype=torch.float64, device=dev).reshape(n).contiguous()
    lg = torch.as_tensor(logs, dtype=torch.float64, device=dev).reshape(n, 2).T.contiguous()
    sc = _sc_now(batch)
    F = torch.empty((NP, 5, n), dtype=torch.float64, device=dev)
    A, B, C = (torch.empty((NP, 25, n), dtype=torch.float64, device=dev) for _ in range(3))
    U = torch.empty((NP, 10, n), dtype=torch.float64, device=dev)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = lib.ddh_assemble_ions(ctypes.byref(b), yl.data_ptr(), Vd.data_ptr(), lg.data_ptr(),
                               0 if bernoulli_name == "exact" else 1, F.data_ptr(), A.data_ptr(),
                               B.data_ptr(), C.data_ptr(), U.data_ptr(), stream)
    _lib.check(rc, "ddh_assemble_ions")
    torch.cuda.current_stream(dev).synchronize()
    return {
        "F": F.permute(2, 0, 1), "A": A.reshape(NP, 5, 5, n).permute(3, 0, 1, 2),
        "B": B.reshape(NP, 5, 5, n).permute(3, 0, 1, 2),
        "C": C.reshape(NP, 5, 5, n).permute(3, 0, 1, 2),
        "U": U.reshape(NP, 5, 2, n).permute(3, 0, 1, 2),
    }


def external_voltages(batch, voltages, J):
    """``V_ext = V_int + J * R_series`` (reference: src/jaxdd/batch.py:1104-1108)."""
    Vt = torch.as_tensor(np.asarray(voltages, dtype=float), device=J.device)
    return Vt[None, :] + J * batch["R_series"][:, None]
