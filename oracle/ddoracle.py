"""ctypes binding of the FP64 CPU oracle (libddoracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg load this module; the
product package (jaxdd_amd) never does. It mirrors the reference's batched entry point
(src/jaxdd/batch.py:1136-1162 `solve_batch` with `_passes` :1194-1208 and `_stragglers`
:1111-1118) on top of the C restatement of `_solve_one` (batch.py:937-1006), and
`solve_jv` (src/jaxdd/solver.py:2041-2060).
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

from jaxdd_amd.constants import Q

HERE = Path(__file__).resolve().parent
LIB = HERE / "libddoracle.so"

_P = ctypes.POINTER(ctypes.c_double)


class _Device(ctypes.Structure):
    _fields_ = [("NP", ctypes.c_int32)] + [
        (name, _P) for name in (
            "x", "h", "eps", "offset_n", "offset_p", "N_c", "ni", "E_c", "N_D", "N_A",
            "mu_n", "mu_p", "k_direct", "Nt", "Cn", "Cp", "nt0", "pt0", "G",
            "a_mask", "c_mask", "a_fixed", "c_fixed")
    ] + [("iface", ctypes.POINTER(ctypes.c_uint8))] + [
        (name, ctypes.c_double) for name in (
            "L", "Vt", "g_frac", "V0", "VL", "W_L", "W_R", "ln_n_left", "ln_n_right",
            "ln_p_left", "ln_p_right", "left_elec", "a_total", "c_total", "V_pre", "J_scale",
            "G_shunt")
    ] + [("pre_cond", ctypes.c_int32), ("fix_ions", ctypes.c_int32), ("mu_a", _P), ("mu_c", _P),
         ("tt", _P)]


class _Opts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("start_iter", ctypes.c_int32), ("strict", ctypes.c_int32),
                ("bernoulli", ctypes.c_int32)]


BERNOULLI = {"exact": 0, "simss": 1}


class _TOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("bernoulli", ctypes.c_int32), ("zimt_current", ctypes.c_int32),
                ("failure_mode", ctypes.c_int32)]


def build() -> Path:
    """Compile the oracle with its Makefile (gcc) and return the library's path."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    lib.ddo_bernoulli.restype = ctypes.c_double
    lib.ddo_bernoulli.argtypes = [ctypes.c_double, ctypes.c_int, _P]
    lib.ddo_current.restype = ctypes.c_double
    lib.ddo_current.argtypes = [ctypes.POINTER(_Device), _P, ctypes.c_double]
    lib.ddo_transient.restype = ctypes.c_int
    lib.ddo_transient.argtypes = [ctypes.POINTER(_Device), _P, _P, _P, ctypes.c_int,
                                  ctypes.POINTER(_TOpts), _P, _P, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p]
    lib.ddo_assemble_step.restype = ctypes.c_int
    lib.ddo_assemble_step.argtypes = [ctypes.POINTER(_Device), _P, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, _P, _P, _P, ctypes.c_int, ctypes.c_int,
                                      _P, _P, _P, _P, _P, _P, _P, _P]
    lib.ddo_noise_step.restype = ctypes.c_double
    lib.ddo_noise_step.argtypes = [ctypes.POINTER(_Device), _P, _P, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_double]
    lib.ddo_solve_one.argtypes = [ctypes.POINTER(_Device), _P, ctypes.c_int,
                                  ctypes.POINTER(_Opts), _P, ctypes.POINTER(ctypes.c_uint8),
                                  ctypes.POINTER(ctypes.c_int32)]
    lib.ddo_solve_one_states.argtypes = lib.ddo_solve_one.argtypes + [_P]
    lib.ddo_solve_jv.argtypes = [ctypes.POINTER(_Device), _P, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, _P, ctypes.POINTER(ctypes.c_uint8),
                                 ctypes.POINTER(ctypes.c_int32)]
    lib.ddo_assemble.argtypes = [ctypes.POINTER(_Device), _P, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, _P, _P, _P, _P, _P]
    lib.ddo_assemble_ions.argtypes = [ctypes.POINTER(_Device), _P, ctypes.c_double, _P,
                                      ctypes.c_int, _P, _P, _P, _P, _P]
    lib.ddo_block_thomas.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def j_scale(dev) -> float:
    """The current scale (batch.py:96, 115): the generation's current, at least 1 A/m^2."""
    gen = Q * dev.L * np.sum(dev.h[:-1] * dev.G[:-1]) / np.sum(dev.h[:-1])
    return max(gen, 1.0)


class _Pinned:
    """A device's arrays as contiguous float64, kept alive with the ctypes struct."""

    def __init__(self, dev, g_frac=None):
        NP = len(dev.x) - 2
        arrays = {
            "x": dev.x, "h": dev.h, "eps": dev.eps, "offset_n": dev.offset_n,
            "offset_p": dev.offset_p, "N_c": dev.N_c, "ni": dev.ni, "E_c": dev.E_c,
            "N_D": dev.N_D, "N_A": dev.N_A, "mu_n": dev.mu_n, "mu_p": dev.mu_p,
            "k_direct": dev.k_direct, "Nt": dev.srh_Nt, "Cn": dev.srh_Cn, "Cp": dev.srh_Cp,
            "nt0": dev.srh_nt0, "pt0": dev.srh_pt0, "G": dev.G,
            "a_mask": np.asarray(dev.anion_mask, dtype=float),
            "c_mask": np.asarray(dev.cation_mask, dtype=float),
            "a_fixed": dev.anion_fixed, "c_fixed": dev.cation_fixed,
            "mu_a": np.zeros(NP + 2) if dev.mu_anion is None else dev.mu_anion,
            "mu_c": np.zeros(NP + 2) if dev.mu_cation is None else dev.mu_cation,
            "tt": np.zeros(NP + 2) if dev.srh_type is None else dev.srh_type,
        }
        self.keep = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in arrays.items()}
        iface = np.zeros(NP + 1, dtype=np.uint8)
        for i in dev.i1[:-1]:
            iface[i] = 1
        self.keep["iface"] = iface
        s = _Device()
        s.NP = NP
        for k, v in self.keep.items():
            if k == "iface":
                s.iface = v.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            else:
                setattr(s, k, _ptr(v))
        s.L, s.Vt = dev.L, dev.Vt
        s.g_frac = dev.G_frac if g_frac is None else g_frac
        s.V0, s.VL, s.W_L, s.W_R = dev.V0, dev.VL, dev.W_L, dev.W_R
        s.ln_n_left, s.ln_n_right = np.log(dev.n_left), np.log(dev.n_right)
        s.ln_p_left, s.ln_p_right = np.log(dev.p_left), np.log(dev.p_right)
        s.left_elec = float(dev.left_elec)
        s.a_total, s.c_total = dev.anion_total, dev.cation_total
        s.V_pre = dev.V_pre
        s.J_scale = j_scale(dev)
        s.G_shunt = 1.0 / dev.R_shunt if dev.R_shunt > 0 else 0.0
        s.pre_cond, s.fix_ions = int(dev.pre_cond), int(dev.fix_ions)
        self.struct = s


def solve_one(dev, V_sorted, tol=1e-10, max_iter=60, start_iter=60, strict=False,
              bernoulli="exact", g_frac=None, return_states=False):
    """One device's curve at ascending voltages: (J without the shunt, converged, iterations);
    ``return_states``: also each voltage's final iterate (nV x NP x 3)."""
    V = np.ascontiguousarray(V_sorted, dtype=np.float64)
    n = len(V)
    J = np.zeros(n)
    ok = np.zeros(n, dtype=np.uint8)
    its = np.zeros(n, dtype=np.int32)
    pin = _Pinned(dev, g_frac)
    o = _Opts(tol, max_iter, start_iter, int(strict), BERNOULLI[bernoulli])
    states = np.zeros((n, len(dev.x) - 2, 3))
    rc = lib().ddo_solve_one_states(ctypes.byref(pin.struct), _ptr(V), n, ctypes.byref(o),
                                    _ptr(J), ok.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                    its.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                    _ptr(states))
    if rc != 0:
        raise ValueError(f"ddo_solve_one: error {rc}")
    if return_states:
        return J, ok.astype(bool), its, states
    return J, ok.astype(bool), its


def _one_pass(devices, V_sorted, tol, max_iter, start_iter, strict, bernoulli, g_frac):
    out = [solve_one(d, V_sorted, tol, max_iter, start_iter, strict, bernoulli,
                     None if g_frac is None else g_frac[k]) for k, d in enumerate(devices)]
    return (np.stack([o[0] for o in out]), np.stack([o[1] for o in out]),
            np.stack([o[2] for o in out]))


def solve_batch(devices, voltages, tol=1e-10, max_iter=60, first_max_iter=6,
                bernoulli_name="exact", strict=False, g_frac=None):
    """The reference's `solve_batch` semantics (batch.py:1136-1208) on the CPU, device by
    device: sorted voltages, two passes (``first_max_iter``, then the stragglers with
    ``max_iter``; the start always with ``max_iter``), results in the caller's order, the
    shunt added. Returns ``(J, converged, iterations)``."""
    V = np.asarray(voltages, dtype=float)
    order = np.argsort(V, kind="stable")
    Vs = V[order]
    if first_max_iter is None or first_max_iter >= max_iter:
        J, ok, its = _one_pass(devices, Vs, tol, max_iter, max_iter, strict, bernoulli_name, g_frac)
    else:
        J, ok, its = _one_pass(devices, Vs, tol, first_max_iter, max_iter, strict, bernoulli_name,
                               g_frac)
        idx = np.flatnonzero(~ok.all(axis=1))
        if idx.size:
            sub = [devices[k] for k in idx]
            gsub = None if g_frac is None else [g_frac[k] for k in idx]
            J2, ok2, its2 = _one_pass(sub, Vs, tol, max_iter, max_iter, strict, bernoulli_name, gsub)
            J[idx], ok[idx], its[idx] = J2, ok2, its2
    inv = np.argsort(order, kind="stable")
    J, ok, its = J[:, inv], ok[:, inv], its[:, inv]
    shunt = np.array([1.0 / d.R_shunt if d.R_shunt > 0 else 0.0 for d in devices])
    return J + shunt[:, None] * V[None, :], ok, its


def solve_jv(dev, voltages, g_frac=None, tol=1e-10, max_iter=200, bernoulli_name="exact"):
    """`solve_jv` (solver.py:2041-2060) on the CPU: (J with the shunt, converged, iterations).
    Raises RuntimeError where the pre-bias does not converge, as the reference does."""
    V = np.ascontiguousarray(voltages, dtype=np.float64)
    n = len(V)
    J = np.zeros(n)
    ok = np.zeros(n, dtype=np.uint8)
    its = np.zeros(n, dtype=np.int32)
    pin = _Pinned(dev, g_frac)
    rc = lib().ddo_solve_jv(ctypes.byref(pin.struct), _ptr(V), n, tol, max_iter,
                            BERNOULLI[bernoulli_name], _ptr(J),
                            ok.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                            its.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if rc == -2:
        raise RuntimeError(f"no convergence at the pre-bias {dev.V_pre} V")
    if rc != 0:
        raise ValueError(f"ddo_solve_jv: error {rc}")
    return J, ok.astype(bool), its


def bernoulli(x, kind="exact"):
    """(B(x), B'(x)) elementwise."""
    x = np.atleast_1d(np.asarray(x, dtype=float))
    B, dB = np.empty_like(x), np.empty_like(x)
    d = ctypes.c_double()
    for i, v in enumerate(x):
        B[i] = lib().ddo_bernoulli(float(v), BERNOULLI[kind], ctypes.byref(d))
        dB[i] = d.value
    return B, dB


def assemble(dev, full, g_frac=None, la=0.0, lc=0.0, bernoulli_name="exact"):
    """`_assemble` at a full state (NP + 2 x 3): F (NP x 3), A, B, C (NP x 3 x 3), U (NP x 3 x 2)."""
    NP = len(dev.x) - 2
    full = np.ascontiguousarray(full, dtype=np.float64)
    F = np.zeros((NP, 3))
    A, B, C = (np.zeros((NP, 3, 3)) for _ in range(3))
    U = np.zeros((NP, 3, 2))
    pin = _Pinned(dev, g_frac)
    lib().ddo_assemble(ctypes.byref(pin.struct), _ptr(full), pin.struct.g_frac, la, lc,
                       BERNOULLI[bernoulli_name], _ptr(F), _ptr(A), _ptr(B), _ptr(C), _ptr(U))
    return F, A, B, C, U


def assemble_ions(dev, full5, logs, g_frac=None, bernoulli_name="exact"):
    """`_assemble_ions` at dti = 0 (src/jaxdd/solver.py:1364-1473) at a full state (NP + 2 x 5:
    V, ln n, ln p, ln a, ln c) with the slots' logarithms ``logs`` (anion, cation): F (NP x 5),
    A, B, C (NP x 5 x 5), U (NP x 5 x 2)."""
    NP = len(dev.x) - 2
    full5 = np.ascontiguousarray(full5, dtype=np.float64)
    lg = np.ascontiguousarray(logs, dtype=np.float64)
    F = np.zeros((NP, 5))
    A, B, C = (np.zeros((NP, 5, 5)) for _ in range(3))
    U = np.zeros((NP, 5, 2))
    pin = _Pinned(dev, g_frac)
    lib().ddo_assemble_ions(ctypes.byref(pin.struct), _ptr(full5), pin.struct.g_frac, _ptr(lg),
                            BERNOULLI[bernoulli_name], _ptr(F), _ptr(A), _ptr(B), _ptr(C),
                            _ptr(U))
    return F, A, B, C, U


def block_thomas(A, B, C, d):
    """`_block_thomas` for one system: blocks (NP x b x b), right-hand sides (NP x b x k)."""
    NP, b, _ = A.shape
    k = d.shape[-1]
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (A, B, C, d)]
    x = np.zeros((NP, b, k))
    rc = lib().ddo_block_thomas(NP, b, k, *(_ptr(a) for a in arrs), _ptr(x))
    if rc != 0:
        raise ValueError("ddo_block_thomas: bad shape")
    return x


def noise_step(dev, y, Vint, logs=(0.0, 0.0), g_frac=None, bernoulli_name="exact",
               eps=np.finfo(float).eps):
    """`_noise_step` (solver.py:1728-1751) at the interior state ``y`` (NP x 3): the Newton
    step the residual's round-off alone gives, with round-off unit ``eps``."""
    pin = _Pinned(dev, g_frac)
    y = np.ascontiguousarray(y, dtype=np.float64)
    lg = np.ascontiguousarray(logs, dtype=np.float64)
    return lib().ddo_noise_step(ctypes.byref(pin.struct), _ptr(y), _ptr(lg), float(Vint),
                                pin.struct.g_frac, BERNOULLI[bernoulli_name], float(eps))


def transient(dev, t, V, G_frac, tol=1e-10, max_iter=200, bernoulli_name="exact",
              zimt_current=False, failure_mode=1):
    """`solve_transient` for one device (transient.py:270-464) through the steps ``t`` (s),
    internal voltages ``V`` and generation fractions ``G_frac``: a dict of per-step ``J``
    (internal current), ``J_ext`` (with the shunt), ``converged``, ``iterations``, ``kept``."""
    t, V, G = (np.ascontiguousarray(a, dtype=np.float64) for a in (t, V, G_frac))
    n = len(t)
    out = {"J": np.zeros(n), "J_ext": np.zeros(n), "converged": np.zeros(n, np.uint8),
           "iterations": np.zeros(n, np.int32), "kept": np.zeros(n, np.uint8)}
    pin = _Pinned(dev)
    o = _TOpts(tol, max_iter, BERNOULLI[bernoulli_name], int(zimt_current), int(failure_mode))
    rc = lib().ddo_transient(ctypes.byref(pin.struct), _ptr(t), _ptr(V), _ptr(G), n,
                             ctypes.byref(o), _ptr(out["J"]), _ptr(out["J_ext"]),
                             out["converged"].ctypes.data, out["iterations"].ctypes.data,
                             out["kept"].ctypes.data)
    if rc == -5:
        raise NotImplementedError("the transient restatement takes no charged bulk traps")
    if rc != 0:
        raise ValueError(f"ddo_transient: error {rc}")
    out["converged"] = out["converged"].astype(bool)
    out["kept"] = out["kept"].astype(bool)
    return out


def assemble_step(dev, full, g_frac, dti, full_prev, fb, logs=(0.0, 0.0),
                  bernoulli_name="exact", zimt=False, rows=False):
    """One time step's system at ``full`` ((NP + 2) x nc, nc = 3 or 5 with the ions'
    logarithms): F (NP x nc), A, B, C (NP x nc x nc), U (NP x nc x 2), the transient current J
    and the ions' constraints g (2), the step fixed by ``dti``, the state before it
    ``full_prev`` and the traps' occupation ``fb`` (NP; None: the traps at their steady
    occupation, instantaneous). ``rows``: also the constraints' rows W (NP x 2), each the
    derivative in the point's own ion logarithm."""
    full = np.ascontiguousarray(full, dtype=np.float64)
    nc = full.shape[1]
    NP = full.shape[0] - 2
    F = np.zeros((NP, nc))
    A, B, C = (np.zeros((NP, nc, nc)) for _ in range(3))
    U = np.zeros((NP, nc, 2))
    J = np.zeros(1)
    g = np.zeros(2)
    W = np.zeros((NP, 2))
    pin = _Pinned(dev)
    prev = np.ascontiguousarray(full_prev, dtype=np.float64)
    f = None if fb is None else _ptr(np.ascontiguousarray(fb, dtype=np.float64))
    lg = np.ascontiguousarray(logs, dtype=np.float64)
    rc = lib().ddo_assemble_step(ctypes.byref(pin.struct), _ptr(full), nc, float(g_frac),
                                 float(dti), _ptr(prev), f, _ptr(lg),
                                 BERNOULLI[bernoulli_name], int(zimt), _ptr(F), _ptr(A),
                                 _ptr(B), _ptr(C), _ptr(U), _ptr(J), _ptr(g), _ptr(W))
    if rc != 0:
        raise ValueError(f"ddo_assemble_step: error {rc}")
    out = (F, A, B, C, U, float(J[0]), g)
    return out + (W,) if rows else out


def current(dev, full, g_frac=None):
    pin = _Pinned(dev, g_frac)
    full = np.ascontiguousarray(full, dtype=np.float64)
    return lib().ddo_current(ctypes.byref(pin.struct), _ptr(full), pin.struct.g_frac)


def solve_states(dev, voltages, tol=1e-10, max_iter=60, bernoulli_name="exact", g_frac=None):
    """`solve_states` for one device (src/jaxdd/batch.py:1165-1191 with `_observed`
    :1009-1030): the states at ``voltages`` on every point and what photoluminescence measures
    of them, per voltage (the caller's order), as a dict of NumPy arrays: V, n, p, qfls, anion,
    cation (nV x NP + 2), implied_voc, plqy, J (with the shunt), converged, iterations (nV).
    One pass with ``max_iter`` (the two passes' result)."""
    import dataclasses

    V = np.asarray(voltages, dtype=float).ravel()
    order = np.argsort(V, kind="stable")
    J, ok, its, st = solve_one(dev, V[order], tol, max_iter, max_iter, False, bernoulli_name,
                               g_frac, return_states=True)
    inv = np.argsort(order, kind="stable")
    J, ok, its, st = J[inv], ok[inv], its[inv], st[inv]
    nV, NP = len(V), len(dev.x) - 2
    full = np.zeros((nV, NP + 2, 3))
    full[:, 1:-1] = st
    # the contacts (batch.py:255-259 `_bounds`)
    full[:, 0] = np.stack([dev.V0 + 0.5 * dev.left_elec * V,
                           np.full(nV, np.log(dev.n_left)), np.full(nV, np.log(dev.p_left))], -1)
    full[:, -1] = np.stack([dev.VL - 0.5 * dev.left_elec * V,
                            np.full(nV, np.log(dev.n_right)), np.full(nV, np.log(dev.p_right))],
                           -1)
    gf = dev.G_frac if g_frac is None else g_frac
    if dev.pre_cond and dev.fix_ions:
        # the freeze (batch.py:859-872): the free-ion profile of the pre-bias's steady state
        free = dataclasses.replace(dev, pre_cond=False, fix_ions=False)
        pre = solve_states(free, [dev.V_pre], tol, max_iter, bernoulli_name, g_frac)
        anion = np.repeat(pre["anion"], nV, axis=0)
        cation = np.repeat(pre["cation"], nV, axis=0)
    else:
        anion, cation = _ions_on_grid(dev, full[:, :, 0])
    qfls, implied, plqy = observed(dev, full, gf)
    shunt = 1.0 / dev.R_shunt if dev.R_shunt > 0 else 0.0
    return {"V": full[..., 0], "n": np.exp(full[..., 1]), "p": np.exp(full[..., 2]),
            "qfls": qfls, "anion": anion, "cation": cation, "implied_voc": implied,
            "plqy": plqy, "J": J + shunt * V, "converged": ok, "iterations": its}


def observed(dev, full, g_frac):
    """What photoluminescence measures of states ``full`` (nV x NP + 2 x 3): the quasi-Fermi
    level splitting on every point (batch.py:1183-1189), the implied V_oc and the PLQY of
    direct recombination (`_observed`, batch.py:1021-1029: control volumes
    solver.py:1952-1954, the photons against their equilibrium rate and the generation)."""
    n, p = np.exp(full[..., 1]), np.exp(full[..., 2])
    ni2 = dev.ni ** 2
    qfls = dev.Vt * np.log(n * p / ni2)
    vol = 0.5 * dev.L * (dev.h[:-2] + dev.h[1:-1])
    k = dev.k_direct[1:-1]
    rad = np.sum(vol * k * n[:, 1:-1] * p[:, 1:-1], axis=1)
    rad0 = np.sum(vol * k * ni2[1:-1])
    gen = g_frac * np.sum(vol * dev.G[1:-1])
    nV = full.shape[0]
    with np.errstate(divide="ignore", invalid="ignore"):
        implied = dev.Vt * np.log(rad / rad0) if rad0 > 0 else np.full(nV, np.nan)
        plqy = (rad - rad0) / gen if gen > 0 else np.full(nV, np.nan)
    return qfls, implied, plqy


def _ions_on_grid(dev, Vfull):
    """Anions and cations on every point at potentials ``Vfull`` (nV x NP + 2): fixed plus
    Boltzmann where mobile (solver.py:951-961) with each region's normalisation in closed
    form (`_ion_logs` :1695-1712, trapezoidal number :1241-1252), the contacts' points their
    neighbours' (`_ions_on_grid` :2096-2102)."""
    hL = dev.h[:-1] * dev.L
    out = []
    for sign, mask, total, fixed in ((1.0, dev.anion_mask, dev.anion_total, dev.anion_fixed),
                                     (-1.0, dev.cation_mask, dev.cation_total, dev.cation_fixed)):
        m = np.asarray(mask, dtype=float)
        prof = m * np.exp(sign * Vfull / dev.Vt)
        number = np.sum(0.5 * (prof[:, 1:] + prof[:, :-1]) * (m[1:] * m[:-1]) * hL, axis=1)
        logs = np.log(total / number) if total > 0 else np.zeros(len(Vfull))
        dens = fixed + m * np.exp(logs[:, None] + sign * Vfull / dev.Vt)
        dens[:, 0], dens[:, -1] = dens[:, 1], dens[:, -2]
        out.append(dens)
    return out[0], out[1]


__all__ = ["assemble", "assemble_ions", "bernoulli", "block_thomas", "build", "current", "j_scale", "lib",
           "assemble_step", "noise_step", "observed", "transient", "solve_batch", "solve_jv", "solve_one", "solve_states"]

