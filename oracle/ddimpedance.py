"""The impedance of one device at a steady state. TEST INFRASTRUCTURE ONLY: only tests/,
``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use it; the product package
(jaxdd_amd) never does.

A restatement of src/jaxdd/impedance.py:230-360 (`_admittance`, both branches) and
:371-409 (`solve_impedance`'s arithmetic after the steady state). At a steady state a small
internal voltage `dV e^(iwt)` drives `(J + lap M) dy = -(dF/dV_int) dV`:

- `J`: the steady residual's Jacobian, the C oracle's `assemble` (its first row's A and its last
  row's C couple to the contacts, whose potentials move by +-left_elec/2 per volt: the drive);
- `M`: implicit Euler's time terms per unit of `dti`, `-fac n sC` and `-fac p sC` on the
  continuity rows' own logarithms (impedance.py:268-271);
- `lap`: `iw`, or over implicit Euler's steps of `dt` `(1 - e^(-iw dt))/dt` (:398);
- the bulk traps' occupation as a state with its own rate (`_trap_terms` :63-87,
  `_beyond_steady` :90-101, `_step_transfer` :104-117, `_trap_blocks` :214-228): beyond the
  steady (instantaneous) response `df0 = drive . dy / a` it lags by `h = -lap/(lap + a)`
  (continuous) or implicit Euler's exact exponential step over `dt`, in both carriers' rates.

The current is a transient's (src/jaxdd/batch.py:314-403 `_current_parts` with a time step:
each point's control volume as the width of its rate, the densities' change, the displacement
current), linearised as impedance.py:316-358 does, `dJ = J_y dy + J_V dV + lap (J_y(1) - J_y(0))
dy` plus the traps' rates beyond the steady ones; here every derivative is the complex step of
the restated current (Im J(y + i e v)/e, exact to rounding for a real-analytic J), not a
derivation of its own. The linear systems are solved by the reference's block elimination
(`_block_thomas`, each block inverted by its adjugate) in complex arithmetic.

With mobile ions (impedance.py:242-262) the ions' logarithms are unknowns of their own (5 x 5
blocks, the C oracle's `assemble_step` at `dti` = 1 from the steady state itself, the traps at
their steady occupation): `M` gains each ion row's time term per unit of `dti` (its row scale
included), the step's own `dti` = 1 is taken back out of `B` (`lap - 1`), and each species'
number is kept by a slot (`U`, the row's derivative in the slot's log) with its trapezoidal number
as a bordered constraint (`W`), closed by the 2 x 2 Schur complement. The current adds each ion
region's current (batch.py:421-485 `_ion_currents`) to the carriers' and the displacement
current. One region per species, not reaching a contact (the C oracle's subset).
"""

from __future__ import annotations

import numpy as np

from jaxdd_amd.constants import Q
from oracle.ddoracle import assemble, assemble_step

_EPS = 1e-30  # the complex step, relative to the direction's largest component


def step_transfer(a, lap, dt=None):
    """`(df - df0)/df0` of a trap level of rate `a` at `lap`, in the rates and in the charge
    (impedance.py:104-117): continuous both `-lap/(lap + a)`; over implicit Euler's steps of
    `dt` the mean over a step and its end, `E = e^(-a dt)`, `m = (1 - E)/(a dt)`."""
    a = np.asarray(a, dtype=float)
    if dt is None:
        h = -lap / (lap + a)
        return h, h
    x = a * dt
    E, m = np.exp(-x), -np.expm1(-x) / x
    den = 1.0 - E * (1.0 - lap * dt)
    return -m * lap * dt / den, -E * lap * dt / den


def _interior(dev):
    """Per interior point: `fac` and the continuity rows' scale `sC` (the C oracle's
    `problem_init`), the trap fields."""
    NP = len(dev.x) - 2
    h, L = np.asarray(dev.h, float), dev.L
    i = np.arange(1, NP + 1)
    fac = 0.5 * L * L * h[i] * h[i - 1] * (h[i] + h[i - 1])
    mu = np.maximum(np.asarray(dev.mu_n, float)[i], np.asarray(dev.mu_p, float)[i])
    sC = 1.0 / (mu * dev.Vt * np.asarray(dev.N_c, float)[i] * h[i])
    traps = {k: np.asarray(getattr(dev, "srh_" + k), float)[i] for k in ("Nt", "Cn", "Cp", "nt0", "pt0")}
    return fac, sC, traps


def trap_terms(dev, n, p):
    """impedance.py:63-87 for one level per point (interior n, p): the rate `a` (1 where no
    traps), the drive per unit of ln n, ln p over `a` (`unit`, NP x 2), and `dR` (2 x NP): the
    occupation's part in the electrons' and the holes' rates."""
    _, _, t = _interior(dev)
    Nt, Cn, Cp, nt0, pt0 = t["Nt"], t["Cn"], t["Cp"], t["nt0"], t["pt0"]
    has = Nt > 0
    a = np.where(has, Cn * (n + nt0) + Cp * (p + pt0), 1.0)
    b = Cn * n + Cp * pt0
    f = b / a
    unit = np.stack([Cn * n * (1.0 - f), -f * Cp * p], -1) / a[:, None]
    dR = np.where(has, np.stack([-Nt * Cn * (n + nt0), Nt * Cp * (p + pt0)]), 0.0)
    return a, np.where(has[:, None], unit, 0.0), dR


def ion_regions(dev):
    """Per species (anions, cations): its mask on the full grid, its mobility on the points
    (1 where none), its total, and its region's first and last point (-1 without ions).
    `ValueError` for several regions of a species or one reaching a contact."""
    NP = len(dev.x) - 2
    out = []
    for mask, mu, total in ((dev.anion_mask, dev.mu_anion, dev.anion_total),
                            (dev.cation_mask, dev.mu_cation, dev.cation_total)):
        m = np.asarray(mask, dtype=float)
        mu = np.zeros(NP + 2) if mu is None else np.asarray(mu, dtype=float)
        idx = np.flatnonzero(m > 0)
        first = last = -1
        if total > 0 and idx.size:
            first, last = int(idx[0]), int(idx[-1])
            if idx.size != last - first + 1:
                raise ValueError("several regions of one species")
            if first == 0 or last == NP + 1:
                raise ValueError("an ion region reaching a contact")
        out.append({"mask": m, "mu": np.where(mu > 0, mu, 1.0), "mu_raw": mu,
                    "total": float(total), "first": first, "last": last})
    return out


def _ion_currents(dev, full, dti, before, width, regions):
    """Each ion region's current on the intervals (NP + 1) over a time step (batch.py:421-485
    `_ion_currents` for one region per species, no sites, no `curr_diff`): the integral form
    with the densities' change as its integrand; 0 outside the region."""
    NP = len(dev.x) - 2
    x = np.asarray(dev.x, float)
    dx = np.diff(x)
    Vt, L = dev.Vt, dev.L
    V = full[:, 0]
    iface = np.zeros(NP + 1, bool)
    for k in dev.i1[:-1]:
        iface[k] = True
    out = np.zeros(NP + 1, dtype=np.result_type(full, float))
    for k, reg in enumerate(regions):
        if reg["first"] < 0:
            continue
        sn = -1.0 if k == 0 else 1.0
        m, mu = reg["mask"], reg["mu"]
        dens = np.exp(full[:, 3 + k])
        U = dti * (dens - before[k])
        ev = np.exp(sn * V / Vt)
        pair = m[1:] * m[:-1]
        int_u = np.concatenate([[0.0], np.cumsum(0.5 * dx * (U[:-1] + U[1:]) * pair)])
        single = np.sum(0.5 * dx * (ev[:-1] / mu[:-1] + ev[1:] / mu[1:]) * pair)
        double = np.sum(0.5 * dx * (ev[:-1] * int_u[:-1] / mu[:-1] + ev[1:] * int_u[1:] / mu[1:]) * pair)
        first, last = reg["first"], reg["last"]
        ends = sn * (dens[last] * ev[last] - dens[first] * ev[first])
        K = (Vt * ends - sn * double) / single
        e = np.arange(NP + 1)
        step = np.where(e > first, sn * Q * L * width * U[: NP + 1], 0.0) * pair
        J = Q * K + np.cumsum(step)
        J = np.where(iface, np.roll(J, 1), J)
        out = out + np.where(pair > 0, J, 0.0)
    return out


def transient_current(dev, full, g_frac, dti, before, dU=None, width=None, ions=None):
    """The device's current over a time step (batch.py:300-311, 314-403 with `tr`), without
    interface traps or field-dependent rates: `full` (NP + 2 x 3, or x 5 with the ions'
    logarithms; may be complex), `before` = (n, p, V) of the state before the step (NP + 2
    each), `dU` = each carrier's rate beyond the steady traps' (two NP + 2 arrays), `width` the
    rates' widths (default: each point's control volume, batch.py:406-419; `dev.h` gives the
    steady state's). ``ions``: (regions, (a, c) densities before the step) adds the mobile
    ions' currents."""
    NP = len(dev.x) - 2
    x, h = np.asarray(dev.x, float), np.asarray(dev.h, float)
    Vt, L = dev.Vt, dev.L
    hh = h[: NP + 1]
    if width is None:
        width = np.concatenate([hh[:1], 0.5 * (hh[:-1] + hh[1:])])
    V, n, p = full[:, 0], np.exp(full[:, 1]), np.exp(full[:, 2])
    kd, ni2 = np.asarray(dev.k_direct, float), np.asarray(dev.ni, float) ** 2
    Nt, Cn, Cp = (np.asarray(getattr(dev, "srh_" + k), float) for k in ("Nt", "Cn", "Cp"))
    nt0, pt0 = np.asarray(dev.srh_nt0, float), np.asarray(dev.srh_pt0, float)
    den = np.where(Nt > 0, Cn * (n + nt0) + Cp * (p + pt0), 1.0)
    R = kd * (n * p - ni2) + np.where(Nt > 0, Nt * Cn * Cp * (n * p - nt0 * pt0) / den, 0.0)
    G = np.asarray(dev.G, float) * g_frac
    dx = np.diff(x[: NP + 1])
    iface = np.zeros(NP + 1, bool)
    for k in dev.i1[:-1]:
        iface[k] = True
    sign = -float(dev.left_elec)

    def carrier(sn, off, dens, mu, before_d, dUc):
        U = R - G + (0.0 if dUc is None else dUc)
        U = U * np.r_[0.0, np.ones(NP), 0.0]
        change = dti * (dens - before_d)
        change = change * np.r_[0.0, np.ones(NP + 1)]
        U = U + change
        ev = np.exp(sn * (V[: NP + 1] + off[: NP + 1]) / Vt)
        int_u = np.concatenate([[0.0], np.cumsum(0.5 * dx * (U[:NP] + U[1 : NP + 1]))])
        single = np.sum(0.5 * dx * (ev[:-1] / mu[:NP] + ev[1:] / mu[1 : NP + 1]))
        double = np.sum(0.5 * dx * (ev[:-1] * int_u[:-1] / mu[:NP] + ev[1:] * int_u[1:] / mu[1 : NP + 1]))
        K = (Vt * (sn * dens[NP + 1] * np.exp(sn * (V[NP + 1] + off[NP + 1]) / Vt)
                   - sn * dens[0] * np.exp(sn * (V[0] + off[0]) / Vt)) - sn * double) / single
        J = Q * K + np.concatenate([[0.0], np.cumsum(sn * Q * L * width[1 : NP + 1] * U[1 : NP + 1])])
        return np.where(iface, np.roll(J, 1), J)

    bn, bp, bV = before
    dUn, dUp = (None, None) if dU is None else dU
    total = sign * carrier(-1.0, np.asarray(dev.offset_n, float), n, np.asarray(dev.mu_n, float), bn, dUn)
    total = total + sign * carrier(1.0, np.asarray(dev.offset_p, float), p, np.asarray(dev.mu_p, float), bp, dUp)
    if ions is not None:
        regions, before_ions = ions
        total = total + sign * _ion_currents(dev, full, dti, before_ions, width, regions)
    eps = np.asarray(dev.eps, float)
    total = total + sign * eps[: NP + 1] * (np.diff(V) - np.diff(bV)) * dti / (L * h[: NP + 1])
    return np.sum(hh * total) / np.sum(hh)


def _cstep(f, scale):
    """The derivative of the real-analytic `f(t)` at t = 0 by the complex step."""
    e = _EPS / max(scale, 1e-300)
    return f(1j * e).imag / e


def _inverse3(m):
    """3x3 inverses by the adjugate (src/jaxdd/solver.py:1481-1502 `_inverse3`), the reference's
    block inverse, for complex blocks as well."""
    a, b, c = m[..., 0, 0], m[..., 0, 1], m[..., 0, 2]
    d, e, f = m[..., 1, 0], m[..., 1, 1], m[..., 1, 2]
    g, h, i = m[..., 2, 0], m[..., 2, 1], m[..., 2, 2]
    co = np.stack([
        np.stack([e * i - f * h, c * h - b * i, b * f - c * e], -1),
        np.stack([f * g - d * i, a * i - c * g, c * d - a * f], -1),
        np.stack([d * h - e * g, b * g - a * h, a * e - b * d], -1),
    ], -2)
    det = a * co[..., 0, 0] + b * co[..., 1, 0] + c * co[..., 2, 0]
    return co / det[..., None, None]


def _thomas(A, B, C, d, pivoted=False):
    """`_block_thomas` (src/jaxdd/solver.py:1534-1560) for a batch of systems: A, B, C
    (..., NP, 3, 3), d (..., NP, 3); each block inverted by `_inverse3`, or (``pivoted``) by
    NumPy's pivoted solves: the same elimination in exact arithmetic, whose difference
    measures how far the problem's conditioning lets rounding move the result."""
    NP = A.shape[-3]
    cp = np.empty(np.broadcast_shapes(A.shape, B.shape), dtype=complex)
    dp = np.empty(np.broadcast_shapes(d.shape, B.shape[:-1]), dtype=complex)
    for i in range(NP):
        Bi, di = B[..., i, :, :], d[..., i, :]
        if i:
            Bi = Bi - A[..., i, :, :] @ cp[..., i - 1, :, :]
            di = di - (A[..., i, :, :] @ dp[..., i - 1, :, None])[..., 0]
        if pivoted:
            cp[..., i, :, :] = np.linalg.solve(Bi, np.broadcast_to(C[..., i, :, :], Bi.shape))
            dp[..., i, :] = np.linalg.solve(Bi, di[..., None])[..., 0]
            continue
        m_inv = _inverse3(Bi)
        cp[..., i, :, :] = m_inv @ C[..., i, :, :]
        dp[..., i, :] = (m_inv @ di[..., None])[..., 0]
    x = np.empty_like(dp)
    x[..., NP - 1, :] = dp[..., NP - 1, :]
    for i in range(NP - 2, -1, -1):
        x[..., i, :] = dp[..., i, :] - (cp[..., i, :, :] @ x[..., i + 1, :, None])[..., 0]
    return x


def _inverse2(m):
    """2x2 inverses by the adjugate (solver.py:1498-1502 `_inverse2`)."""
    a, b, c, d = m[..., 0, 0], m[..., 0, 1], m[..., 1, 0], m[..., 1, 1]
    inv = 1.0 / (a * d - b * c)
    return np.stack([np.stack([d, -b], -1), np.stack([-c, a], -1)], -2) * inv[..., None, None]


def _inverse5(m):
    """5x5 inverses by their partition (solver.py:1505-1517 `_inverse5`): the carriers' 3x3 block
    by its adjugate, the ions' 2x2 Schur complement by its own."""
    P, Qm, R, S = m[..., :3, :3], m[..., :3, 3:], m[..., 3:, :3], m[..., 3:, 3:]
    Pi = _inverse3(P)
    PiQ, RPi = Pi @ Qm, R @ Pi
    Si = _inverse2(S - R @ PiQ)
    top = np.concatenate([Pi + PiQ @ (Si @ RPi), -(PiQ @ Si)], -1)
    bottom = np.concatenate([-(Si @ RPi), Si], -1)
    return np.concatenate([top, bottom], -2)


def _thomas_cols(A, B, C, D, pivoted=False):
    """`_block_thomas` (solver.py:1534-1560) for 5x5 blocks and several right-hand sides: A, B,
    C (..., NP, 5, 5), D (..., NP, 5, k); each block inverted by `_inverse5`, or (``pivoted``) by
    NumPy's pivoted solves."""
    NP = A.shape[-3]
    shape_c = np.broadcast_shapes(A.shape, B.shape)
    cp = np.empty(shape_c, dtype=complex)
    dp = np.empty(np.broadcast_shapes(D.shape[:-3], B.shape[:-3]) + (NP,) + D.shape[-2:], dtype=complex)
    for i in range(NP):
        Bi, di = B[..., i, :, :], D[..., i, :, :]
        if i:
            Bi = Bi - A[..., i, :, :] @ cp[..., i - 1, :, :]
            di = di - A[..., i, :, :] @ dp[..., i - 1, :, :]
        if pivoted:
            cp[..., i, :, :] = np.linalg.solve(Bi, np.broadcast_to(C[..., i, :, :], Bi.shape))
            dp[..., i, :, :] = np.linalg.solve(Bi, np.broadcast_to(di, Bi.shape[:-1] + di.shape[-1:]))
            continue
        m_inv = _inverse5(Bi)
        cp[..., i, :, :] = m_inv @ C[..., i, :, :]
        dp[..., i, :, :] = m_inv @ di
    x = np.empty_like(dp)
    x[..., NP - 1, :, :] = dp[..., NP - 1, :, :]
    for i in range(NP - 2, -1, -1):
        x[..., i, :, :] = dp[..., i, :, :] - cp[..., i, :, :] @ x[..., i + 1, :, :]
    return x


def ion_state(dev, full, anion, cation):
    """The 5-column full state of the ion branch (impedance.py:248-252): `full` (NP + 2 x 3) and
    the ions' logarithms, each species' density on the grid (``anion``, ``cation``, NP + 2)
    where it is mobile, 0 elsewhere (the contacts' too: no region reaches one)."""
    NP = full.shape[0] - 2
    out = np.zeros((NP + 2, 5))
    out[:, :3] = full
    for k, (mask, dens) in enumerate(((dev.anion_mask, anion), (dev.cation_mask, cation))):
        m = np.asarray(mask, float)[1:-1] > 0
        out[1:-1, 3 + k] = np.where(m, np.log(np.where(m, dens[1:-1], 1.0)), 0.0)
    return out


def ion_masses(dev, full5):
    """`_assemble_ions`' time terms per unit of `dti` (solver.py:1467-1472, 1436-1441): NP x 5,
    `-fac n sC`, `-fac p sC` on the continuity rows and `-fac a scale` on each ion's row in its
    region, `scale = 1/(a_prev (fac dti + h_(i-1) V_T mu_i + h_i V_T mu_(i-1)))` at `dti` = 1
    with the state before the step the steady one; 0 in Poisson's row."""
    NP = full5.shape[0] - 2
    fac, sC, _ = _interior(dev)
    h = np.asarray(dev.h, float)
    i = np.arange(1, NP + 1)
    M = np.zeros((NP, 5))
    M[:, 1] = -fac * np.exp(full5[1:-1, 1]) * sC
    M[:, 2] = -fac * np.exp(full5[1:-1, 2]) * sC
    for k, reg in enumerate(ion_regions(dev)):
        if reg["first"] < 0:
            continue
        mu = reg["mu_raw"]
        dens = np.exp(full5[1:-1, 3 + k])
        scale = 1.0 / (dens * (fac + h[i - 1] * dev.Vt * mu[i] + h[i] * dev.Vt * mu[i - 1]))
        M[:, 3 + k] = np.where(reg["mask"][i] > 0, -fac * dens * scale, 0.0)
    return M


def admittance_ions(dev, full5, laps, g_frac=None, dt=None, bernoulli_name="exact",
                    response=False, pivoted=False):
    """`dJ/dV_int` (complex, per `lap`) of one device with its ions' logarithms as unknowns
    (impedance.py:242-262 and the bordered solve :301-308): ``full5`` (NP + 2 x 5, `ion_state`)
    the steady state. Also for a device without ions (its ion rows held, its slots identities):
    the 3 x 3 branch's result. ``response``: also `dy` (frequencies x NP x 5)."""
    laps = np.atleast_1d(np.asarray(laps, dtype=complex))
    g = dev.G_frac if g_frac is None else g_frac
    full5 = np.asarray(full5, dtype=float)
    NP = full5.shape[0] - 2
    regions = ion_regions(dev)
    F, A, B, C, U, _, _, W = assemble_step(dev, full5, g, 1.0, full5, None,
                                           bernoulli_name=bernoulli_name, rows=True)
    M = ion_masses(dev, full5)
    half = 0.5 * float(dev.left_elec)
    dF = np.zeros((NP, 5))
    dF[0] += A[0][:, 0] * half
    dF[-1] += C[-1][:, 0] * (-half)
    C = C.copy()
    C[-1] = 0.0
    fac, sC, _ = _interior(dev)
    n, p = np.exp(full5[1:-1, 1]), np.exp(full5[1:-1, 2])
    a, unit, dR = trap_terms(dev, n, p)
    fs = -fac * sC
    h_rate, _ = step_transfer(a[None, :], laps[:, None], dt)
    nf = len(laps)
    Bl = np.broadcast_to(B, (nf,) + B.shape).astype(complex).copy()
    for c in range(5):
        Bl[..., c, c] += (laps[:, None] - 1.0) * M[:, c]
    for c in (0, 1):
        for k in (0, 1):
            Bl[..., 1 + c, 1 + k] += fs * dR[c] * unit[:, k] * h_rate
    rhs = np.concatenate([-dF[:, :, None], U], -1)  # NP x 5 x 3
    X = _thomas_cols(A.astype(complex), Bl, C.astype(complex),
                     np.broadcast_to(rhs, (nf,) + rhs.shape).astype(complex), pivoted)
    XF, XU = X[..., 0], X[..., 1:]
    held = np.array([0.0 if r["first"] >= 0 else 1.0 for r in regions])
    Wf = np.zeros((NP, 5, 2))
    Wf[:, 3, 0], Wf[:, 4, 1] = W[:, 0], W[:, 1]
    schur = np.diag(held)[None] - np.einsum("irk,firl->fkl", Wf, XU)
    z = np.linalg.solve(schur, (-np.einsum("irk,fir->fk", Wf, XF))[..., None])[..., 0]
    dy = XF - np.einsum("firl,fl->fir", XU, z)

    before = (np.exp(full5[:, 1]), np.exp(full5[:, 2]), full5[:, 0].copy())
    ions_before = tuple(np.exp(full5[:, 3 + k]) * r["mask"] for k, r in enumerate(regions))
    ions = (regions, ions_before)

    def linear(dy_real, dV, dti):
        v = np.zeros_like(full5)
        v[1:-1] = dy_real
        v[0, 0], v[-1, 0] = half * dV, -half * dV
        if not np.any(v):
            return 0.0
        scale = float(np.max(np.abs(v)))
        return _cstep(lambda t: transient_current(dev, full5 + t * v, g, dti, before, ions=ions),
                      scale)

    def rates(dUn, dUp):
        scale = float(max(np.max(np.abs(dUn)), np.max(np.abs(dUp)), 1e-300))
        return _cstep(lambda t: transient_current(dev, full5 + 0j, g, 0.0, before,
                                                  dU=(t * dUn, t * dUp), ions=ions), scale)

    out = np.empty(nf, dtype=complex)
    for k, lap in enumerate(laps):
        d = dy[k]
        J0 = linear(d.real, 1.0, 0.0) + 1j * linear(d.imag, 0.0, 0.0)
        J1 = linear(d.real, 1.0, 1.0) + 1j * linear(d.imag, 0.0, 1.0)
        df = unit[:, 0] * h_rate[k] * d[:, 1] + unit[:, 1] * h_rate[k] * d[:, 2]
        pad = lambda v: np.concatenate([[0.0], v, [0.0]])  # noqa: E731
        dUn, dUp = pad(dR[0] * df), pad(dR[1] * df)
        Ju = rates(dUn.real, dUp.real) + 1j * rates(dUn.imag, dUp.imag)
        out[k] = J0 + lap * (J1 - J0) + Ju
    return (out, dy) if response else out


def admittance(dev, full, laps, g_frac=None, dt=None, bernoulli_name="exact", response=False,
               pivoted=False):
    """`dJ/dV_int` (complex, per `lap` of `laps`) of one device without mobile ions at its
    steady state `full` (NP + 2 x 3: V, ln n, ln p; the contacts at its internal voltage).
    ``dt``: implicit Euler's steps (`laps` theirs, the traps' `step_transfer` over them).
    ``response``: also `dy` (frequencies x NP x 3). ``pivoted``: `_thomas`'s."""
    laps = np.atleast_1d(np.asarray(laps, dtype=complex))
    g = dev.G_frac if g_frac is None else g_frac
    full = np.asarray(full, dtype=float)
    NP = full.shape[0] - 2
    _, A, B, C, _ = assemble(dev, full, g, bernoulli_name=bernoulli_name)
    fac, sC, _ = _interior(dev)
    half = 0.5 * float(dev.left_elec)
    dF = np.zeros((NP, 3))
    dF[0] += A[0][:, 0] * half
    dF[-1] += C[-1][:, 0] * (-half)
    C = C.copy()
    C[-1] = 0.0  # the last row's coupling is the contact's, not an unknown's
    n, p = np.exp(full[1:-1, 1]), np.exp(full[1:-1, 2])
    a, unit, dR = trap_terms(dev, n, p)
    fs = -fac * sC
    h_rate, _ = step_transfer(a[None, :], laps[:, None], dt)  # freq x NP
    Bl = np.broadcast_to(B, (len(laps),) + B.shape).astype(complex)
    Bl[..., 1, 1] += laps[:, None] * fs * n
    Bl[..., 2, 2] += laps[:, None] * fs * p
    for c in (0, 1):
        for k in (0, 1):
            Bl[..., 1 + c, 1 + k] += fs * dR[c] * unit[:, k] * h_rate
    dy = _thomas(A.astype(complex), Bl, C.astype(complex), np.broadcast_to(-dF, (len(laps), NP, 3)),
                 pivoted)

    before = (np.exp(full[:, 1]), np.exp(full[:, 2]), full[:, 0].copy())

    def linear(dy_real, dV, dti):
        v = np.zeros_like(full)
        v[1:-1] = dy_real
        v[0, 0], v[-1, 0] = half * dV, -half * dV
        scale = float(np.max(np.abs(v))) if np.any(v) else 1.0
        if not np.any(v):
            return 0.0
        return _cstep(lambda t: transient_current(dev, full + t * v, g, dti, before), scale)

    def rates(dUn, dUp):
        scale = float(max(np.max(np.abs(dUn)), np.max(np.abs(dUp)), 1e-300))
        return _cstep(lambda t: transient_current(dev, full + 0j, g, 0.0, before,
                                                  dU=(t * dUn, t * dUp)), scale)

    out = np.empty(len(laps), dtype=complex)
    for k, lap in enumerate(laps):
        d = dy[k]
        J0 = linear(d.real, 1.0, 0.0) + 1j * linear(d.imag, 0.0, 0.0)
        J1 = linear(d.real, 1.0, 1.0) + 1j * linear(d.imag, 0.0, 1.0)
        df = unit[:, 0] * h_rate[k] * d[:, 1] + unit[:, 1] * h_rate[k] * d[:, 2]
        pad = lambda v: np.concatenate([[0.0], v, [0.0]])  # noqa: E731
        dUn, dUp = pad(dR[0] * df), pad(dR[1] * df)
        Ju = rates(dUn.real, dUp.real) + 1j * rates(dUn.imag, dUp.imag)
        out[k] = J0 + lap * (J1 - J0) + Ju
    return (out, dy) if response else out


def laps_of(frequencies, dt=None):
    """`iw`, or implicit Euler's `(1 - e^(-iw dt))/dt` (impedance.py:397-398)."""
    w = 2 * np.pi * np.asarray(frequencies, dtype=float)
    return 1j * w if dt is None else -np.expm1(-1j * w * dt) / dt


def full_state(dev, y, Vint):
    """The interior state `y` (NP x 3) with the contacts at `Vint` (src/jaxdd/batch.py:255-267)."""
    half = 0.5 * float(dev.left_elec) * Vint
    left = [dev.V0 + half, np.log(dev.n_left), np.log(dev.p_left)]
    right = [dev.VL - half, np.log(dev.n_right), np.log(dev.p_right)]
    return np.concatenate([[left], np.asarray(y, float), [right]])


def geometric_capacitance(dev):
    """The layers' dielectrics in series, `1/integral dx/eps` (the trapezoid over the grid)."""
    x, eps = np.asarray(dev.x, float), np.asarray(dev.eps, float)
    return 1.0 / np.sum(np.diff(x) * 0.5 * (1 / eps[:-1] + 1 / eps[1:]))


__all__ = ["admittance", "admittance_ions", "full_state", "geometric_capacitance", "ion_masses",
           "ion_regions", "ion_state", "laps_of", "step_transfer", "trap_terms", "transient_current"]
