"""Independent NumPy statements used as checks by the tests (not product code)."""

from __future__ import annotations

import numpy as np

from jaxdd_amd.constants import Q


def initial_full(dev, Vint):
    """The linear start on all NP + 2 points (src/jaxdd/solver.py:1882-1893)."""
    e = dev.left_elec
    x, L = dev.x, dev.L
    V = dev.V0 + 0.5 * e * Vint + x * (dev.VL - dev.V0 - e * Vint) / L
    EF = dev.W_L - x * (dev.W_L - dev.W_R) / L
    n = dev.N_c * np.exp(-(EF - dev.E_c) / dev.Vt)
    p = dev.ni ** 2 / n
    full = np.stack([V, np.log(n), np.log(p)], -1)
    full[0] = [dev.V0 + 0.5 * e * Vint, np.log(dev.n_left), np.log(dev.p_left)]
    full[-1] = [dev.VL - 0.5 * e * Vint, np.log(dev.n_right), np.log(dev.p_right)]
    return full


def _B(x):
    x = np.asarray(x, dtype=float)
    small = np.abs(x) < 1e-2
    safe = np.where(small, 1.0, x)
    return np.where(small, 1 - x / 2 + x * x / 12 - x ** 4 / 720, safe / np.expm1(safe))


def residual_by_rows(dev, full, la, lc):
    """Each interior row's residual from its three points (the reference's `_row`,
    src/jaxdd/solver.py:964-1006): Poisson with the permittivity jump, and each carrier's
    continuity with the Scharfetter-Gummel coefficients of its two intervals."""
    NP = len(dev.x) - 2
    i = np.arange(1, NP + 1)
    h, L, eps, Vt = dev.h, dev.L, dev.eps, dev.Vt
    hm, hp = h[i - 1], h[i]
    fac_m = (eps[i] - eps[i - 1]) / (hm ** 2 * L ** 2) - 2 * eps[i] * (
        1 / (hp * (hp + hm) * L ** 2) + 1 / (hm * (hp + hm) * L ** 2))
    fac_l = 2 * eps[i] / (hm * (hp + hm) * L ** 2) - (eps[i] - eps[i - 1]) / (hm ** 2 * L ** 2)
    fac_u = 2 * eps[i] / (hp * (hp + hm) * L ** 2)
    fac = 0.5 * L ** 2 * hp * hm * (hp + hm)
    V = full[:, 0]
    n, p = np.exp(full[:, 1]), np.exp(full[:, 2])
    anion = dev.anion_fixed[i] + dev.anion_mask[i] * np.exp(la + V[i] / Vt)
    cation = dev.cation_fixed[i] + dev.cation_mask[i] * np.exp(lc - V[i] / Vt)
    poisson = fac_l * V[i - 1] + fac_m * V[i] + fac_u * V[i + 1] - Q * (
        n[i] - p[i] + dev.N_A[i] - dev.N_D[i] + anion - cation)
    gn, gp = (V + dev.offset_n) / Vt, (V + dev.offset_p) / Vt
    mun_m, mun_p, mup_m, mup_p = dev.mu_n[i - 1], dev.mu_n[i], dev.mu_p[i - 1], dev.mu_p[i]
    lo_n = hp * mun_m * Vt * _B(gn[i - 1] - gn[i])
    up_n = hm * mun_p * Vt * _B(gn[i + 1] - gn[i])
    m_n = -(hm * mun_p * Vt * _B(gn[i] - gn[i + 1]) + hp * mun_m * Vt * _B(gn[i] - gn[i - 1]))
    lo_p = hp * mup_m * Vt * _B(gp[i] - gp[i - 1])
    up_p = hm * mup_p * Vt * _B(gp[i] - gp[i + 1])
    m_p = -(hm * mup_p * Vt * _B(gp[i + 1] - gp[i]) + hp * mup_m * Vt * _B(gp[i - 1] - gp[i]))
    nn, pp = n[i], p[i]
    R = dev.k_direct[i] * (nn * pp - dev.ni[i] ** 2)
    Nt = dev.srh_Nt[i]
    den = dev.srh_Cn[i] * (nn + dev.srh_nt0[i]) + dev.srh_Cp[i] * (pp + dev.srh_pt0[i])
    srh = Nt * dev.srh_Cn[i] * dev.srh_Cp[i] * (nn * pp - dev.srh_nt0[i] * dev.srh_pt0[i])
    R = R + np.where(Nt > 0, srh / np.where(Nt > 0, den, 1.0), 0.0)
    gen = dev.G[i] * dev.G_frac
    elec = lo_n * n[i - 1] + m_n * nn + up_n * n[i + 1] + fac * (gen - R)
    hole = lo_p * p[i - 1] + m_p * pp + up_p * p[i + 1] + fac * (gen - R)
    sP = 1 / (Q * dev.N_c[i])
    sC = 1 / (np.maximum(dev.mu_n[i], dev.mu_p[i]) * Vt * dev.N_c[i] * hp)
    return np.stack([poisson * sP, elec * sC, hole * sC], -1)
