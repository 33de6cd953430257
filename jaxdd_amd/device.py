"""One device on its grid: the static arrays the steady-state solve reads (host side, NumPy).

Reference interface mirrored: src/jaxdd/device.py (`make_grid` :236-248, `_graded_layout`
:160-214, `Device` :251-382, `Structure` :385-438, `device_parameters` :576-801,
`device_arrays` :877-1063, `build_device` :1066-1079), restricted to the physics of the
batched J-V path (Boltzmann carriers, direct and neutral bulk SRH recombination, generation
from a file or uniform per layer, one mobile ionic region per species, ohmic contacts).

A device splits into its structure (the grid's points per layer, which layers generate, hold
traps or ions, the pre-bias sequence) and its parameters (the numbers). `device_arrays` makes
the arrays from both; the synthetic batch generator calls it with parameters it draws.
"""

from __future__ import annotations

import io
from collections.abc import Sequence
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from .constants import EPS_0, K_B, Q
from .simss_input import Setup, ion_region_problems, ion_regions

__all__ = [
    "GRIDS",
    "GRID_SPLITS",
    "Device",
    "Structure",
    "build_device",
    "device_arrays",
    "device_parameters",
    "layer_points",
    "make_grid",
]

#: SIMsalabim's `MinGridPointsPerLayer`.
MIN_GRID_POINTS_PER_LAYER = 5
#: How `layer_points` evaluates the share of the points: in exact rational arithmetic from the
#: input's decimals, or as SimSS 5.36's 80-bit extended reals round it.
GRID_SPLITS = ("exact", "simss")
#: The grids `build_device` takes: graded towards every layer's ends (the default), or
#: SIMsalabim's `Make_Grid`.
GRIDS = ("simss", "graded")
#: The graded grid's points per e-fold of the distance to a layer's end, and the distance (m)
#: where the grading starts (reference: device.py:139-141).
GRADED_POINTS_PER_EFOLD = 6.0
GRADED_SMALLEST = 3e-12
#: Bisection steps placing the graded grid's points (the integral is monotone).
_BISECTION_STEPS = 100
#: Significand bits of the x87 extended type SimSS computes in.
_EXTENDED_BITS = 64


def _to_extended(q: Fraction) -> Fraction:
    """``q`` rounded to the nearest value with a 64-bit significand, ties to even."""
    if q == 0:
        return q
    negative = q < 0
    q = -q if negative else q
    # exponent e with 2^e <= q < 2^(e+1)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    if q < Fraction(2) ** e:
        e -= 1
    unit = Fraction(2) ** (e - _EXTENDED_BITS + 1)
    scaled = q / unit
    whole, rest = divmod(scaled.numerator, scaled.denominator)
    rest = Fraction(rest, scaled.denominator)
    if rest > Fraction(1, 2) or (rest == Fraction(1, 2) and whole % 2 == 1):
        whole += 1
    value = whole * unit
    return -value if negative else value


def layer_points(thicknesses: Sequence[str | float], NP: int, split: str = "exact") -> list[int]:
    """Points `Make_Grid` gives each layer beyond its first (reference: device.py:73-100):
    ``max(5, FLOOR((NP - n_thin*5) * L/(Ltot - L_thin)))``, a layer thin where ``NP*L/Ltot``
    is below 5. ``split`` "exact" evaluates the share in rationals from the decimal values,
    "simss" rounds each operation to an extended real, as SimSS does."""
    if split not in GRID_SPLITS:
        raise ValueError(f"split {split!r}: one of {GRID_SPLITS}")
    r = (lambda q: q) if split == "exact" else _to_extended
    values = [r(Fraction(t if isinstance(t, str) else repr(float(t)))) for t in thicknesses]
    total = Fraction(0)
    for v in values:
        total = r(total + v)
    thin = [v for v in values if r(r(NP * v) / total) < MIN_GRID_POINTS_PER_LAYER]
    thin_total = Fraction(0)
    for v in thin:
        thin_total = r(thin_total + v)
    free = NP - len(thin) * MIN_GRID_POINTS_PER_LAYER
    rest = r(total - thin_total)
    counts = []
    for v in values:
        share = r(free * r(v / rest))
        counts.append(max(MIN_GRID_POINTS_PER_LAYER, share.numerator // share.denominator))
    return counts


def _sub_grid(m: int, grad: float) -> np.ndarray:
    """`Make_Sub_Grid` before scaling: ``m + 1`` exponentially graded spacings, symmetric."""
    h = np.ones(m + 1)
    half = round(0.5 * m)
    for i in range(half + 1):
        h[i] = np.exp(grad * (1.0 / half - 1.0 / (i + 1)))
        h[m - i] = h[i]
    return h


def _assemble_layout(counts, NP, spacing_of):
    """Shared bookkeeping of both layouts: ``(i0, i1, lid, raw, norm)``, each layer's
    unscaled spacings from ``spacing_of(k, m)`` for its ``m`` spacings beyond the first."""
    n_layers = len(counts)
    raw = np.zeros(NP + 2)
    norm = np.zeros(n_layers)
    lid = np.zeros(NP + 2, dtype=int)
    i0, i1 = [], []
    start = 0
    for k in range(n_layers):
        finish = start + int(counts[k])
        i0.append(start)
        i1.append(finish)
        if k == n_layers - 1:
            finish = NP + 1
        raw[start:finish + 1] = spacing_of(k, finish - start)
        norm[k] = raw[start:finish + 1].sum()
        lid[start:finish + 1] = k
        start = finish + 1
    i1[-1] = NP + 1
    return i0, i1, lid, raw, norm


def _simss_layout(thicknesses, NP, grad, split):
    counts = layer_points(thicknesses, NP, split)
    return _assemble_layout(counts, NP, lambda k, m: _sub_grid(m, grad))


def _graded_layout(thicknesses, NP, beta=GRADED_POINTS_PER_EFOLD, d0=GRADED_SMALLEST):
    """A layout graded geometrically towards both ends of every layer (reference:
    device.py:160-214): point density ``1/H + beta/(d + d0)`` from each end, ``H`` the total
    thickness over ``NP``; a layer's points beyond its first are its share of the density's
    integral (at least 5), and inside the layer the points equidistribute that integral."""
    L = np.array([float(t) for t in thicknesses])
    H = L.sum() / NP
    n_layers = len(L)

    def integral(x, Lk):
        return x / H + beta * (np.log1p(x / d0) + np.log1p(Lk / d0) - np.log1p((Lk - x) / d0))

    weight = np.array([integral(Lk, Lk) for Lk in L])
    shares = (NP + 2 - n_layers) * weight / weight.sum()
    counts = np.maximum(MIN_GRID_POINTS_PER_LAYER, np.floor(shares).astype(int))
    excess = int(counts.sum()) + n_layers - (NP + 2)
    if excess > 0:
        counts[np.argmax(counts)] -= excess

    def spacing_of(k, m_beyond):
        m = m_beyond + 1  # spacings this layer holds
        target = np.arange(m + 1) * weight[k] / m
        lo, hi = np.zeros(m + 1), np.full(m + 1, L[k])
        for _ in range(_BISECTION_STEPS):
            mid = 0.5 * (lo + hi)
            below = integral(mid, L[k]) < target
            lo = np.where(below, mid, lo)
            hi = np.where(below, hi, mid)
        u = 0.5 * (lo + hi)
        u[0], u[-1] = 0.0, L[k]
        return np.diff(u)

    return _assemble_layout(counts, NP, spacing_of)


def _neumaier_sum(values) -> float:
    """Compensated sum (Neumaier), as Python 3.12+ ``sum`` of floats (reference:
    device.py:817-825): the total thickness is the same number both ways."""
    total, comp = 0.0, 0.0
    for v in values:
        v = float(v)
        s = total + v
        comp += (total - s) + v if abs(total) >= abs(v) else (v - s) + total
        total = s
    return total + comp


def _spacing(L, total, i0, lid, raw, norm):
    """``(x, h)``: each layer's spacing scaled to its thickness over the total, then equalised
    over the three spacings around each interface (reference: device.py:225-233)."""
    h = raw * (L[lid] / (total * norm[lid]))
    for j in i0[1:]:
        h[j - 2:j + 1] = (h[j] + h[j - 1] + h[j - 2]) / 3.0
    x = np.concatenate([[0.0], np.cumsum(total * h[:-1])])
    return x, h


def make_grid(thicknesses, NP: int, grad: float, split: str = "exact"):
    """``(x, h, i0, i1, lid)`` as SIMsalabim's `Make_Grid` (reference: device.py:236-248)."""
    i0, i1, lid, raw, norm = _simss_layout(thicknesses, NP, grad, split)
    L = np.array([float(t) for t in thicknesses])
    x, h = _spacing(L, _neumaier_sum(L), i0, lid, raw, norm)
    return x, h, i0, i1, lid


@dataclass(frozen=True)
class Device:
    """Every static array of one device on its ``NP + 2`` points (SI units). Field names and
    meanings are the reference's (src/jaxdd/device.py:251-382); the optional physics this path
    does not solve is absent (always None / off)."""

    x: np.ndarray
    h: np.ndarray
    L: float
    T: float
    Vt: float
    eps: np.ndarray
    E_c: np.ndarray
    E_v: np.ndarray
    N_c: np.ndarray
    ni: np.ndarray
    N_D: np.ndarray
    N_A: np.ndarray
    mu_n: np.ndarray  # on the intervals: mu_n[i] between x[i] and x[i+1]
    mu_p: np.ndarray
    k_direct: np.ndarray
    srh_Nt: np.ndarray
    srh_Cn: np.ndarray
    srh_Cp: np.ndarray
    srh_nt0: np.ndarray
    srh_pt0: np.ndarray
    G: np.ndarray
    offset_n: np.ndarray
    offset_p: np.ndarray
    V0: float
    VL: float
    n_left: float
    n_right: float
    p_left: float
    p_right: float
    W_L: float
    W_R: float
    G_frac: float
    i0: tuple[int, ...]
    i1: tuple[int, ...]
    anion_mask: np.ndarray
    cation_mask: np.ndarray
    anion_total: float
    cation_total: float
    anion_fixed: np.ndarray
    cation_fixed: np.ndarray
    pre_cond: bool
    V_pre: float
    fix_ions: bool
    R_series: float = 0.0
    R_shunt: float = -1.0
    left_elec: int = -1
    mu_anion: np.ndarray | None = None
    mu_cation: np.ndarray | None = None
    curr_diff: bool = False
    free: tuple[bool, bool] = (False, False)
    statistics: str = "boltzmann"

    @property
    def NP(self) -> int:
        return len(self.x) - 2

    @property
    def lid(self) -> np.ndarray:
        """Each point's layer (0-based), from ``i0``/``i1``."""
        out = np.zeros(len(self.x), dtype=int)
        for k, (a, b) in enumerate(zip(self.i0, self.i1)):
            out[a:b + 1] = k
        return out


@dataclass(frozen=True)
class Structure:
    """What a device's parameters do not change (reference: device.py:385-438): the grid's
    layout, which layers generate, hold traps or hold mobile or fixed ions, the pre-bias."""

    NP: int
    i0: tuple[int, ...]
    i1: tuple[int, ...]
    lid: np.ndarray
    raw: np.ndarray
    norm: np.ndarray
    layer_gen: np.ndarray
    traps: np.ndarray
    anion_mobile: np.ndarray
    anion_fixed: np.ndarray
    cation_mobile: np.ndarray
    cation_fixed: np.ndarray
    pre_cond: bool
    fix_ions: bool
    sfb: tuple[bool, bool] = (False, False)
    gen_mode: str = "file"  # "file" (a profile) or "none" (each layer's G_ehp)
    left_elec: int = -1
    extra: dict = field(default_factory=dict)


#: Per-layer parameters `device_arrays` reads (SIMsalabim's names).
LAYER_PARAMETERS = (
    "L", "eps_r", "E_c", "E_v", "N_c", "N_D", "N_A", "mu_n", "mu_p", "nu_int_n", "nu_int_p",
    "k_direct", "N_t_bulk", "C_n_bulk", "C_p_bulk", "E_t_bulk", "N_anion", "N_cation",
    "mu_anion", "mu_cation", "G_ehp",
)
_OPTIONAL = {"nu_int_n", "nu_int_p", "C_n_bulk", "C_p_bulk", "E_t_bulk", "G_ehp"}


def _sfb(N_A, N_D, E_c, E_v, N_c, Vt):
    """`ObtainWorkFunction` for 'sfb' (reference: device.py:555-573, Boltzmann branch): the
    adjacent layer's equilibrium Fermi level from its net doping, mid-gap without any."""
    net = N_A - N_D
    if net > 0:
        return E_v - Vt * np.log(N_c / abs(net))
    if net < 0:
        return E_c + Vt * np.log(N_c / abs(net))
    return 0.5 * (E_c + E_v)


def device_parameters(setup: Setup, grid_split: str = "exact",
                      grid: str = "graded") -> tuple[Structure, dict]:
    """A SIMsalabim input as its `Structure` and parameters (reference: device.py:576-801):
    a dict of `LAYER_PARAMETERS` (arrays over the layers) and ``T``, ``W_L``, ``W_R``,
    ``G_frac``, ``V_pre``, ``gen_x``, ``gen_G``, ``R_series``, ``R_shunt``."""
    layers = setup.layers
    if setup.i("leftElec") not in (-1, 1):
        raise ValueError("leftElec must be -1 (the left electrode the cathode) or 1 (the anode)")
    NP, grad = setup.i("NP"), setup.f("grad")
    if grid not in GRIDS:
        raise ValueError(f"grid {grid!r}: one of {GRIDS}")
    thick = [ly.raw["L"] for ly in layers]
    if grid == "graded":
        i0, i1, lid, raw, norm = _graded_layout(thick, NP)
    else:
        i0, i1, lid, raw, norm = _simss_layout(thick, NP, grad, grid_split)
    held = bool(setup.i("fixIons", 0)) and not bool(setup.i("preCond", 0))
    problems = ion_region_problems(setup)
    if problems:
        raise NotImplementedError("; ".join(problems))

    def mobile_layers(species):
        out = np.zeros(len(layers), dtype=bool)
        for first, last, moves in ion_regions(setup, species):
            if moves and not held:
                out[first:last + 1] = True
        return out

    def flags(test):
        return np.array([bool(test(ly)) for ly in layers])

    sfb = tuple(setup.raw[k].strip().lower() == "sfb" for k in ("W_L", "W_R"))
    mode = setup.raw.get("genProfile", "none").strip().lower()
    mode = "none" if mode == "none" else "file"
    structure = Structure(
        NP=NP, i0=tuple(i0), i1=tuple(i1), lid=lid, raw=raw, norm=norm,
        layer_gen=flags(lambda ly: ly.i("layerGen")),
        traps=flags(lambda ly: float(ly.raw.get("N_t_bulk", "0")) != 0),
        anion_mobile=mobile_layers("anion"),
        anion_fixed=flags(lambda ly: ly.f("N_anion") != 0 and (ly.f("mu_anion") == 0 or held)),
        cation_mobile=mobile_layers("cation"),
        cation_fixed=flags(lambda ly: ly.f("N_cation") != 0 and (ly.f("mu_cation") == 0 or held)),
        pre_cond=bool(setup.i("preCond", 0)), fix_ions=bool(setup.i("fixIons", 0)),
        sfb=sfb, gen_mode=mode, left_elec=setup.i("leftElec"),
    )
    for k, ly, on_sfb in (("W_L", layers[0], sfb[0]), ("W_R", layers[-1], sfb[1])):
        if not on_sfb and not ly.f("E_c") <= setup.f(k) <= ly.f("E_v"):
            raise ValueError(f"{k} = {setup.f(k)} eV: outside the gap of its layer "
                             f"({ly.f('E_c')} to {ly.f('E_v')} eV)")
    for j in np.flatnonzero(structure.traps):
        e_t = float(layers[j].raw.get("E_t_bulk", "0"))
        if not layers[j].f("E_c") < e_t < layers[j].f("E_v"):
            raise ValueError(f"layer {j + 1}: a bulk trap level outside the gap ({e_t} eV)")

    def value(ly, name):
        return float(ly.raw.get(name, "0")) if name in _OPTIONAL else ly.f(name)

    params = {name: np.array([value(ly, name) for ly in layers]) for name in LAYER_PARAMETERS}
    if mode == "file":
        table = np.loadtxt(io.StringIO(setup.text("genProfile")), skiprows=1)
    else:
        table = np.zeros((2, 2))
    params.update(
        T=setup.f("T"), G_frac=setup.f("G_frac", 1.0),
        W_L=0.0 if sfb[0] else setup.f("W_L"), W_R=0.0 if sfb[1] else setup.f("W_R"),
        V_pre=setup.f("Vpre", 0.0), gen_x=table[:, 0].copy(), gen_G=table[:, 1].copy(),
        R_series=float(setup.raw.get("R_series", "0")),
        R_shunt=float(setup.raw.get("R_shunt", "-1")),
    )
    return structure, params


def _profile_generation(gen_x, gen_G, x, lid, layer_gen, total):
    """`Init_Generation_Profile` from a file (reference: device.py:804-814): the file's axis
    stretched to the device, linear interpolation, each point taking the first interval whose
    end lies beyond it; zero in layers that do not generate and beyond the axis's end."""
    a = gen_x * (total / gen_x[-1])
    j = np.searchsorted(a[1:], x, side="right")
    jj = np.minimum(j, len(a) - 2)
    g = gen_G[jj] + (x - a[jj]) * (gen_G[jj + 1] - gen_G[jj]) / (a[jj + 1] - a[jj])
    return np.where((j < len(a) - 1) & layer_gen[lid], g, 0.0)


def device_arrays(structure: Structure, p: dict) -> Device:
    """Every array of a device from its structure and parameters (reference:
    device.py:877-1063, the subset this path solves)."""
    st = structure
    NP, lid, i0, i1 = st.NP, st.lid, st.i0, st.i1
    Ls = np.asarray(p["L"], dtype=float)
    total = _neumaier_sum(Ls)
    x, h = _spacing(Ls, total, i0, lid, st.raw, st.norm)
    T = float(p["T"])
    Vt = K_B * T / Q
    interior = np.zeros(NP + 2, dtype=bool)
    interior[1:-1] = True
    E_c, E_v, N_c = p["E_c"][lid], p["E_v"][lid], p["N_c"][lid]
    ni = N_c * np.exp(-0.5 * (E_v - E_c) / Vt)
    iface = np.zeros(NP + 2, dtype=bool)
    iface[list(i1[:-1])] = True

    def interleaved(name, nu_name):  # the interface interval's mobility L*h*nu/Vt
        return np.where(iface, (total * h) * p[nu_name][lid] / Vt, p[name][lid])

    k_direct = np.where(interior, p["k_direct"][lid], 0.0)
    trap = interior & st.traps[lid]
    E_t = p["E_t_bulk"]
    nt0 = p["N_c"] * np.exp((p["E_c"] - E_t) / Vt)
    pt0 = p["N_c"] * np.exp((E_t - p["E_v"]) / Vt)
    srh = {key: np.where(trap, v[lid], 0.0) for key, v in (
        ("Nt", p["N_t_bulk"]), ("Cn", p["C_n_bulk"]), ("Cp", p["C_p_bulk"]),
        ("nt0", nt0), ("pt0", pt0))}
    if st.gen_mode == "none":
        G = p["G_ehp"][lid]
    else:
        G = _profile_generation(p["gen_x"], p["gen_G"], x, lid, st.layer_gen, total)
    G = np.where(interior, G, 0.0)
    ions = {}
    for species in ("anion", "cation"):
        conc = p[f"N_{species}"]
        mobile = getattr(st, f"{species}_mobile")
        fixed = getattr(st, f"{species}_fixed")
        ions[species] = (mobile[lid], float(np.sum(np.where(mobile, Ls * conc, 0.0))),
                         np.where(fixed[lid], conc[lid], 0.0))
    dos = Vt * np.log(N_c / N_c[0])
    W_L, W_R = float(p["W_L"]), float(p["W_R"])
    if st.sfb[0]:
        W_L = _sfb(p["N_A"][0], p["N_D"][0], p["E_c"][0], p["E_v"][0], p["N_c"][0], Vt)
    if st.sfb[1]:
        W_R = _sfb(p["N_A"][-1], p["N_D"][-1], p["E_c"][-1], p["E_v"][-1], p["N_c"][-1], Vt)
    return Device(
        x=x, h=h, L=total, T=T, Vt=Vt,
        eps=p["eps_r"][lid] * EPS_0, E_c=E_c, E_v=E_v, N_c=N_c, ni=ni,
        N_D=p["N_D"][lid], N_A=p["N_A"][lid],
        mu_n=interleaved("mu_n", "nu_int_n"), mu_p=interleaved("mu_p", "nu_int_p"),
        k_direct=k_direct,
        srh_Nt=srh["Nt"], srh_Cn=srh["Cn"], srh_Cp=srh["Cp"],
        srh_nt0=srh["nt0"], srh_pt0=srh["pt0"], G=G,
        offset_n=E_c - E_c[0] + dos, offset_p=E_v - E_v[0] - dos,
        V0=0.5 * (-W_L + W_R), VL=0.5 * (W_L - W_R),
        n_left=N_c[0] * np.exp(-(W_L - E_c[0]) / Vt),
        n_right=N_c[-1] * np.exp(-(W_R - E_c[-1]) / Vt),
        p_left=N_c[0] * np.exp((W_L - E_v[0]) / Vt),
        p_right=N_c[-1] * np.exp(-(E_v[-1] - W_R) / Vt),
        W_L=W_L, W_R=W_R, G_frac=float(p["G_frac"]),
        i0=tuple(i0), i1=tuple(i1),
        anion_mask=ions["anion"][0], cation_mask=ions["cation"][0],
        anion_total=ions["anion"][1], cation_total=ions["cation"][1],
        anion_fixed=ions["anion"][2], cation_fixed=ions["cation"][2],
        pre_cond=st.pre_cond, V_pre=float(p["V_pre"]), fix_ions=st.fix_ions,
        R_series=float(p.get("R_series", 0.0)), R_shunt=float(p.get("R_shunt", -1.0)),
        left_elec=st.left_elec,
        mu_anion=p["mu_anion"][lid], mu_cation=p["mu_cation"][lid],
    )


def build_device(setup: Setup, grid_split: str = "exact", grid: str = "graded") -> Device:
    """The device of a SIMsalabim input in NumPy (reference: device.py:1066-1079): ``grid``
    "graded" (the default) or "simss" (SIMsalabim's `Make_Grid`, for a point-by-point
    comparison with SimSS), ``grid_split`` as in `layer_points`."""
    return device_arrays(*device_parameters(setup, grid_split, grid))


def with_parameters(structure: Structure, params: dict, **changes) -> Device:
    """The device of ``structure`` with some parameters changed (the synthetic generator's
    and the gradient checks' way to vary one device)."""
    return device_arrays(structure, dict(params, **changes))


__all__ += ["with_parameters"]
