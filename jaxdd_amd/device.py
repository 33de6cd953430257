"""One device on its grid: the static arrays the steady-state solve reads (host side, NumPy).

Reference interface kept (src/jaxdd/device.py): `Device` (:251-382) and `Structure`
(:385-438) with their field names, `device_parameters` (:576-801), `device_arrays`
(:877-1063), `build_device` (:1066-1079), `make_grid`, `layer_points`, `GRIDS`,
`GRID_SPLITS`. The physics is the subset the batched J-V path solves: Boltzmann carriers,
direct and neutral single-level bulk SRH recombination, generation from a profile file or
uniform per layer, one mobile ionic region per species (or fixed ions), ohmic contacts with
given or flat-band ('sfb') work functions, series and shunt resistance.

The build runs in three stages, each a function of the one before:

1. `device_parameters`: the input as a `Structure` (everything discrete: the grid layout,
   which layers generate, hold traps, mobile or fixed ions, the pre-bias sequence) and a dict
   of numbers (per-layer arrays named as SIMsalabim names them, and device scalars);
2. `_LayerValues`: the per-layer numbers turned into the derived per-layer quantities the
   equations use (intrinsic density, trap reference densities, band offsets, contacts);
3. `device_arrays`: those spread over the grid's points (a gather by each point's layer), the
   interface intervals' mobilities and the generation profile added.

At a fixed structure every array is a smooth function of the numbers, which is what the
gradient path differentiates.
"""

from __future__ import annotations

import io
from dataclasses import dataclass, field

import numpy as np

from .constants import EPS_0, K_B, Q
from .grid import (GRADED_POINTS_PER_EFOLD, GRADED_SMALLEST, GRID_SPLITS, GRIDS, Layout,
                   graded_layout, layer_points, make_grid, simss_layout, stack_sum)
from .simss_input import Setup, ion_region_problems, ion_regions

__all__ = [
    "GRADED_POINTS_PER_EFOLD",
    "GRADED_SMALLEST",
    "GRIDS",
    "GRID_SPLITS",
    "LANGEVIN_MOBILITIES",
    "LAYER_PARAMETERS",
    "Device",
    "Structure",
    "build_device",
    "device_arrays",
    "device_parameters",
    "layer_points",
    "make_grid",
    "with_parameters",
]


@dataclass(frozen=True)
class Device:
    """Every static array of one device on its ``NP + 2`` points (SI units, energies in eV).
    Field names and meanings are the reference's (src/jaxdd/device.py:251-382); the optional
    physics outside this path is absent (its switches fixed off)."""

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
    # per point: the bulk traps' charge type where they are (1 donor, -1 acceptor, 0 neutral
    # and where there are none; reference: device.py:940-948 `trap_charge`, `trap_cwe`)
    srh_type: np.ndarray | None = None

    @property
    def NP(self) -> int:
        return len(self.x) - 2

    @property
    def lid(self) -> np.ndarray:
        """Each point's layer (0-based)."""
        return np.repeat(np.arange(len(self.i0)),
                         np.asarray(self.i1) - np.asarray(self.i0) + 1)


@dataclass(frozen=True)
class Structure:
    """The discrete part of a device (reference: device.py:385-438): the grid's layout (``i0``,
    ``i1``, ``lid``; ``raw`` the unscaled spacing of each point and ``norm`` its sum per
    layer), which layers generate, hold traps, mobile or fixed ions, the pre-bias sequence,
    which work functions are flat-band, the generation's source and the cathode's side."""

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
    langevin: np.ndarray | None = None  # per layer: direct recombination by Langevin's rate
    langevin_mobility: str = "layer"  # `LANGEVIN_MOBILITIES`
    trap_type: np.ndarray | None = None  # per layer: the bulk traps' `bulkTrapType` (1, -1, 0)

    @property
    def layout(self) -> Layout:
        return Layout(NP=self.NP, i0=tuple(self.i0), i1=tuple(self.i1), lid=self.lid,
                      unit=self.raw, unit_sum=self.norm)

    @staticmethod
    def on_layout(layout: Layout, **flags) -> "Structure":
        """A structure on a grid layout with the given flags (the other fields)."""
        return Structure(NP=layout.NP, i0=layout.i0, i1=layout.i1, lid=layout.lid,
                         raw=layout.unit, norm=layout.unit_sum, **flags)


#: Per-layer parameters `device_arrays` reads (SIMsalabim's names).
LAYER_PARAMETERS = (
    "L", "eps_r", "E_c", "E_v", "N_c", "N_D", "N_A", "mu_n", "mu_p", "nu_int_n", "nu_int_p",
    "k_direct", "N_t_bulk", "C_n_bulk", "C_p_bulk", "E_t_bulk", "N_anion", "N_cation",
    "mu_anion", "mu_cation", "G_ehp", "preLangevin",
)
#: Layer parameters an input may leave out (0 then; `preLangevin` 1).
_DEFAULT_ZERO = frozenset({"nu_int_n", "nu_int_p", "C_n_bulk", "C_p_bulk", "E_t_bulk", "G_ehp"})
#: The mobilities Langevin's rate `preLangevin q (mu_n + mu_p) / eps` takes at a point: "layer"
#: its own layer's; "simss" those of the interval after it (at a layer's last point before an
#: interface the interface's, L h nu / V_T), as SIMsalabim takes them.
LANGEVIN_MOBILITIES = ("layer", "simss")


# ---- stage 1: the input as a structure and numbers ----------------------------------------
def _flat_band_level(net_acceptors, E_c, E_v, N_c, Vt):
    """SIMsalabim's 'sfb' work function (its `ObtainWorkFunction`, Boltzmann): the contact
    layer's equilibrium Fermi level, a band's distance ``Vt ln(N_c/|N_A - N_D|)`` inside the
    band of the majority carrier, mid-gap when undoped."""
    if net_acceptors == 0:
        return 0.5 * (E_c + E_v)
    depth = Vt * np.log(N_c / abs(net_acceptors))
    return E_v - depth if net_acceptors > 0 else E_c + depth


def _mobile_layers(setup: Setup, species: str, held: bool) -> np.ndarray:
    """Layers whose ions of ``species`` move (none when the input holds the ions)."""
    out = np.zeros(len(setup.layers), dtype=bool)
    if not held:
        for first, last, moves in ion_regions(setup, species):
            out[first:last + 1] |= moves
    return out


def device_parameters(setup: Setup, grid_split: str = "exact", grid: str = "graded",
                      langevin: str = "layer") -> tuple[Structure, dict]:
    """A SIMsalabim input as its `Structure` and numbers (reference: device.py:576-801): a dict
    of `LAYER_PARAMETERS` (arrays over the layers) and the device's ``T``, ``W_L``, ``W_R``,
    ``G_frac``, ``V_pre``, ``gen_x``, ``gen_G``, ``R_series``, ``R_shunt``. ``grid``: one of
    `GRIDS`; ``grid_split``: the SIMsalabim grid's share rule (`layer_points`); ``langevin``:
    the mobilities of Langevin's rate in layers with `useLangevin` (`LANGEVIN_MOBILITIES`)."""
    if grid not in GRIDS:
        raise ValueError(f"grid {grid!r}: one of {GRIDS}")
    if langevin not in LANGEVIN_MOBILITIES:
        raise ValueError(f"langevin {langevin!r}: one of {LANGEVIN_MOBILITIES}")
    layers = setup.layers
    if setup.i("leftElec") not in (-1, 1):
        raise ValueError("leftElec must be -1 (the left electrode the cathode) or 1 (the anode)")
    problems = ion_region_problems(setup)
    if problems:
        raise NotImplementedError("; ".join(problems))
    NP = setup.i("NP")
    thicknesses = [ly.raw["L"] for ly in layers]
    layout = (graded_layout(thicknesses, NP) if grid == "graded"
              else simss_layout(thicknesses, NP, setup.f("grad"), grid_split))

    def number(ly, name):
        if name == "preLangevin" and name not in ly.raw:
            return 1.0
        if name in _DEFAULT_ZERO and name not in ly.raw:
            return 0.0
        return ly.f(name)

    num = {name: np.array([number(ly, name) for ly in layers]) for name in LAYER_PARAMETERS}
    held = bool(setup.i("fixIons", 0)) and not bool(setup.i("preCond", 0))

    def fixed(species):  # immobile ions, or every ion when the input holds them
        return (num[f"N_{species}"] != 0) & ((num[f"mu_{species}"] == 0) | held)

    sfb = tuple(setup.raw[k].strip().lower() == "sfb" for k in ("W_L", "W_R"))
    gen_mode = "none" if setup.raw.get("genProfile", "none").strip().lower() == "none" else "file"
    structure = Structure.on_layout(
        layout,
        layer_gen=np.array([bool(ly.i("layerGen")) for ly in layers]),
        traps=num["N_t_bulk"] != 0,
        anion_mobile=_mobile_layers(setup, "anion", held), anion_fixed=fixed("anion"),
        cation_mobile=_mobile_layers(setup, "cation", held), cation_fixed=fixed("cation"),
        pre_cond=bool(setup.i("preCond", 0)), fix_ions=bool(setup.i("fixIons", 0)),
        sfb=sfb, gen_mode=gen_mode, left_elec=setup.i("leftElec"),
        langevin=np.array([int(float(ly.raw.get("useLangevin", "0"))) == 1 for ly in layers]),
        langevin_mobility=langevin,
        trap_type=np.array([int(float(ly.raw.get("bulkTrapType", "0"))) for ly in layers]),
    )
    for name, ly, flat in (("W_L", layers[0], sfb[0]), ("W_R", layers[-1], sfb[1])):
        if not flat and not ly.f("E_c") <= setup.f(name) <= ly.f("E_v"):
            raise ValueError(f"{name} = {setup.f(name)} eV: outside the gap of its layer "
                             f"({ly.f('E_c')} to {ly.f('E_v')} eV)")
    outside = structure.traps & ~((num["E_c"] < num["E_t_bulk"]) & (num["E_t_bulk"] < num["E_v"]))
    for j in np.flatnonzero(outside):
        raise ValueError(f"layer {j + 1}: a bulk trap level outside the gap "
                         f"({num['E_t_bulk'][j]} eV)")
    if gen_mode == "file":
        profile = np.loadtxt(io.StringIO(setup.text("genProfile")), skiprows=1, ndmin=2)
    else:
        profile = np.zeros((2, 2))
    num.update(
        T=setup.f("T"), G_frac=setup.f("G_frac", 1.0),
        W_L=0.0 if sfb[0] else setup.f("W_L"), W_R=0.0 if sfb[1] else setup.f("W_R"),
        V_pre=setup.f("Vpre", 0.0), gen_x=profile[:, 0].copy(), gen_G=profile[:, 1].copy(),
        R_series=float(setup.raw.get("R_series", "0")),
        R_shunt=float(setup.raw.get("R_shunt", "-1")),
    )
    return structure, num


# ---- stage 2: derived per-layer quantities ------------------------------------------------
@dataclass(frozen=True)
class _LayerValues:
    """Per-layer quantities of the equations (arrays over the layers) and the contacts."""

    Vt: float
    ni: np.ndarray
    nt0: np.ndarray  # SRH reference densities of the trap level
    pt0: np.ndarray
    offset_n: np.ndarray  # generalised potentials' offsets from the left contact's layer
    offset_p: np.ndarray
    W_L: float
    W_R: float

    @staticmethod
    def of(st: Structure, p: dict) -> "_LayerValues":
        Vt = K_B * float(p["T"]) / Q
        Ec, Ev, Nc = (np.asarray(p[k], dtype=float) for k in ("E_c", "E_v", "N_c"))
        Et = np.asarray(p["E_t_bulk"], dtype=float)
        dos = Vt * np.log(Nc / Nc[0])  # the density of states' part of the band offsets
        net = np.asarray(p["N_A"], dtype=float) - np.asarray(p["N_D"], dtype=float)
        W = [float(p["W_L"]), float(p["W_R"])]
        for side, k in ((0, 0), (1, -1)):
            if st.sfb[side]:
                W[side] = _flat_band_level(net[k], Ec[k], Ev[k], Nc[k], Vt)
        return _LayerValues(
            Vt=Vt, ni=Nc * np.exp(-0.5 * (Ev - Ec) / Vt),
            nt0=Nc * np.exp((Ec - Et) / Vt), pt0=Nc * np.exp((Et - Ev) / Vt),
            offset_n=(Ec - Ec[0]) + dos, offset_p=(Ev - Ev[0]) - dos, W_L=W[0], W_R=W[1],
        )


# ---- stage 3: the arrays on the grid's points ---------------------------------------------
def _generation(st: Structure, p: dict, x: np.ndarray, total: float) -> np.ndarray:
    """Generation rate per point at G_frac = 1: each layer's uniform ``G_ehp``, or the profile
    file's rate interpolated linearly on its axis stretched to the stack, each point in the
    interval k = #{a_1..a_{m-1} <= x} (extrapolated before a_1), zero at or beyond the axis's
    end and in layers that do not generate; none at the contacts (reference:
    device.py:804-814)."""
    if st.gen_mode == "none":
        G = np.asarray(p["G_ehp"], dtype=float)[st.lid]
    else:
        gx, gG = (np.asarray(p[k], dtype=float) for k in ("gen_x", "gen_G"))
        axis = gx * (total / gx[-1])
        j = np.searchsorted(axis[1:], x, side="right")
        k = np.minimum(j, len(axis) - 2)
        G = gG[k] + (x - axis[k]) * (gG[k + 1] - gG[k]) / (axis[k + 1] - axis[k])
        G = np.where(st.layer_gen[st.lid] & (j < len(axis) - 1), G, 0.0)
    G = G.copy()
    G[0] = G[-1] = 0.0
    return G


def device_arrays(structure: Structure, p: dict) -> Device:
    """Every array of a device from its structure and numbers (reference:
    device.py:877-1063, the physics of this path)."""
    st = structure
    lid = st.lid
    L = np.asarray(p["L"], dtype=float)
    total = stack_sum(L)
    x, h = st.layout.place(L, total)
    lv = _LayerValues.of(st, p)
    Vt = lv.Vt
    P = len(x)
    interior = np.ones(P, dtype=bool)
    interior[[0, -1]] = False

    def on_points(a):
        return np.asarray(a, dtype=float)[lid]

    # the interval after each layer's last point (but the device's) crosses an interface,
    # where the carriers' mobility is that of the interface's velocity, L h nu / V_T
    crossing = np.zeros(P, dtype=bool)
    crossing[list(st.i1[:-1])] = True

    def interval_mobility(mu, nu):
        return np.where(crossing, total * h * on_points(p[nu]) / Vt, on_points(p[mu]))

    trap = interior & st.traps[lid]

    def in_traps(a):
        return np.where(trap, on_points(a), 0.0)

    # direct recombination: each layer's k_direct, or where the layer uses Langevin's rate,
    # preLangevin q (mu_n + mu_p) / eps (SIMsalabim's Calc_Langevin_Factor); none at the contacts
    k_direct = on_points(p["k_direct"])
    if st.langevin is not None and np.any(st.langevin):
        if st.langevin_mobility == "simss":
            mob_n = interval_mobility("mu_n", "nu_int_n")
            mob_p = interval_mobility("mu_p", "nu_int_p")
        else:
            mob_n, mob_p = on_points(p["mu_n"]), on_points(p["mu_p"])
        rate = on_points(p["preLangevin"]) * Q / (on_points(p["eps_r"]) * EPS_0)
        k_direct = np.where(st.langevin[lid], rate * (mob_n + mob_p), k_direct)

    ions = {}
    for sp in ("anion", "cation"):
        mobile, fixed = getattr(st, f"{sp}_mobile"), getattr(st, f"{sp}_fixed")
        N = np.asarray(p[f"N_{sp}"], dtype=float)
        ions[sp] = dict(mask=mobile[lid], total=float(np.sum(L * N * mobile)),
                        fixed=np.where(fixed[lid], N[lid], 0.0))
    Ec, Ev, Nc = on_points(p["E_c"]), on_points(p["E_v"]), on_points(p["N_c"])
    W_L, W_R = lv.W_L, lv.W_R
    return Device(
        x=x, h=h, L=total, T=float(p["T"]), Vt=Vt,
        eps=EPS_0 * on_points(p["eps_r"]), E_c=Ec, E_v=Ev, N_c=Nc, ni=lv.ni[lid],
        N_D=on_points(p["N_D"]), N_A=on_points(p["N_A"]),
        mu_n=interval_mobility("mu_n", "nu_int_n"), mu_p=interval_mobility("mu_p", "nu_int_p"),
        k_direct=np.where(interior, k_direct, 0.0),
        srh_Nt=in_traps(p["N_t_bulk"]), srh_Cn=in_traps(p["C_n_bulk"]),
        srh_Cp=in_traps(p["C_p_bulk"]), srh_nt0=in_traps(lv.nt0), srh_pt0=in_traps(lv.pt0),
        G=_generation(st, p, x, total),
        offset_n=lv.offset_n[lid], offset_p=lv.offset_p[lid],
        V0=0.5 * (W_R - W_L), VL=0.5 * (W_L - W_R),
        # ohmic contacts: Boltzmann densities at the electrodes' work functions
        n_left=Nc[0] * np.exp((Ec[0] - W_L) / Vt), n_right=Nc[-1] * np.exp((Ec[-1] - W_R) / Vt),
        p_left=Nc[0] * np.exp((W_L - Ev[0]) / Vt), p_right=Nc[-1] * np.exp((W_R - Ev[-1]) / Vt),
        W_L=W_L, W_R=W_R, G_frac=float(p["G_frac"]),
        i0=tuple(st.i0), i1=tuple(st.i1),
        anion_mask=ions["anion"]["mask"], cation_mask=ions["cation"]["mask"],
        anion_total=ions["anion"]["total"], cation_total=ions["cation"]["total"],
        anion_fixed=ions["anion"]["fixed"], cation_fixed=ions["cation"]["fixed"],
        pre_cond=st.pre_cond, V_pre=float(p["V_pre"]), fix_ions=st.fix_ions,
        R_series=float(p.get("R_series", 0.0)), R_shunt=float(p.get("R_shunt", -1.0)),
        left_elec=st.left_elec,
        mu_anion=on_points(p["mu_anion"]), mu_cation=on_points(p["mu_cation"]),
        srh_type=(None if st.trap_type is None or not np.any(st.traps & (st.trap_type != 0))
                  else in_traps(np.where(st.traps, st.trap_type, 0).astype(float))),
    )


def build_device(setup: Setup, grid_split: str = "exact", grid: str = "graded",
                 langevin: str = "layer") -> Device:
    """The device of a SIMsalabim input (reference: device.py:1066-1079): ``grid`` "graded"
    (the default) or "simss" (SIMsalabim's own grid, to compare point by point with SimSS),
    ``grid_split`` as in `layer_points`, ``langevin`` as in `device_parameters`."""
    return device_arrays(*device_parameters(setup, grid_split, grid, langevin))


def with_parameters(structure: Structure, params: dict, **changes) -> Device:
    """The device of ``structure`` with some numbers changed."""
    return device_arrays(structure, dict(params, **changes))
