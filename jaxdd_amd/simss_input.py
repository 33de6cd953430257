"""SIMsalabim 5.36 input files: `simulation_setup.txt`, the layer files it names, and ZimT's
time-step file.

The file format is SIMsalabim's: ``name = value`` per line, ``*`` starting a comment, any
order. Values stay strings until a number is asked for, so that a parameter outside this path
is refused by name (`check_supported`) rather than dropped.

Reference interface kept (src/jaxdd/simss_input.py): `parse_parameters`, `Layer`, `Setup`,
`read_setup`, `setup_from_texts`, `simss_voltages`, `ion_regions`, `ion_region_problems`,
`check_supported`, `TVG`, `read_tvg`, with their error classes and messages where a caller
may match them.
"""

from __future__ import annotations

import io
import os
import re
from collections.abc import Callable, Mapping
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .constants import K_B, Q

__all__ = [
    "TVG",
    "Layer",
    "Setup",
    "check_supported",
    "ion_region_problems",
    "ion_regions",
    "parse_parameters",
    "read_setup",
    "read_tvg",
    "setup_from_texts",
    "simss_voltages",
]

# name = value, the value up to a comment; names are anything but '=' and '*'
_ASSIGNMENT = re.compile(r"^\s*([^=*]*?)\s*=\s*([^*]*?)\s*(?:\*.*)?$")


def parse_parameters(text: str) -> dict[str, str]:
    """``{name: value}`` of a parameter file's text; lines without an assignment before any
    comment are skipped, and a name given twice keeps its last value."""
    found = (_ASSIGNMENT.match(line) for line in text.splitlines())
    return {m.group(1): m.group(2) for m in found if m and m.group(1)}


class _Numbers:
    """Typed access to a parameter mapping ``raw``."""

    raw: dict[str, str]

    def f(self, name: str, default: float | None = None) -> float:
        """Parameter ``name`` as a float; ``default`` where the file has none."""
        if name not in self.raw and default is not None:
            return float(default)
        return float(self.raw[name])

    def i(self, name: str, default: int | None = None) -> int:
        """Parameter ``name`` as an integer (SIMsalabim writes some as reals)."""
        return int(self.f(name, default))


@dataclass(frozen=True)
class Layer(_Numbers):
    """One layer's parameters as read (SI units, energies in eV)."""

    raw: dict[str, str]


@dataclass(frozen=True)
class Setup(_Numbers):
    """The setup file and its layers, from the left electrode. The files it names (the
    generation profile, ZimT's time steps) come from ``files`` (name to text) where given,
    else from ``directory``."""

    raw: dict[str, str]
    layers: tuple[Layer, ...]
    directory: Path | None = None
    files: Mapping[str, str] | None = None

    def text(self, name: str) -> str:
        """The text of the file parameter ``name`` names."""
        return self.read(self.raw[name], name)

    def read(self, file: str, name: str | None = None) -> str:
        """The text of ``file`` (named by parameter ``name``, for the error)."""
        if self.files is not None and file in self.files:
            return self.files[file]
        if self.directory is None:
            label = file if name is None else f"{name} = {file}"
            raise FileNotFoundError(f"{label}: not among the setup's files")
        return (self.directory / file).read_text()


def _setup(raw: dict[str, str], fetch: Callable[[str], str], **where) -> Setup:
    """A setup whose layer files l1, l2, ... (consecutive from 1) ``fetch`` returns."""
    names = []
    while f"l{len(names) + 1}" in raw:
        names.append(raw[f"l{len(names) + 1}"])
    layers = tuple(Layer(parse_parameters(fetch(n))) for n in names)
    return Setup(raw=raw, layers=layers, **where)


def read_setup(path: str | os.PathLike) -> Setup:
    """The setup file at ``path`` with the layer files it names (beside it)."""
    path = Path(path)
    folder = path.parent
    return _setup(parse_parameters(path.read_text()), lambda n: (folder / n).read_text(),
                  directory=folder)


def setup_from_texts(files: Mapping[str, str], setup: str = "simulation_setup.txt") -> Setup:
    """`read_setup` from file names mapped to texts, the setup's under ``setup``."""
    texts = dict(files)
    return _setup(parse_parameters(texts[setup]), texts.__getitem__, files=texts)


# ---- SimSS's voltages -----------------------------------------------------------------------
def _measured_voltages(setup: Setup) -> np.ndarray:
    table = np.genfromtxt(io.StringIO(setup.text("expJV")), names=True)
    if "Vext" not in (table.dtype.names or ()):
        raise ValueError(f"expJV = {setup.raw['expJV']}: no column Vext")
    V = np.atleast_1d(table["Vext"]).astype(float)
    if V.size < 4:
        raise ValueError("Not enough experimental data points (SimSS needs 4).")
    return V


def _uniform_voltages(setup: Setup, vmin: float, vmax: float) -> np.ndarray:
    step = setup.f("Vstep")
    if step <= 0:
        raise ValueError("Vstep should be positive.")
    # SimSS counts the steps that fit, with a margin for a range that is a whole number of them
    count = 1 + int(np.trunc((vmax - vmin) / step + 1e-10))
    return vmin + step * np.arange(count)


def _accumulating_voltages(setup: Setup, vmin: float, vmax: float) -> np.ndarray:
    acc, count = setup.f("Vacc"), setup.i("NJV")
    if vmin <= acc <= vmax:
        raise ValueError("Invalid Vacc selected, must be outside [Vmin, Vmax].")
    if abs(vmin - vmax) < 1e-10:
        raise ValueError("Do not use Vdist=2 when Vmin = Vmax.")
    if count < 2:
        raise ValueError(f"NJV = {count}: Vdist = 2 needs at least 2 voltages")
    # the distance to Vacc shrinks geometrically from (Vacc - Vmin) to (Vacc - Vmax)
    far, near = acc - vmin, acc - vmax
    fraction = 1 - np.arange(count) / (count - 1)
    return acc - near * np.exp(fraction * np.log(far / near))


def simss_voltages(setup: Setup) -> np.ndarray:
    """The internal voltages SimSS runs, in its order (its `Init_Voltages_and_Tasks`):
    ``Vdist = 1`` from Vmin to Vmax by Vstep, ``Vdist = 2`` NJV voltages crowding towards
    Vacc; reversed for ``Vscan = -1``; with ``useExpData`` the ``Vext`` column of ``expJV``.
    SimSS's own checks raise `ValueError`."""
    if setup.i("useExpData", 0):
        return _measured_voltages(setup)
    vdist, vscan = setup.i("Vdist"), setup.i("Vscan")
    vmin, vmax = setup.f("Vmin"), setup.f("Vmax")
    rules = {1: _uniform_voltages, 2: _accumulating_voltages}
    if vdist not in rules:
        raise ValueError(f"Vdist = {vdist}: SimSS has 1 (uniform) and 2 (logarithmic)")
    if vscan not in (-1, 1):
        raise ValueError("Vscan must be either -1 or 1.")
    if vmin > vmax:
        raise ValueError("Vmin should not be greater than Vmax.")
    V = rules[vdist](setup, vmin, vmax)
    return np.ascontiguousarray(V[::vscan])


# ---- ionic regions --------------------------------------------------------------------------
def ion_regions(setup: Setup, species: str) -> list[tuple[int, int, bool]]:
    """SIMsalabim's ionic regions of ``species`` ("anion" or "cation"; its
    `Init_Ionic_Region`) as ``(first, last, mobile)`` layer indices from 0.

    A layer takes part when its ions may enter it (``ionsMayEnter``) or it holds immobile
    ions. Along a run of such layers a region closes at a layer whose ions do not move, at a
    layer followed by one they may not enter, and at the run's end; it is kept if the ions
    counted since the last close are not zero, mobile if the closing layer's ions move. As in
    SimSS 5.36 every region of a run starts at the run's first layer. A layer with mobile ions
    that they may not enter is an error (`ValueError`), as in SimSS."""
    layers = setup.layers
    regions: list[tuple[int, int, bool]] = []
    run_start: int | None = None
    counted = 0.0
    for j, ly in enumerate(layers):
        mu, N = ly.f(f"mu_{species}"), ly.f(f"N_{species}")
        enter = bool(ly.i("ionsMayEnter"))
        if mu * N > 0 and not enter:
            raise ValueError(f"layer {j + 1}: If a layer contains mobile ions, then ionsMayEnter "
                             "(of that layer) must be 1.")
        if not (enter or (mu == 0 and N > 0)):
            run_start, counted = None, 0.0
            continue
        run_start = j if run_start is None else run_start
        counted += ly.f("L") * N
        last = j + 1 == len(layers)
        if last or mu == 0 or not layers[j + 1].i("ionsMayEnter"):
            if counted != 0:
                regions.append((run_start, j, mu > 0))
            counted = 0.0
    return regions


def ion_region_problems(setup: Setup) -> list[str]:
    """The ionic regions this path does not take: SimSS's two readings of multi-layer regions
    that contradict the input (a region taking in a layer the ions may not enter; a region
    whose last layer's mobility 0 holds every ion of it), and more than one mobile region per
    species (the solve carries one normalisation slot per species)."""
    problems = []
    for species in ("anion", "cation"):
        regions = ion_regions(setup, species)
        for first, last, mobile in regions:
            span = list(enumerate(setup.layers[first:last + 1], start=first + 1))
            if len(span) < 2:
                continue
            where = f"{species}s: SimSS's region of layers {first + 1}-{last + 1}"
            closed = [k for k, ly in span if not ly.i("ionsMayEnter")]
            if closed:
                problems.append(f"{where} takes in layer {closed[0]}, which they may not enter")
            moving = any(ly.f(f"mu_{species}") * ly.f(f"N_{species}") > 0 for _, ly in span)
            if moving and not mobile:
                problems.append(f"{where} has mobility 0 in its last layer, so it holds them all")
        if sum(mobile for *_, mobile in regions) > 1:
            problems.append(f"{species}s: more than one mobile ionic region (ion slots > 2)")
    return problems


# ---- what this path solves -----------------------------------------------------------------
#: Layer switches whose other values turn on physics outside this path, with their off value.
_LAYER_SWITCHES = {
    "mobnDep": "0", "mobpDep": "0", "fieldDepG": "0",
    "intTrapFile": "none", "bulkTrapFile": "none",
}


def _switched_on(value: str, off: str) -> bool:
    if off == "none":
        return value.strip().lower() != "none"
    return float(value) != float(off)


def _device_problems(setup: Setup) -> list[str]:
    raw = setup.raw
    out = []
    form = float(raw.get("currDiffInt", 2))
    if form not in (1.0, 2.0):
        out.append(f"currDiffInt = {raw['currDiffInt']} (SimSS has 1 and 2)")
    elif form == 1.0:
        out.append("currDiffInt = 1 (the differential current): not on this path")
    if float(raw.get("R_series", 0)) < 0 or float(raw.get("R_shunt", -1)) == 0:
        out.append("R_series < 0 or R_shunt = 0 (SimSS refuses both)")
    out += [f"{k} = {raw[k]}: finite surface recombination is not on this path"
            for k in ("S_n_L", "S_p_L", "S_n_R", "S_p_R") if float(raw.get(k, "-1")) >= 0]
    out += [f"{k} = free: a free end is not on this path"
            for k in ("W_L", "W_R") if raw.get(k, "").strip().lower() == "free"]
    if raw.get("genProfile", "none").strip().lower() == "calc":
        out.append("genProfile = calc: the transfer-matrix generation is not on this path")
    return out


def _layer_problems(j: int, ly: Layer, last: bool) -> list[str]:
    out = []
    kind = ly.raw.get("bulkTrapType", "0").strip()
    if float(kind) not in (-1.0, 0.0, 1.0):
        out.append(f"layer {j}: bulkTrapType = {kind} (-1 acceptor, 0 neutral, 1 donor)")
    for name, off in _LAYER_SWITCHES.items():
        value = ly.raw.get(name, off)
        if _switched_on(value, off):
            out.append(f"layer {j}: {name} = {value.strip()} is not on this path")
    if not last and float(ly.raw.get("N_t_int", "0")) != 0:
        out.append(f"layer {j}: N_t_int = {ly.raw['N_t_int']} (interface traps) is not on this "
                   "path")
    out += [f"layer {j}: {name} (Auger) is not on this path"
            for name in ("C_n_Auger", "C_p_Auger") if float(ly.raw.get(name, "0")) != 0]
    out += [f"layer {j}: {name} (ion sites) is not on this path"
            for name in ("N_max_anion", "N_max_cation")
            if np.isfinite(float(ly.raw.get(name, "inf")))]
    return out


def check_supported(setup: Setup) -> None:
    """Raise `NotImplementedError` naming every input outside this path: the steady state (and
    ZimT's transients) of Boltzmann carriers with direct and neutral single-level bulk SRH
    recombination, a generation profile from a file or uniform per layer, ohmic contacts, one
    mobile ionic region per species (free, frozen after a pre-bias, or held), series and shunt
    resistance, either electrode the cathode."""
    n = len(setup.layers)
    problems = _device_problems(setup)
    for j, ly in enumerate(setup.layers, start=1):
        problems += _layer_problems(j, ly, j == n)
    problems += ion_region_problems(setup)
    if problems:
        raise NotImplementedError("outside what jaxdd_amd solves:\n  " + "\n  ".join(problems))


# ---- ZimT's time steps ---------------------------------------------------------------------
#: ln of the largest x87 extended real, SIMsalabim's bound on |V|/V_T (ZimT checks 1.95x it).
_LN_MAX_EXTENDED = 11356.523406294143


@dataclass(frozen=True)
class TVG:
    """ZimT's time steps (its `tVGFile`): per step the time ``t`` (s), the voltage ``Vext`` (V;
    NaN at an open circuit, ``oc``), ``G_frac`` and ``track`` (0-3, ZimT's automatic stop);
    ``dti`` = 1/dt from the step before, 0 at t = 0 (a steady state). With them the setup's
    ``failureMode`` (0 stop at a failed step, 1 keep it, 2 skip it), ``convVar`` and the
    tolerances ZimT compares states with (``tolPois`` V, ``tolDens``, ``tolCurr`` A/m^2)."""

    t: np.ndarray
    Vext: np.ndarray
    oc: np.ndarray
    G_frac: np.ndarray
    track: np.ndarray
    dti: np.ndarray
    failure_mode: int = 1
    conv_var: int = 1
    tol_pois: float = 1e-6
    tol_dens: float = 1e-6
    tol_curr: float = 1e-3


def _tvg_rows(lines, name: str):
    """The step rows after the header, up to the end or an empty line: (t, V, oc, G, track)."""
    for line in lines:
        fields = line.split()
        if not fields:
            return
        t, v, g, track = (fields + ["", "", "", ""])[:4]
        try:
            oc = v.lower() == "oc"
            row = (float(t), np.nan if oc else float(v), oc, float(g), int(track))
            if track not in ("0", "1", "2", "3"):
                raise ValueError
        except ValueError:
            raise ValueError(f"Error while reading from file {name}; offending line: "
                             f"{line!r}") from None
        yield row


def read_tvg(setup: Setup) -> TVG:
    """ZimT 5.36's `tVGFile`: the lines before the header ``t Vext G_frac Track`` are skipped,
    then one step per line up to the end or the first empty line (columns past the fourth
    ignored). ZimT's checks raise `ValueError`: the header, a first step at t = 0, times that
    increase or return to 0, voltages within SIMsalabim's range, Track in 0-3, and a Track 1
    (not on the first line, which ZimT forgets) before any Track 2 or 3."""
    name = setup.raw["tVGFile"]
    lines = iter(setup.text("tVGFile").splitlines())
    if not any("".join(line.lower().split()).startswith("tvextg_fractrack") for line in lines):
        raise ValueError(f"Could not find correct header 't Vext G_frac Track' in {name}.")
    rows = list(_tvg_rows(lines, name))
    if not rows or rows[0][0] != 0:
        raise ValueError("tVGFile did not specify steady-state (t=0)")
    limit = 1.95 * _LN_MAX_EXTENDED * K_B * setup.f("T") / Q
    dti, last = [], 0.0
    for k, (t, v, oc, _, track) in enumerate(rows):
        if t > last:
            dti.append(1.0 / (t - last))
        elif t == 0:
            dti.append(0.0)
        else:
            raise ValueError("Time steps must either be positive, or you should go back to t=0 "
                             f"(steady-state); at t = {t}")
        if not oc and abs(v) > limit:
            raise ValueError(f"V is too {'small' if v < 0 else 'large'}; at t = {t}")
        last = t
    tracks = [r[4] for r in rows]
    first_reference = next((k for k, tr in enumerate(tracks) if tr == 1 and k > 0), len(rows))
    if any(tr in (2, 3) for tr in tracks[1:first_reference]) or tracks[0] in (2, 3):
        raise ValueError("Error in tVGFile: before setting Track=2 or 3, you need to have at "
                         "least one Track=1")
    raw = setup.raw
    failure_mode, conv_var = int(raw.get("failureMode", 1)), int(raw.get("convVar", 1))
    if failure_mode not in (0, 1, 2):
        raise ValueError("Invalid failureMode selected.")
    if conv_var not in (1, 2, 3, 4):
        raise ValueError("Invalid convVar selected.")
    t, V, oc, G, track = (np.array(col) for col in zip(*rows))
    return TVG(t=t.astype(float), Vext=V.astype(float), oc=oc.astype(bool),
               G_frac=G.astype(float), track=track.astype(int), dti=np.array(dti),
               failure_mode=failure_mode, conv_var=conv_var,
               tol_pois=float(raw.get("tolPois", 1e-6)),
               tol_dens=float(raw.get("tolDens", 1e-6)),
               tol_curr=float(raw.get("tolCurr", 1e-3)))
