"""SIMsalabim 5.36 input files: `simulation_setup.txt` and the layer files it names.

Host-side reader for the devices the steady-state J-V path solves. The format is SIMsalabim's:
one ``name = value * comment`` per line, ``*`` starts a comment, order free. Values stay
strings until used, so that a parameter outside this path can be refused by name
(`check_supported`) instead of being dropped.

Reference interface mirrored: src/jaxdd/simss_input.py:40-333 (`parse_parameters`, `Layer`,
`Setup`, `read_setup`, `setup_from_texts`, `simss_voltages`, `ion_regions`, `check_supported`).
"""

from __future__ import annotations

import io
import os
from collections.abc import Mapping
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = [
    "Layer",
    "Setup",
    "check_supported",
    "ion_region_problems",
    "ion_regions",
    "parse_parameters",
    "read_setup",
    "setup_from_texts",
    "simss_voltages",
]


def parse_parameters(text: str) -> dict[str, str]:
    """``{name: value}`` of a parameter file's text (reference: simss_input.py:40-50)."""
    out: dict[str, str] = {}
    for line in text.splitlines():
        body = line.partition("*")[0]
        name, eq, value = body.partition("=")
        if not eq:
            continue
        name = name.strip()
        if name:
            out[name] = value.strip()
    return out


@dataclass(frozen=True)
class Layer:
    """One layer's parameters as read (SI units, energies in eV)."""

    raw: dict[str, str]

    def f(self, name: str) -> float:
        return float(self.raw[name])

    def i(self, name: str) -> int:
        return int(float(self.raw[name]))


@dataclass(frozen=True)
class Setup:
    """The setup file with its layers in order from the left electrode. Files it names (the
    generation profile) come from ``files`` (name to text) where given, else from
    ``directory`` (reference: simss_input.py:69-104)."""

    raw: dict[str, str]
    layers: tuple[Layer, ...]
    directory: Path | None = None
    files: Mapping[str, str] | None = None

    def f(self, name: str, default: float | None = None) -> float:
        if default is not None and name not in self.raw:
            return float(default)
        return float(self.raw[name])

    def i(self, name: str, default: int | None = None) -> int:
        return int(self.f(name, default))

    def text(self, name: str) -> str:
        """The text of the file that parameter ``name`` names."""
        return self.read(self.raw[name], name)

    def read(self, file: str, name: str | None = None) -> str:
        if self.files is not None and file in self.files:
            return self.files[file]
        if self.directory is None:
            what = f"{name} = {file}" if name else file
            raise FileNotFoundError(f"{what}: not among the setup's files")
        return (self.directory / file).read_text()


def _layers(raw: Mapping[str, str], read) -> tuple[Layer, ...]:
    out = []
    k = 1
    while f"l{k}" in raw:
        out.append(Layer(parse_parameters(read(raw[f"l{k}"]))))
        k += 1
    return tuple(out)


def read_setup(path: str | os.PathLike) -> Setup:
    """The setup file and every layer file ``l1``, ``l2``, ... beside it."""
    path = Path(path)
    raw = parse_parameters(path.read_text())
    return Setup(raw=raw, layers=_layers(raw, lambda f: (path.parent / f).read_text()),
                 directory=path.parent)


def setup_from_texts(files: Mapping[str, str], setup: str = "simulation_setup.txt") -> Setup:
    """`read_setup` from a mapping of file names to texts (no files on disk)."""
    raw = parse_parameters(files[setup])
    return Setup(raw=raw, layers=_layers(raw, lambda f: files[f]), files=dict(files))


def simss_voltages(setup: Setup) -> np.ndarray:
    """SimSS's internal voltages in its order (`Init_Voltages_and_Tasks`; reference:
    simss_input.py:107-146): ``Vmin`` to ``Vmax`` by ``Vstep`` (``Vdist = 1``) or ``NJV``
    accumulating towards ``Vacc`` (``Vdist = 2``), reversed for ``Vscan = -1``."""
    if "useExpData" in setup.raw and setup.i("useExpData"):
        table = np.genfromtxt(io.StringIO(setup.text("expJV")), names=True)
        if "Vext" not in (table.dtype.names or ()):
            raise ValueError(f"expJV = {setup.raw['expJV']}: no column Vext")
        V = np.atleast_1d(table["Vext"]).astype(float)
        if V.size < 4:
            raise ValueError("Not enough experimental data points (SimSS needs 4).")
        return V
    vdist, vscan = setup.i("Vdist"), setup.i("Vscan")
    vmin, vmax = setup.f("Vmin"), setup.f("Vmax")
    if vdist not in (1, 2):
        raise ValueError(f"Vdist = {vdist}: SimSS has 1 (uniform) and 2 (logarithmic)")
    if abs(vscan) != 1:
        raise ValueError("Vscan must be either -1 or 1.")
    if vmin > vmax:
        raise ValueError("Vmin should not be greater than Vmax.")
    if vdist == 1:
        step = setup.f("Vstep")
        if step <= 0:
            raise ValueError("Vstep should be positive.")
        count = int(np.trunc((vmax - vmin) / step + 1e-10)) + 1
        V = vmin + step * np.arange(count)
    else:
        acc, count = setup.f("Vacc"), setup.i("NJV")
        if vmin <= acc <= vmax:
            raise ValueError("Invalid Vacc selected, must be outside [Vmin, Vmax].")
        if abs(vmin - vmax) < 1e-10:
            raise ValueError("Do not use Vdist=2 when Vmin = Vmax.")
        if count < 2:
            raise ValueError(f"NJV = {count}: Vdist = 2 needs at least 2 voltages")
        d = acc - vmax
        V = acc - d * np.exp((1 - np.arange(count) / (count - 1)) * np.log((acc - vmin) / d))
    return V[::-1].copy() if vscan == -1 else V


def ion_regions(setup: Setup, species: str) -> list[tuple[int, int, bool]]:
    """SIMsalabim's ionic regions of ``species`` (`Init_Ionic_Region`; reference:
    simss_input.py:242-270): ``(first, last, mobile)`` layer indices. A region is a run of
    adjacent layers that ions may enter (or one layer of immobile ions) holding some ions; it
    ends at a layer whose ions do not move or before one they may not enter. As in SimSS, a
    region found does not reset the run's start."""
    regions: list[tuple[int, int, bool]] = []
    start, total = None, 0.0
    layers = setup.layers
    for j, ly in enumerate(layers):
        mob, conc = ly.f(f"mu_{species}"), ly.f(f"N_{species}")
        if mob * conc > 0 and not ly.i("ionsMayEnter"):
            raise ValueError(f"layer {j + 1}: If a layer contains mobile ions, then ionsMayEnter "
                             "(of that layer) must be 1.")
        if ly.i("ionsMayEnter") or (mob == 0 and conc > 0):
            total += ly.f("L") * conc
            if start is None:
                start = j
        else:
            start, total = None, 0.0
            continue
        closes = j == len(layers) - 1 or not layers[j + 1].i("ionsMayEnter") or mob == 0
        if closes:
            if total != 0:
                regions.append((start, j, mob > 0))
            total = 0.0
    return regions


def ion_region_problems(setup: Setup) -> list[str]:
    """The ionic regions this path does not take (reference: simss_input.py:294-309): SimSS's
    two quirks with regions spanning several layers, and (this path's own limit) more than one
    mobile region per species, which needs more than two ion normalisation slots."""
    problems = []
    for species in ("anion", "cation"):
        regions = ion_regions(setup, species)
        for first, last, mobile in regions:
            if first == last:
                continue
            span = [setup.layers[k] for k in range(first, last + 1)]
            where = f"{species}s: SimSS's region of layers {first + 1}-{last + 1}"
            closed = [first + 1 + k for k, ly in enumerate(span) if not ly.i("ionsMayEnter")]
            if closed:
                problems.append(f"{where} takes in layer {closed[0]}, which they may not enter")
            if not mobile and any(ly.f(f"mu_{species}") * ly.f(f"N_{species}") > 0 for ly in span):
                problems.append(f"{where} has mobility 0 in its last layer, so it holds them all")
        if sum(1 for r in regions if r[2]) > 1:
            problems.append(f"{species}s: more than one mobile ionic region (ion slots > 2)")
    return problems


#: Layer parameters whose non-default values switch on physics outside this path.
_LAYER_OFF = {
    "mobnDep": "0", "mobpDep": "0", "fieldDepG": "0", "useLangevin": "0",
    "bulkTrapType": "0", "intTrapFile": "none", "bulkTrapFile": "none",
}


def check_supported(setup: Setup) -> None:
    """Raise `NotImplementedError` naming every input outside this path (reference:
    simss_input.py:312-333, narrowed to the steady-state solve this package runs on the GPU:
    Boltzmann carriers, direct and neutral bulk SRH recombination, a generation profile from
    a file or uniform per layer, ohmic contacts, one mobile ionic region per species, free or
    frozen after a pre-bias or held, series and shunt resistance, either electrode the
    cathode)."""
    problems = []
    raw = setup.raw
    if float(raw.get("currDiffInt", 2)) not in (1.0, 2.0):
        problems.append(f"currDiffInt = {raw['currDiffInt']} (SimSS has 1 and 2)")
    elif float(raw.get("currDiffInt", 2)) == 1.0:
        problems.append("currDiffInt = 1 (the differential current): not on this path")
    if float(raw.get("R_series", 0)) < 0 or float(raw.get("R_shunt", -1)) == 0:
        problems.append("R_series < 0 or R_shunt = 0 (SimSS refuses both)")
    for k in ("S_n_L", "S_p_L", "S_n_R", "S_p_R"):
        if float(raw.get(k, "-1")) >= 0:
            problems.append(f"{k} = {raw[k]}: finite surface recombination is not on this path")
    for k in ("W_L", "W_R"):
        if raw.get(k, "").strip().lower() == "free":
            problems.append(f"{k} = free: a free end is not on this path")
    if raw.get("genProfile", "none").strip().lower() == "calc":
        problems.append("genProfile = calc: the transfer-matrix generation is not on this path")
    for j, ly in enumerate(setup.layers, start=1):
        for name, off in _LAYER_OFF.items():
            if name == "bulkTrapType" and float(ly.raw.get("N_t_bulk", "0")) == 0:
                continue  # a trap type without traps does nothing
            value = ly.raw.get(name, off).strip()
            same = value.lower() == off if off == "none" else float(value) == float(off)
            if not same:
                problems.append(f"layer {j}: {name} = {value} is not on this path")
        if float(ly.raw.get("N_t_int", "0")) != 0 and j < len(setup.layers):
            problems.append(f"layer {j}: N_t_int = {ly.raw['N_t_int']} (interface traps) is not "
                            "on this path")
        for name in ("C_n_Auger", "C_p_Auger"):
            if float(ly.raw.get(name, "0")) != 0:
                problems.append(f"layer {j}: {name} (Auger) is not on this path")
        for name in ("N_max_anion", "N_max_cation"):
            if np.isfinite(float(ly.raw.get(name, "inf"))):
                problems.append(f"layer {j}: {name} (ion sites) is not on this path")
    problems += ion_region_problems(setup)
    if problems:
        raise NotImplementedError("outside what jaxdd_amd solves:\n  " + "\n  ".join(problems))
