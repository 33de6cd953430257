"""Shared fixtures. GPU tests are marked ``gpu`` and run the HIP kernels through the C-ABI;
everything else runs on the CPU (the oracle, the host logic, the kernels' CPU replay)."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
for p in (ROOT, ROOT / "tests" / "kernel_emu"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

#: The SimSS 5.36 runs copied from the reference's tests/data, with the tolerance (of the
#: curve's largest |J|) its own tests hold them to (src/../tests/test_solver.py:194-204,
#: :254-272, :878-887, :1316-1358, :1556-1575).
FIXTURES = {
    "thin_l1_light": 1e-5,
    "none_light": 1e-5,
    "frozen_1e17_x1_thin_light": 2e-5,
    "steady_1e17_thin_light": 2e-5,
    "held_thin_light": 1e-9,
    "frozen_thick_htl_dark": 2e-5,
    "resist_light": 1e-5,
    # six voltages 0.1 mV apart around open circuit (SimSS at tolPois 1e-12): the curve's
    # largest |J| is 1.1 A/m^2, 4e-3 of the photocurrent; 3e-7 of it is 1e-9 of the latter
    "thin_l1_voc_light": 3e-7,
    # acceptor bulk traps (3.9e21 m^-3, mid-gap) charged in the absorber, SimSS at tolerances
    # of 1e-9 (tests/test_solver.py:909-922)
    "bulk_acceptor_light": 1e-9,
}


def open_circuit_voltage(V, J):
    """The open-circuit voltage of a curve sampled around it: the root, inside the samples, of
    the cubic least-squares fit of J(V) (the reference's reading of the Voc run,
    tests/test_solver.py:750-764)."""
    m = np.mean(V)
    r = np.roots(np.polyfit(np.asarray(V) - m, J, 3))
    r = r[np.abs(r.imag) == 0].real + m
    r = r[(r > np.min(V)) & (r < np.max(V))]
    assert r.size == 1, r
    return float(r[0])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def fixture_device(case: str):
    from jaxdd_amd.device import build_device
    from jaxdd_amd.simss_input import read_setup

    return build_device(read_setup(GOLDEN / case / "simulation_setup.txt"), grid="simss")


def fixture_inputs(case: str):
    """(structure, numbers) of a SimSS fixture's device on SIMsalabim's grid."""
    from jaxdd_amd.device import device_parameters
    from jaxdd_amd.simss_input import read_setup

    return device_parameters(read_setup(GOLDEN / case / "simulation_setup.txt"), grid="simss")


#: The SimSS runs of one stack (3 layers, 400 points), nine of them generating by the
#: profile file, one by each layer's G_ehp; one with charged (acceptor) bulk traps.
SIMSS_RUNS = ("bulk_acceptor_light", "frozen_1e17_x1_thin_light", "frozen_thick_htl_dark",
              "held_thin_light", "langevin_light", "logv_light", "none_light", "resist_light",
              "steady_1e17_thin_light", "thin_l1_light")
#: Numbers the profile's stretch and scale depend on: the thicknesses and G_frac.
PROFILE_WRT = (("L", 0), ("L", 1), ("L", 2), "G_frac")


def fixture_jv(case: str):
    return np.genfromtxt(GOLDEN / case / "JV.dat", names=True)


@pytest.fixture(scope="session")
def oracle():
    from oracle import ddoracle

    ddoracle.build()
    return ddoracle


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "the GPU tests need the MI355X"
    from jaxdd_amd import _lib

    _lib.load()  # the HIP library must be there: no fallback
    return torch


@pytest.fixture(scope="session")
def emu():
    import emu as _emu

    _emu.build()
    return _emu


def curve_scale(dev, J) -> float:
    """The reference's per-curve scale: the generation's current plus the curve's largest |J|."""
    from jaxdd_amd.batch import j_scale

    return j_scale(dev) + float(np.max(np.abs(J)))


#: A solve that takes this many Newton steps wanders before it converges (or fails): there the
#: outcome depends on the last bits of the arithmetic (a 1e-15 change of one input flips it in
#: the oracle itself), so converged flags are compared on the solves both sides finish sooner.
CHAOTIC_STEPS = 20


def flags_agree(ok, its, ok_ref, its_ref):
    """Converged flags equal wherever neither side took CHAOTIC_STEPS or more; returns
    (agree, number of chaotic entries)."""
    calm = (its < CHAOTIC_STEPS) & (its_ref < CHAOTIC_STEPS)
    return bool(np.array_equal(ok[calm], ok_ref[calm])), int((~calm).sum())


def chaotic_devices(oracle, devs, V, J_ref, rel=1e-13, tol=1e-7):
    """Devices whose oracle curve moves by more than ``tol`` of its scale when the generation
    rate changes by ``rel``: there a start or pre-bias solve lands on one of several nearby
    solutions by the last bits of the arithmetic (frozen ions keep whichever profile the
    pre-bias found), so no implementation can be held to the oracle's curve."""
    import dataclasses

    moved = [dataclasses.replace(d, G=d.G * (1.0 + rel)) for d in devs]
    J2, _, _ = oracle.solve_batch(moved, V, first_max_iter=None)
    return np.array([np.max(np.abs(J2[k] - J_ref[k])) > tol * curve_scale(d, J_ref[k])
                     for k, d in enumerate(devs)])


#: Newton solves the oracle finishes in at most this many steps follow one path: their step
#: counts are compared exactly (longer solves wander, and last-bit differences of the
#: arithmetic change their count by a step or two).
SHORT_STEPS = 8


def short_points(devs, V, its, its_ref, wild):
    """(devices x voltages) mask of the points whose Newton step counts are compared exactly:
    the oracle took at most SHORT_STEPS, neither side CHAOTIC_STEPS, the device is not among
    the ``wild`` ones, and it is not the scan's first voltage, whose solve starts from the
    converged start state so that every step is round-off and its count a matter of the last
    bits (the lowest voltage, or the pre-bias's with a pre-bias)."""
    V = np.asarray(V)
    first = [int(np.searchsorted(V, d.V_pre - 1e-9)) if d.pre_cond else 0 for d in devs]
    mask = (its_ref <= SHORT_STEPS) & (its < CHAOTIC_STEPS) & ~np.asarray(wild)[:, None]
    mask[np.arange(len(devs)), np.minimum(first, len(V) - 1)] = False
    return mask


#: Named numbers of the free-ion stack (jaxdd_amd.builder's): the ions' densities, thicknesses,
#: a mobility, a permittivity, a doping, traps, generation, the work function, temperature and
#: the generation's factor.
ION_WRT = [("N_anion", 1), ("N_cation", 1), ("L", 1), ("L", 0), ("mu_n", 1), ("eps_r", 1),
           ("N_A", 2), ("N_t_bulk", 1), ("G_ehp", 1), "W_L", "T", "G_frac"]


#: Named numbers of Langevin's rate (jaxdd_amd.builder's): the prefactors, the mobilities and
#: permittivity it takes, an interface velocity (the interval's mobility at a layer's last
#: point), a thickness.
LANGEVIN_WRT = [("preLangevin", 1), ("preLangevin", 0), ("mu_n", 1), ("mu_p", 2), ("eps_r", 1),
                ("nu_int_n", 0), ("L", 1)]


def langevin_rows(n, NP, mode, ions=False):
    """(structure, numbers) of ``n`` drawn stacks with Langevin's rate in every layer
    (`preLangevin` 1e-3, 1e-2, 1e-3, as the SimSS run `langevin_light`) at the mobilities
    ``mode`` ("layer" or "simss")."""
    import dataclasses

    from jaxdd_amd import synthetic as S

    t = S.draw(n)
    out = []
    for k in range(n):
        st, p = S.row_inputs(t, k, NP, ions, False)
        st = dataclasses.replace(st, langevin=np.ones(len(st.i0), bool), langevin_mobility=mode)
        out.append((st, dict(p, preLangevin=np.array([1e-3, 1e-2, 1e-3]))))
    return out


def bumped(p, w, h):
    """The numbers ``p`` with ``w`` (a name or a (name, layer) pair) moved by ``h``."""
    q = {k: (np.array(v, dtype=float) if isinstance(v, np.ndarray) else v) for k, v in p.items()}
    if isinstance(w, str):
        q[w] = float(q[w]) + h
    else:
        arr = np.array(q[w[0]], dtype=float)
        arr[w[1]] += h
        q[w[0]] = arr
    return q


def number_of(p, w):
    """The number ``w`` of ``p``."""
    return float(p[w]) if isinstance(w, str) else float(np.asarray(p[w[0]])[w[1]])


def named_fd_errors(oracle, structs, params, wrt, V, J, grad, per_voltage=False):
    """Per device and number, max over V of |dJ/dx - central differences| |x| / the curve's
    current scale (relative step 1e-5, the oracle's curves at tol 1e-12); ``per_voltage``: the
    errors (devices, numbers, voltages) instead of their maxima."""
    from jaxdd_amd.batch import j_scale
    from jaxdd_amd.device import device_arrays

    err = np.zeros((len(structs), len(wrt)) + ((len(V),) if per_voltage else ()))
    for d, (st, p) in enumerate(zip(structs, params)):
        cscale = j_scale(device_arrays(st, p)) + np.max(np.abs(J[d]))
        for k, w in enumerate(wrt):
            x = number_of(p, w)
            h = 1e-5 * abs(x)
            Jpm = [oracle.solve_batch([device_arrays(st, bumped(p, w, s * h))], V, tol=1e-12,
                                      strict=True, first_max_iter=None, max_iter=200)[0][0]
                   for s in (1.0, -1.0)]
            fd = (Jpm[0] - Jpm[1]) / (2 * h)
            e = np.abs(grad[d, :, k] - fd) * abs(x) / cscale
            err[d, k] = e if per_voltage else np.max(e)
    return err
