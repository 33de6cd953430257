"""Seeded synthetic device batches for the parity tests and the benchmark.

Four workloads of the batched steady-state J-V path, each a p-i-n stack (ETL / perovskite /
HTL) built by `device.device_arrays` from drawn parameters:

- ``c1``: one device, NP = 256, ETL 30 nm / perovskite 300 nm / HTL 30 nm, FP64, no ions.
- ``c2``: NP = 400, per-device thicknesses, mobilities, dopings, bulk trap density and
  generation drawn from the seed; no ions.
- ``c3``: as ``c2`` plus mobile anions and cations in the perovskite, either free at every
  voltage (``variant="free"``) or settled at a 1.2 V pre-bias and frozen (``"frozen"``).
- ``c4``: as ``c2`` (the gradient workload's devices).

The fixed values the workload description leaves open are pinned to the reference's
``tests/data/thin_l1_light`` inputs (transport-layer band edges, work functions, interface
velocities, direct recombination, temperature). The perovskite is p-doped (``N_A``).
Voltages: 61 from -0.2 V to 1.3 V. Seed: 20261107 (the reference's benchmark seed,
benchmarks/run.py:63).
"""

from __future__ import annotations

import hashlib

import numpy as np

from .device import Structure, device_arrays
from .grid import graded_layout

SEED = 20261107
VOLTAGES = np.linspace(-0.2, 1.3, 61)

#: Per-layer values shared by every workload (ETL, perovskite, HTL).
BASE = {
    "eps_r": np.array([5.0, 24.0, 3.5]),
    "E_c": np.array([4.0, 4.0, 3.0]),
    "E_v": np.array([6.6, 5.6, 5.6]),
    "N_c": np.array([2e24, 2e24, 2e24]),
    "nu_int_n": np.array([1e5, 1e5, 1e5]),
    "nu_int_p": np.array([1e5, 1e5, 1e5]),
    "k_direct": np.array([4.7281398e-17, 4.7281398e-17, 4.7281398e-17]),
    "C_n_bulk": np.array([1e-14, 1e-14, 1e-14]),
    "C_p_bulk": np.array([1e-14, 1e-14, 1e-14]),
    "E_t_bulk": np.array([4.8, 4.8, 4.8]),
    "mu_anion": np.zeros(3),
    "mu_cation": np.zeros(3),
    "N_anion": np.zeros(3),
    "N_cation": np.zeros(3),
}
SCALARS = {"T": 298.15, "W_L": 4.275, "W_R": 5.325, "G_frac": 1.0, "V_pre": 0.0,
           "R_series": 0.0, "R_shunt": -1.0, "gen_x": np.zeros(2), "gen_G": np.zeros(2)}
#: Mobile ion mobility (m^2/Vs), as the reference's `steady_1e17_thin_light`.
ION_MOBILITY = 1e-12


def _structure(L, NP, ions=False, pre_cond=False, fix_ions=False) -> Structure:
    traps = np.array([False, True, False])
    mobile = np.array([False, ions, False])
    return Structure.on_layout(
        graded_layout([repr(float(x)) for x in L], NP),
        layer_gen=np.array([False, True, False]), traps=traps,
        anion_mobile=mobile, anion_fixed=np.zeros(3, bool),
        cation_mobile=mobile, cation_fixed=np.zeros(3, bool),
        pre_cond=pre_cond, fix_ions=fix_ions, gen_mode="none", left_elec=-1,
    )


def c1_device():
    """The one-device workload: 30/300/30 nm, NP = 256, uniform G = 4e27 in the perovskite,
    SRH tau = 1e-7 s (N_t = 1e21 m^-3, C = 1e-14 m^3/s at mid-gap), no doping, no ions."""
    L = np.array([30e-9, 300e-9, 30e-9])
    p = dict(BASE, **SCALARS)
    p.update(L=L, mu_n=np.array([1e-6, 1e-4, 1e-6]), mu_p=np.array([1e-6, 1e-4, 1e-6]),
             N_D=np.zeros(3), N_A=np.zeros(3), N_t_bulk=np.array([0.0, 1e21, 0.0]),
             G_ehp=np.array([0.0, 4e27, 0.0]))
    return device_arrays(_structure(L, 256), p)


#: Columns of the drawn table: one row of uniform numbers per device, so that device k is
#: the same whatever the batch size (a prefix of a larger batch is the smaller batch).
COLUMNS = 16


def draw(n: int, seed: int = SEED) -> dict:
    """The parameter table of ``n`` devices: per device 16 uniform numbers mapped to thicknesses
    (uniform), 6 mobilities, 3 dopings, the bulk trap density, the generation rate and the
    ion density (log-uniform)."""
    u = np.random.default_rng(seed).random((n, COLUMNS))

    def lin(col, lo, hi):
        return lo + (hi - lo) * u[:, col]

    def logu(cols, lo, hi):
        return 10.0 ** (np.log10(lo) + (np.log10(hi) - np.log10(lo)) * u[:, cols])

    return {
        "L_etl": lin(0, 20e-9, 60e-9),
        "L_pvk": lin(1, 200e-9, 800e-9),
        "L_htl": lin(2, 20e-9, 60e-9),
        "mu": logu(slice(3, 9), 1e-8, 1e-3),  # (mu_n, mu_p) of ETL, perovskite, HTL
        "dop": logu(slice(9, 12), 1e18, 1e23),  # ETL N_D, perovskite N_A, HTL N_A
        "N_t": logu(12, 1e19, 1e22),
        "G": logu(13, 1e26, 1e28),
        "N_ion": logu(14, 1e22, 1e24),
    }


def table_digest(table: dict) -> str:
    """SHA-256 of a drawn parameter table (the workload's fingerprint)."""
    h = hashlib.sha256()
    for k in sorted(table):
        h.update(k.encode())
        h.update(np.ascontiguousarray(table[k], dtype=np.float64).tobytes())
    return h.hexdigest()


def device_from_row(t: dict, k: int, NP: int = 400, ions: bool = False, frozen: bool = False):
    """Device ``k`` of a drawn table: ions in the perovskite if ``ions``; with ``frozen`` a
    1.2 V pre-bias after which the ions are held."""
    return device_arrays(*row_inputs(t, k, NP, ions, frozen))


def row_inputs(t: dict, k: int, NP: int = 400, ions: bool = False, frozen: bool = False):
    """(structure, numbers) of device ``k`` of a drawn table (`device_from_row`'s inputs)."""
    L = np.array([t["L_etl"][k], t["L_pvk"][k], t["L_htl"][k]])
    mu = t["mu"][k]
    dop = t["dop"][k]
    p = dict(BASE, **SCALARS)
    p.update(
        L=L, mu_n=mu[0::2].copy(), mu_p=mu[1::2].copy(),
        N_D=np.array([dop[0], 0.0, 0.0]), N_A=np.array([0.0, dop[1], dop[2]]),
        N_t_bulk=np.array([0.0, t["N_t"][k], 0.0]), G_ehp=np.array([0.0, t["G"][k], 0.0]),
    )
    if ions:
        N = t["N_ion"][k]
        p.update(N_anion=np.array([0.0, N, 0.0]), N_cation=np.array([0.0, N, 0.0]),
                 mu_anion=np.array([0.0, ION_MOBILITY, 0.0]),
                 mu_cation=np.array([0.0, ION_MOBILITY, 0.0]))
    if frozen:
        p.update(V_pre=1.2)
    return _structure(L, NP, ions=ions, pre_cond=frozen, fix_ions=frozen), p


def devices(config: str, n: int, seed: int = SEED, variant: str = "free", start: int = 0):
    """``n`` devices of workload ``config`` ("c2", "c3", "c4"): rows ``start`` to
    ``start + n`` of the drawn table."""
    t = draw(start + n, seed)
    ions = config == "c3"
    frozen = ions and variant == "frozen"
    return [device_from_row(t, k, 400, ions, frozen) for k in range(start, start + n)]


def _pack_rows(args):
    from .batch import pack

    config, start, stop, seed, variant, NP = args
    t = draw(stop, seed)
    ions = config == "c3"
    frozen = ions and variant == "frozen"
    return pack([device_from_row(t, k, NP, ions, frozen) for k in range(start, stop)])


def packed(config: str, n: int, seed: int = SEED, variant: str = "free", workers: int = 1,
           chunk: int = 256, NP: int = 400) -> dict:
    """The packed kernel arrays (`batch.pack`) of ``n`` devices of a workload (``NP`` interior
    points), built in ``workers`` processes. Call it before the process touches the GPU: the
    workers are forked."""
    from .batch import merge

    jobs = [(config, s, min(s + chunk, n), seed, variant, NP) for s in range(0, n, chunk)]
    if workers <= 1:
        parts = [_pack_rows(j) for j in jobs]
    else:
        import multiprocessing as mp

        pool = mp.get_context("fork").Pool(workers)
        try:
            parts = pool.map(_pack_rows, jobs)
        finally:
            pool.close()  # let the workers exit on their own (no SIGTERM from terminate())
            pool.join()
    return merge(parts)
