"""`solve_states` (src/jaxdd/batch.py:1165-1191): the oracle's restatement pinned on the CPU
by what the mathematics fixes, then the HIP path (ddh_solve_batch_states through the C-ABI)
against it on the GPU."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import CHAOTIC_STEPS, chaotic_devices, fixture_device


def _trapezoid_number(dev, dens, mask):
    m = np.asarray(mask, dtype=float)
    return np.sum(0.5 * (dens[1:] + dens[:-1]) * m[1:] * m[:-1] * dev.h[:-1] * dev.L)


# ---- the oracle's restatement, pinned ------------------------------------------------------
def test_oracle_states_keep_every_ion_region_at_its_total(oracle):
    """Boltzmann's normalisation puts each species' total into its region at every voltage
    (the trapezoidal number of solver.py:1241-1252), and the states are the solve's own."""
    from jaxdd_amd import synthetic as S

    dev = S.devices("c3", 1, variant="free", start=3)[0]
    V = np.array([0.9, -0.1, 0.4])
    out = oracle.solve_states(dev, V)
    J, ok, _ = oracle.solve_batch([dev], V, first_max_iter=None)
    assert np.array_equal(out["converged"], ok[0]) and np.allclose(out["J"], J[0], rtol=0, atol=0)
    for k in range(len(V)):
        na = _trapezoid_number(dev, out["anion"][k] - dev.anion_fixed, dev.anion_mask)
        nc = _trapezoid_number(dev, out["cation"][k] - dev.cation_fixed, dev.cation_mask)
        assert abs(na / dev.anion_total - 1) < 1e-12 and abs(nc / dev.cation_total - 1) < 1e-12
        # anions pile up where V is high: Boltzmann's sign
        inside = np.flatnonzero(np.asarray(dev.anion_mask))
        i, j = inside[np.argmin(out["V"][k][inside])], inside[np.argmax(out["V"][k][inside])]
        assert out["anion"][k][j] >= out["anion"][k][i]
        assert out["cation"][k][j] <= out["cation"][k][i]


def test_oracle_observables_of_a_uniform_splitting(oracle):
    """A state whose quasi-Fermi levels split by the same D on every point has qfls = D there
    and an implied V_oc of exactly D; at D = 0 the photons balance their equilibrium (plqy 0)."""
    from jaxdd_amd import synthetic as S

    dev = S.devices("c2", 1, start=5)[0]
    rng = np.random.default_rng(3)
    P = len(dev.x)
    for D in (0.0, 0.37, 1.05):
        full = np.zeros((1, P, 3))
        full[0, :, 0] = rng.uniform(-1, 1, P)
        full[0, :, 1] = rng.uniform(-30, 30, P) + np.log(dev.ni)
        full[0, :, 2] = 2 * np.log(dev.ni) + D / dev.Vt - full[0, :, 1]
        qfls, voc, plqy = oracle.observed(dev, full, 1.0)
        assert np.max(np.abs(qfls - D)) < 1e-12 and abs(voc[0] - D) < 1e-12
        if D == 0.0:
            assert abs(plqy[0]) < 1e-12
        else:
            assert plqy[0] > 0


def test_oracle_dark_equilibrium_has_no_splitting(oracle):
    """In the dark at 0 V the device is in equilibrium: no quasi-Fermi level splitting, an
    implied V_oc of 0, and no generation to measure a PLQY against."""
    from jaxdd_amd import synthetic as S

    dev = dataclasses.replace(S.devices("c2", 1, start=2)[0], G_frac=0.0)
    out = oracle.solve_states(dev, [0.0])
    assert out["converged"].all()
    assert np.max(np.abs(out["qfls"])) < 1e-8 and abs(out["implied_voc"][0]) < 1e-8
    assert np.isnan(out["plqy"][0]) and abs(out["J"][0]) < 1e-6


def test_oracle_frozen_states_hold_the_pre_bias_profile(oracle):
    """With ions frozen after the pre-bias every voltage carries the pre-bias's profile (the
    freeze, batch.py:859-872), which keeps each species' total."""
    from jaxdd_amd import synthetic as S

    dev = S.devices("c3", 1, variant="frozen", start=1)[0]
    out = oracle.solve_states(dev, [0.0, 1.0])
    assert np.array_equal(out["anion"][0], out["anion"][1])
    free = dataclasses.replace(dev, pre_cond=False, fix_ions=False)
    pre = oracle.solve_states(free, [dev.V_pre])
    assert np.array_equal(out["anion"][0], pre["anion"][0])
    na = _trapezoid_number(dev, out["anion"][0] - dev.anion_fixed, dev.anion_mask)
    assert abs(na / dev.anion_total - 1) < 1e-12


# ---- the HIP path against the oracle -----------------------------------------------------
def _compare(got, ref, k, dev, *, v_tol=1e-8, rel=1e-7):
    calm = ref["converged"] & (ref["iterations"] < CHAOTIC_STEPS) & (got["iterations"][k] < CHAOTIC_STEPS)
    assert np.array_equal(got["converged"][k][calm], ref["converged"][calm])
    for key in ("V", "qfls"):
        assert np.max(np.abs(got[key][k][calm] - ref[key][calm])) < v_tol, key
    assert np.max(np.abs(got["implied_voc"][k][calm] - ref["implied_voc"][calm])) < v_tol
    for key in ("n", "p", "anion", "cation"):
        g, r = got[key][k][calm], ref[key][calm]
        assert np.all(np.abs(g - r) <= rel * np.abs(r) + 1e-300), key
    g, r = got["plqy"][k][calm], ref["plqy"][calm]
    assert np.array_equal(np.isnan(g), np.isnan(r))
    fin = ~np.isnan(r)
    assert np.all(np.abs(g[fin] - r[fin]) <= rel * np.abs(r[fin]) + 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 2])
def test_solve_states_matches_oracle_on_fixtures(torch_cuda, oracle, lanes):
    from jaxdd_amd.batch import prepare, solve_states

    cases = ["thin_l1_light", "steady_1e17_thin_light", "frozen_1e17_x1_thin_light"]
    devs = [fixture_device(c) for c in cases]
    V = np.array([0.6, -0.2, 0.0, 1.1, 0.95])
    print(f"solve_states fixtures lanes={lanes}", flush=True)
    got = {k: v.cpu().numpy() for k, v in solve_states(prepare(devs, "cuda"), V,
                                                         bernoulli_name="simss", lanes=lanes).items()}
    for k, dev in enumerate(devs):
        ref = oracle.solve_states(dev, V, bernoulli_name="simss")
        assert np.array_equal(got["iterations"][k], ref["iterations"]), cases[k]
        _compare(got, ref, k, dev)


@pytest.mark.gpu
@pytest.mark.parametrize("workload", ["c2", "c3-free", "c3-frozen"])
def test_solve_states_matches_oracle_on_synthetic(torch_cuda, oracle, workload):
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import prepare, solve_states

    config, _, variant = workload.partition("-")
    devs = S.devices(config, 8, variant=variant or "free", start=40)
    V = S.VOLTAGES[::6].copy()
    print(f"solve_states {workload}: 8 devices x {len(V)} voltages", flush=True)
    got = {k: v.cpu().numpy() for k, v in solve_states(prepare(devs, "cuda"), V).items()}
    refs = [oracle.solve_states(d, V) for d in devs]
    skip = np.zeros(len(devs), bool)
    if variant == "frozen":
        skip = chaotic_devices(oracle, devs, V, np.stack([r["J"] for r in refs]))
        print(f"  devices whose pre-bias lands by the last bits (not compared): "
              f"{np.flatnonzero(skip).tolist()}", flush=True)
        assert skip.sum() <= 1
    for k, dev in enumerate(devs):
        if not skip[k]:
            _compare(got, refs[k], k, dev)
