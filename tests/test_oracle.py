"""The CPU oracle pinned before it is trusted: against the SimSS fixtures the reference's own
tests hold (tests/golden), closed forms, dense solves and finite differences."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import FIXTURES, GOLDEN, curve_scale, fixture_device, fixture_jv


# ---- Bernoulli (src/jaxdd/solver.py:95-121; reference tests test_solver.py:50-64) ----------
def test_bernoulli_at_zero(oracle):
    B, dB = oracle.bernoulli(0.0)
    assert B[0] == 1.0
    assert dB[0] == pytest.approx(-0.5)


@pytest.mark.parametrize("x", [-30.0, -1.0, -1e-3, 1e-3, 0.3, 2.0, 30.0])
def test_bernoulli_identity(oracle, x):
    """B(-x) = B(x) + x, and B'(-x) = -B'(x) - 1."""
    Bp, dBp = oracle.bernoulli(x)
    Bm, dBm = oracle.bernoulli(-x)
    assert Bm[0] == pytest.approx(Bp[0] + x, rel=1e-12, abs=1e-14)
    assert dBm[0] == pytest.approx(-dBp[0] - 1.0, rel=1e-10, abs=1e-12)


def test_bernoulli_closed_form_and_derivative(oracle):
    """Against x/expm1(x) in extended precision and a central difference of B."""
    x = np.concatenate([np.linspace(-40, 40, 161), [0.02, -0.02, 5e-3, -5e-3]])
    B, dB = oracle.bernoulli(x)
    xl = x.astype(np.longdouble)
    safe = np.where(x == 0, 1, xl)
    exact = np.where(x == 0, 1, safe / np.expm1(safe))
    assert np.max(np.abs(B - exact.astype(float)) / np.abs(exact.astype(float))) < 1e-13
    h = 1e-6
    fd = (oracle.bernoulli(x + h)[0] - oracle.bernoulli(x - h)[0]) / (2 * h)
    assert np.max(np.abs(dB - fd)) < 1e-8


def test_simss_bernoulli_within_its_own_error(oracle):
    x = np.linspace(-5, 5, 2001)
    assert np.max(np.abs(oracle.bernoulli(x, "simss")[0] / oracle.bernoulli(x)[0] - 1)) < 1e-3


def test_host_bernoulli_matches_oracle(oracle):
    from jaxdd_amd.solver import bernoulli, bernoulli_simss

    x = np.linspace(-20, 20, 801)
    assert np.allclose(bernoulli(x), oracle.bernoulli(x)[0], rtol=1e-14, atol=0)
    assert np.allclose(bernoulli_simss(x), oracle.bernoulli(x, "simss")[0], rtol=1e-14, atol=0)


# ---- block Thomas vs a dense solve (reference tests test_solver.py:1753-1773) --------------
@pytest.mark.parametrize("b,k", [(3, 1), (3, 3), (5, 4)])
def test_block_thomas_is_the_dense_solve(oracle, b, k):
    rng = np.random.default_rng(3)
    NP = 12
    A, C = rng.normal(size=(NP, b, b)), rng.normal(size=(NP, b, b))
    B = rng.normal(size=(NP, b, b)) + 8 * np.eye(b)
    d = rng.normal(size=(NP, b, k))
    T = np.zeros((NP * b, NP * b))
    for i in range(NP):
        T[i * b:(i + 1) * b, i * b:(i + 1) * b] = B[i]
        if i > 0:
            T[i * b:(i + 1) * b, (i - 1) * b:i * b] = A[i]
        if i < NP - 1:
            T[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = C[i]
    ref = np.linalg.solve(T, d.reshape(NP * b, k)).reshape(NP, b, k)
    x = oracle.block_thomas(A, B, C, d)
    assert np.max(np.abs(x - ref)) < 1e-12 * np.max(np.abs(ref))


# ---- assembly: Jacobian by finite differences, residual by rows ---------------------------
def _perturbed_state(dev, seed=1):
    from tests_helpers import initial_full

    rng = np.random.default_rng(seed)
    full = initial_full(dev, 0.3)
    NP = len(dev.x) - 2
    noise = np.zeros_like(full)
    noise[1:-1] = rng.normal(0, 1, (NP, 3)) * np.array([5 * dev.Vt, 2.0, 2.0])
    noise[1:-1:7, 0] = 0.0  # some neighbours at equal potential: the series of B
    return full + noise


@pytest.mark.parametrize("case", ["thin_l1_light", "steady_1e17_thin_light",
                                  "bulk_acceptor_light"])
@pytest.mark.parametrize("bern", ["exact", "simss"])
def test_assembly_jacobian_is_the_residuals_derivative(oracle, case, bern):
    """A, B, C and the ion columns U against central differences of the residual F
    (the reference checks its by-edge assembly against jacfwd of its rows,
    test_solver.py:1578-1617)."""
    dev = fixture_device(case)
    full = _perturbed_state(dev)
    NP = len(dev.x) - 2
    la, lc = 40.0, 41.0
    F, A, B, C, U = oracle.assemble(dev, full, la=la, lc=lc, bernoulli_name=bern)
    rng = np.random.default_rng(7)
    for i in rng.choice(np.arange(1, NP + 1), 12, replace=False):
        for col in range(3):
            h = 1e-7 * max(1.0, abs(full[i, col]))
            fp, fm = full.copy(), full.copy()
            fp[i, col] += h
            fm[i, col] -= h
            dF = (oracle.assemble(dev, fp, la=la, lc=lc, bernoulli_name=bern)[0]
                  - oracle.assemble(dev, fm, la=la, lc=lc, bernoulli_name=bern)[0]) / (2 * h)
            r = i - 1
            for row_pt, blk in ((r, B), (r + 1, A), (r - 1, C)):
                if 0 <= row_pt < NP:
                    got, want = blk[row_pt, :, col], dF[row_pt]
                    scale = np.max(np.abs(blk[row_pt])) + 1e-300
                    # the series switch of SimSS's B makes its derivative piecewise
                    tol = 1e-5 if bern == "exact" else 2e-3
                    assert np.max(np.abs(got - want)) < tol * scale, (case, i, col)
    for k, l0 in ((0, la), (1, lc)):
        h = 1e-6
        args = [la, lc]
        args[k] = l0 + h
        Fp = oracle.assemble(dev, full, la=args[0], lc=args[1], bernoulli_name=bern)[0]
        args[k] = l0 - h
        Fm = oracle.assemble(dev, full, la=args[0], lc=args[1], bernoulli_name=bern)[0]
        dF = (Fp - Fm) / (2 * h)
        mask = (dev.anion_mask if k == 0 else dev.cation_mask)[1:-1]
        got = U[:, :, k] * mask[:, None]
        assert np.max(np.abs(got - dF * mask[:, None])) < 1e-6 * (np.max(np.abs(dF)) + 1e-300)


def test_assembly_residual_by_rows(oracle):
    """The residual by edges equals an independent row-by-row NumPy statement of the discrete
    equations (src/jaxdd/solver.py:964-1006, `_row`)."""
    from tests_helpers import residual_by_rows

    dev = fixture_device("steady_1e17_thin_light")
    full = _perturbed_state(dev, seed=5)
    F = oracle.assemble(dev, full, la=40.0, lc=41.0)[0]
    ref = residual_by_rows(dev, full, 40.0, 41.0)
    scale = np.max(np.abs(ref), axis=0)
    assert np.max(np.abs(F - ref) / scale) < 1e-11


# ---- end to end against SimSS -------------------------------------------------------------
@pytest.mark.parametrize("case", sorted(FIXTURES))
def test_oracle_solve_jv_reproduces_simss(oracle, case):
    dev = fixture_device(case)
    ref = fixture_jv(case)
    J, ok, its = oracle.solve_jv(dev, ref["Vint"], bernoulli_name="simss")
    assert ok.all()
    scale = np.max(np.abs(ref["Jext"]))
    assert np.max(np.abs(J - ref["Jext"])) < FIXTURES[case] * scale


@pytest.mark.parametrize("case", sorted(FIXTURES))
def test_oracle_batch_reproduces_simss(oracle, case):
    dev = fixture_device(case)
    ref = fixture_jv(case)
    J, ok, its = oracle.solve_batch([dev], ref["Vint"], bernoulli_name="simss")
    assert ok.all()
    scale = np.max(np.abs(ref["Jext"]))
    assert np.max(np.abs(J[0] - ref["Jext"])) < FIXTURES[case] * scale


def test_charged_traps_move_the_curve_as_simss(oracle):
    """Acceptor bulk traps charged in the absorber (reference: solver.py:762-773, 1136-1140; the
    start from the solution with them neutral, batch.py:898-900): the oracle's curve is SimSS's
    to 1e-9 of its scale, in at most 8 steps a voltage, where the same traps left neutral miss it
    by 4 % (their charge is what moves it); the transient's restatement refuses them."""
    import dataclasses

    from conftest import fixture_device, fixture_jv

    dev, ref = fixture_device("bulk_acceptor_light"), fixture_jv("bulk_acceptor_light")
    assert set(np.unique(dev.srh_type)) == {-1.0, 0.0}
    scale = np.max(np.abs(ref["Jext"]))
    J, ok, its = oracle.solve_jv(dev, ref["Vint"], bernoulli_name="simss")
    assert ok.all() and its.max() <= 8
    assert np.max(np.abs(J - ref["Jext"])) < 1e-9 * scale
    J0, ok0, _ = oracle.solve_jv(dataclasses.replace(dev, srh_type=None), ref["Vint"],
                                 bernoulli_name="simss")
    assert ok0.all() and np.max(np.abs(J0 - ref["Jext"])) > 0.03 * scale
    with pytest.raises(NotImplementedError, match="charged"):
        oracle.transient(dev, np.array([0.0, 1e-9]), np.array([0.0, 0.1]), np.ones(2))


def test_charged_traps_are_refused_by_the_impedance_and_the_transient():
    """The impedance and the transient take no charged traps (`NotImplementedError` before any
    launch); the GPU builder takes them (their type a held per-layer number)."""
    from conftest import GOLDEN
    from jaxdd_amd import builder as B
    from jaxdd_amd.batch import charged_traps, refuse_charged_traps
    from jaxdd_amd.device import build_device, device_parameters
    from jaxdd_amd.simss_input import check_supported, read_setup

    setup = read_setup(GOLDEN / "bulk_acceptor_light" / "simulation_setup.txt")
    check_supported(setup)
    st, p = device_parameters(setup, grid="simss")
    B.check_buildable(st)
    layers, _ = B.number_arrays([st], [p])
    assert list(layers[:, B.LAYER_FIELDS.index("bulkTrapType"), 0]) == [0.0, -1.0, 0.0]
    dev = build_device(setup, grid="simss")
    assert charged_traps([dev]) and not charged_traps([fixture_device("thin_l1_light")])
    for what in ("the impedance", "the transient"):
        with pytest.raises(NotImplementedError, match=f"{what}: charged"):
            refuse_charged_traps([dev], what)


def test_open_circuit_voltage_is_simss(oracle):
    """SimSS's run at 0.1 mV steps around open circuit: the oracle's curve to 2e-11 A/m^2
    (Jphoto 257 A/m^2), and its Voc (the cubic through the six points) SimSS's to 1e-10 V,
    the reference's bar for its own open circuit."""
    from conftest import fixture_device, fixture_jv, open_circuit_voltage

    dev, ref = fixture_device("thin_l1_voc_light"), fixture_jv("thin_l1_voc_light")
    J, ok, _ = oracle.solve_jv(dev, ref["Vint"], bernoulli_name="simss")
    assert ok.all()
    assert np.max(np.abs(J - ref["Jext"])) < 5e-11
    assert abs(open_circuit_voltage(ref["Vext"], J)
               - open_circuit_voltage(ref["Vext"], ref["Jext"])) < 1e-10


def test_langevin_recombination_reproduces_simss(oracle):
    """SimSS's Langevin recombination in all three layers (`useLangevin`, `preLangevin` 1e-3,
    1e-2, 1e-3; reference tests/test_solver.py:1083-1101): at SIMsalabim's mobilities (those of
    the interval after each point, `langevin="simss"`) within 1e-9 of the curve's scale; at
    each point's own layer's (the default) 1e-5 to 1e-4 off, the points before an interface
    taking the layer's mobility where SimSS takes the interface's. The rate moves the curve by
    per cent: without it the curve is far off."""
    import dataclasses

    from jaxdd_amd.device import build_device, device_arrays, device_parameters
    from jaxdd_amd.simss_input import check_supported, read_setup

    setup = read_setup(GOLDEN / "langevin_light" / "simulation_setup.txt")
    check_supported(setup)
    ref = fixture_jv("langevin_light")
    scale = np.max(np.abs(ref["Jext"]))
    gap = {}
    for mode in ("simss", "layer"):
        J, ok, _ = oracle.solve_jv(build_device(setup, grid="simss", langevin=mode), ref["Vint"],
                                   bernoulli_name="simss")
        assert ok.all()
        gap[mode] = np.max(np.abs(J - ref["Jext"])) / scale
    assert gap["simss"] < 1e-9
    assert 1e-5 < gap["layer"] < 1e-4
    st, p = device_parameters(setup, grid="simss")
    off = device_arrays(dataclasses.replace(st, langevin=np.zeros(3, bool)), p)
    J, _, _ = oracle.solve_jv(off, ref["Vint"], bernoulli_name="simss")
    assert np.max(np.abs(J - ref["Jext"])) / scale > 1e-3


def test_dark_current_is_monotone_and_zero_at_zero(oracle):
    """frozen_thick_htl_dark (reference test_solver.py:1556-1575): monotone, |J(0)| below the
    current tolerance of the scale."""
    dev = fixture_device("frozen_thick_htl_dark")
    ref = fixture_jv("frozen_thick_htl_dark")
    J, ok, _ = oracle.solve_batch([dev], ref["Vint"], bernoulli_name="simss")
    J = J[0]
    scale = oracle.j_scale(dev)
    assert ok.all()
    assert np.all(np.diff(J) > -1e-5 * scale)
    assert abs(J[np.argmin(np.abs(ref["Vint"]))]) < 1e-5 * scale


def test_series_resistance_external_voltage(oracle):
    dev = fixture_device("resist_light")
    ref = fixture_jv("resist_light")
    J, ok, _ = oracle.solve_jv(dev, ref["Vint"], bernoulli_name="simss")
    V_ext = ref["Vint"] + J * dev.R_series
    assert np.max(np.abs(V_ext - ref["Vext"])) < 1e-4


def test_short_circuit_current_below_generation(oracle):
    from jaxdd_amd.constants import Q

    dev = fixture_device("thin_l1_light")
    J, ok, _ = oracle.solve_jv(dev, np.array([0.0]))
    generated = Q * np.sum(dev.h[:-1] * dev.G[:-1]) / np.sum(dev.h[:-1]) * dev.L
    assert ok[0] and -J[0] <= generated * (1 + 1e-9)


# ---- batch semantics (reference tests test_solver.py:1479-1532) ----------------------------
BATCH_CASES = ["thin_l1_light", "frozen_1e17_x1_thin_light", "steady_1e17_thin_light"]
V_COARSE = np.round(np.arange(-0.2, 1.3 + 1e-9, 0.1), 6)


def test_batch_matches_one_device(oracle):
    devs = [fixture_device(c) for c in BATCH_CASES]
    J, ok, _ = oracle.solve_batch(devs, V_COARSE, bernoulli_name="simss")
    assert ok.all()
    for k, dev in enumerate(devs):
        single, ok1, _ = oracle.solve_jv(dev, V_COARSE, bernoulli_name="simss")
        assert ok1.all()
        assert np.max(np.abs(J[k] - single)) < 1e-6 * np.max(np.abs(single))


def test_two_passes_give_the_one_pass_result(oracle):
    devs = [fixture_device(c) for c in BATCH_CASES]
    devs = devs + devs[:2]
    one = oracle.solve_batch(devs, V_COARSE, first_max_iter=None)
    two = oracle.solve_batch(devs, V_COARSE, first_max_iter=6)
    first = oracle.solve_batch(devs, V_COARSE, first_max_iter=6, max_iter=6)
    assert 0 < (~first[1].all(axis=1)).sum() < len(devs)  # some devices are re-solved
    assert np.array_equal(one[1], two[1]) and np.array_equal(one[2], two[2])
    assert np.array_equal(one[0], two[0])


def test_voltage_order_does_not_matter(oracle):
    devs = [fixture_device("frozen_1e17_x1_thin_light")]
    up = oracle.solve_batch(devs, V_COARSE)
    down = oracle.solve_batch(devs, V_COARSE[::-1])
    for a, b in zip(up, down):
        assert np.array_equal(a, b[:, ::-1])


# ---- the 5x5 steady system with the ions as unknowns (solver.py:1364-1473) ---------------
def _ion_state(dev, full3, la, lc, rng, noise):
    """(NP + 2) x 5: the carriers' state and each ion's log density, Boltzmann's of the slot's
    normalisation in its region (plus noise), 0 outside it."""
    V = full3[:, 0]
    am, cm = np.asarray(dev.anion_mask) > 0, np.asarray(dev.cation_mask) > 0
    lna = np.where(am, la + V / dev.Vt + rng.normal(0, noise, V.shape), 0.0)
    lnc = np.where(cm, lc - V / dev.Vt + rng.normal(0, noise, V.shape), 0.0)
    return np.column_stack([full3, lna, lnc])


@pytest.mark.parametrize("case", ["steady_1e17_thin_light", "frozen_1e17_x1_thin_light"])
def test_ion_system_reduces_to_the_3x3_at_boltzmann_ions(oracle, case):
    """Ions in Boltzmann's profiles of the slots' normalisations: the 5x5 carriers' rows are the
    3x3 assembly's with those normalisations, the ions' rows vanish in their regions and hold
    the unknown outside them, and Poisson's V column with its ion columns through
    d ln a/dV = 1/V_T, d ln c/dV = -1/V_T is the 3x3's V column (the chain rule)."""
    dev = fixture_device(case)
    full = _perturbed_state(dev)
    la, lc = 40.0, 41.0
    full5 = _ion_state(dev, full, la, lc, np.random.default_rng(3), 0.0)
    F5, A5, B5, C5, U5 = oracle.assemble_ions(dev, full5, [la, lc])
    F, A, B, C, U = oracle.assemble(dev, full, la=la, lc=lc)
    rs = np.max(np.abs(np.concatenate([A, B, C], axis=2)), axis=2)  # each row's scale
    assert np.all(np.abs(F5[:, :3] - F) <= 1e-12 * (rs + np.abs(F)))
    for got, ref in ((A5, A), (C5, C)):
        assert np.all(np.abs(got[:, :3, :3] - ref) <= 1e-12 * rs[:, :, None])
        assert not got[:, 3:].any() and not got[:, :, 3:].any()
    assert np.all(np.abs(B5[:, 1:3, :3] - B[:, 1:3]) <= 1e-12 * rs[:, 1:3, None])
    chain = B5[:, 0, 0] + (B5[:, 0, 3] - B5[:, 0, 4]) / dev.Vt
    assert np.all(np.abs(chain - B[:, 0, 0]) <= 1e-12 * rs[:, 0])
    assert np.all(np.abs(B5[:, 0, 1:3] - B[:, 0, 1:3]) <= 1e-12 * rs[:, 0, None])
    am, cm = np.asarray(dev.anion_mask)[1:-1] > 0, np.asarray(dev.cation_mask)[1:-1] > 0
    assert am.any() and cm.any() and not am.all()
    assert np.all(np.abs(F5[am, 3]) < 1e-12 * (la + np.abs(full5[1:-1][am, 3])))
    assert np.all(np.abs(F5[cm, 4]) < 1e-12 * (lc + np.abs(full5[1:-1][cm, 4])))
    assert np.array_equal(F5[~am, 3], full5[1:-1][~am, 3])
    assert np.array_equal(F5[~cm, 4], full5[1:-1][~cm, 4])


@pytest.mark.parametrize("bern", ["exact", "simss"])
def test_ion_system_jacobian_is_the_residuals_derivative(oracle, bern):
    """The 5x5 blocks and the slots' columns against central differences of the 5x5 residual,
    at ions off their Boltzmann profiles (test_solver.py:2077-2112 checks the reference's
    blocks against its autodiff)."""
    dev = fixture_device("steady_1e17_thin_light")
    la, lc = 40.0, 41.0
    full5 = _ion_state(dev, _perturbed_state(dev), la, lc, np.random.default_rng(4), 0.4)
    NP = len(dev.x) - 2
    F5, A5, B5, C5, U5 = oracle.assemble_ions(dev, full5, [la, lc], bernoulli_name=bern)
    rng = np.random.default_rng(8)
    for i in rng.choice(np.arange(1, NP + 1), 10, replace=False):
        for col in range(5):
            h = 1e-7 * max(1.0, abs(full5[i, col]))
            fp, fm = full5.copy(), full5.copy()
            fp[i, col] += h
            fm[i, col] -= h
            dF = (oracle.assemble_ions(dev, fp, [la, lc], bernoulli_name=bern)[0]
                  - oracle.assemble_ions(dev, fm, [la, lc], bernoulli_name=bern)[0]) / (2 * h)
            r = i - 1
            for row_pt, blk in ((r, B5), (r + 1, A5), (r - 1, C5)):
                if 0 <= row_pt < NP:
                    got, want = blk[row_pt, :, col], dF[row_pt]
                    scale = np.max(np.abs(blk[row_pt])) + 1e-300  # as the 3x3 test's
                    tol = 1e-5 if bern == "exact" else 2e-3
                    assert np.max(np.abs(got - want)) < tol * scale, (i, col)
    for k in range(2):
        h = 1e-6
        lp, lm = [la, lc], [la, lc]
        lp[k] += h
        lm[k] -= h
        dF = (oracle.assemble_ions(dev, full5, lp, bernoulli_name=bern)[0]
              - oracle.assemble_ions(dev, full5, lm, bernoulli_name=bern)[0]) / (2 * h)
        assert np.max(np.abs(U5[:, :, k] - dF)) < 1e-8


# ---- the noise floor of solve_jv's stop (solver.py:1728-1751) -------------------------------
def _fd_blocks(oracle, dev, full, bern, rel=1e-6):
    """A, B, C (NP x 3 x 3) of the residual by central differences of the oracle's F in every
    value of the full state (the contacts' included)."""
    NP = len(dev.x) - 2
    A, B, C = (np.zeros((NP, 3, 3)) for _ in range(3))
    for j in range(NP + 2):
        for c in range(3):
            h = rel * max(1.0, abs(full[j, c]))
            up, dn = full.copy(), full.copy()
            up[j, c] += h
            dn[j, c] -= h
            dF = (oracle.assemble(dev, up, bernoulli_name=bern)[0]
                  - oracle.assemble(dev, dn, bernoulli_name=bern)[0]) / (2 * h)
            for r in range(NP):
                off = j - (r + 1)  # point j relative to row r's own point r + 1
                if off == -1:
                    A[r, :, c] = dF[r]
                elif off == 0:
                    B[r, :, c] = dF[r]
                elif off == 1:
                    C[r, :, c] = dF[r]
    return A, B, C


@pytest.mark.parametrize("bern", ["exact", "simss"])
def test_noise_step_is_the_round_off_step_of_the_dense_system(oracle, bern):
    """The oracle's `_noise_step`, rebuilt independently: the round-off right-hand side
    eps * sum |dF/dy| |y| from a central-difference Jacobian of the residual, the system solved
    densely (numpy), the step weighted as the Newton step is (V over V_T; the densities'
    logarithms by n / (n + n_floor)). Also exactly linear in eps."""
    from jaxdd_amd import synthetic as S

    t = S.draw(8)
    dev = S.device_from_row(t, 5, 20, False, False)
    NP = 20
    V = 0.8
    _, ok, _, st = oracle.solve_one(dev, [V], bernoulli=bern, return_states=True)
    assert ok.all()
    rng = np.random.default_rng(2)
    for y in (st[0], st[0] + 1e-3 * rng.standard_normal(st[0].shape)):
        full = np.zeros((NP + 2, 3))
        full[1:-1] = y
        full[0] = [dev.V0 + 0.5 * dev.left_elec * V, np.log(dev.n_left), np.log(dev.p_left)]
        full[-1] = [dev.VL - 0.5 * dev.left_elec * V, np.log(dev.n_right), np.log(dev.p_right)]
        Af, Bf, Cf = _fd_blocks(oracle, dev, full, bern)
        _, A, B, C, _ = oracle.assemble(dev, full, bernoulli_name=bern)
        scale = np.max(np.abs(np.concatenate([A, B, C], axis=2)), axis=2)[:, :, None]
        for fd, an in ((Af, A), (Bf, B), (Cf, C)):
            assert np.all(np.abs(fd - an) <= 1e-6 * scale)
        a = np.abs(full)
        eps = np.finfo(float).eps
        rhs = eps * (np.einsum("irc,ic->ir", np.abs(Af), a[:-2])
                     + np.einsum("irc,ic->ir", np.abs(Bf), a[1:-1])
                     + np.einsum("irc,ic->ir", np.abs(Cf), a[2:]))
        T = np.zeros((3 * NP, 3 * NP))
        for i in range(NP):
            T[3 * i:3 * i + 3, 3 * i:3 * i + 3] = B[i]
            if i > 0:
                T[3 * i:3 * i + 3, 3 * i - 3:3 * i] = A[i]
            if i < NP - 1:
                T[3 * i:3 * i + 3, 3 * i + 3:3 * i + 6] = C[i]
        X = np.linalg.solve(T, rhs.ravel()).reshape(NP, 3)
        dens = np.exp(y[:, 1:])
        n_floor = 1e-12 * np.max(dev.N_c)
        want = max(np.max(np.abs(X[:, 0])) / dev.Vt,
                   np.max(np.abs(X[:, 1:]) * dens / (dens + n_floor)))
        got = oracle.noise_step(dev, y, V, bernoulli_name=bern)
        assert got > 0 and abs(got / want - 1) < 1e-6, (got, want)
        assert oracle.noise_step(dev, y, V, bernoulli_name=bern, eps=2 * eps) == 2 * got
        assert abs(oracle.noise_step(dev, y, V, bernoulli_name=bern, eps=3 * eps) / (3 * got)
                   - 1) < 1e-13
